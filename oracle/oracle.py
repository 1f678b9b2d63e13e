"""ctypes wrapper of the CPU restatement in oracle/gpm_oracle.c.

TEST INFRASTRUCTURE: only tests/, __graft_entry__.smoke() and bench.py's CPU
baseline leg use this module, as the checker.  It never backs the product path.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
# the reference's ids (constants.f90:17-34)
BC_CODES = {"neumann": 17, "periodic": 18, "dirichlet": 19, "reflective": 20, "kh": 403, "mach800": 201,
            "doublemach800": 301, "rmi": 401, "dmr": 35, "rt": 402}
FLUX_CODES = {"hllc": 100, "llf": 101, "hll": 102, "unsplit": 103}
INTEGRATORS = {"fe": 8, "ssp_rk2": 9, "ssp_rk3": 10, "ssp_rk4": 12}
PROBLEM_RT = 51
METHODS = {"gp_mood": 2, "fog": 1, "pol_mood": 200, "unlim_pol": 201, "gp_se": 202, "nn_gp_mood": 203,
           "nn_gp_mood_cc": 204}
CASCADES = {"reference": 0, "full": 1}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle library not built: run `make -C oracle`")
        L = ctypes.CDLL(LIB_PATH)
        P, I, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        L.orc_create.restype = P
        L.orc_create.argtypes = [I] * 9 + [D] * 3 + [P] * 5
        L.orc_destroy.argtypes = [P]
        L.orc_set_flux.argtypes = [P, I]
        L.orc_set_integrator.argtypes = [P, I]
        L.orc_set_problem.argtypes = [P, I]
        L.orc_set_method.argtypes = [P, I]
        L.orc_set_t_rk.argtypes = [P, D]
        L.orc_riemann.argtypes = [I, P, P, I, P]
        L.orc_bc.argtypes = [P, P]
        L.orc_setdt.restype = D
        L.orc_setdt.argtypes = [P, P, D]
        L.orc_step.argtypes = [P, P, P, D, D]
        L.orc_run.argtypes = [P, P, I, P, D, P, P, P, P]
        L.orc_step_stages.argtypes = [P, P, D, D, P, P]
        L.orc_count_detected.restype = ctypes.c_long
        L.orc_count_detected.argtypes = [P]
        L.orc_passes.argtypes = [P, P]
        L.orc_hllc.argtypes = [P, P, I, P]
        L.orc_set_initial_orders.argtypes = [P, P]
        L.orc_run2.argtypes = [P, P, I, P, D, P, P, P, P, P, P, P]
        L.orc_set_refnn.argtypes = [P, I, I, P, P, P, P, P, P]
        L.orc_set_niter.argtypes = [P, I]
        L.orc_skip_initial_bc.argtypes = [P]
        L.orc_get_niter.argtypes = [P]
        L.orc_error.argtypes = [P]
        L.orc_nn_forward.argtypes = [P, P, P]
        L.orc_expf_bits.argtypes = [ctypes.c_uint32, ctypes.c_long, P]
        L.orc_pol_table.argtypes = [P, P]
        L.orc_nn_format_input.argtypes = [P, P, I, I, I, P]
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


class Oracle:
    """The reference step on CPU.  ``tables`` is a dict with keys zT, d2x, d2y,
    gw, ixiy in the reference's Fortran layouts (flattened column-major).

    The grid is ``n`` (lf) cells in x by ``nf`` (default ``n``) in y on
    [0, lx] x [0, ly]; arrays are indexed [row (y), column (x), variable]."""

    def __init__(self, n, mord, ngp, bc, tables, dmp=True, u2=True, u2tol=True,
                 cascade="reference", cfl=0.8, flux="hllc", nf=None, lx=1.0, ly=1.0,
                 integrator="ssp_rk3", problem=0, dx=None, dy=None, method="gp_mood"):
        self.n = self.lf = n
        self.nf = n if nf is None else nf
        self.tables = {k: np.ascontiguousarray(v) for k, v in tables.items()}
        t = self.tables
        dx = lx / n if dx is None else dx
        dy = ly / self.nf if dy is None else dy
        self._h = lib().orc_create(n, self.nf, mord, ngp, BC_CODES[bc], int(dmp), int(u2), int(u2tol),
                                   CASCADES[cascade], cfl, dx, dy,
                                   _ptr(t["zT"].astype(np.float64)), _ptr(t["d2x"].astype(np.float64)),
                                   _ptr(t["d2y"].astype(np.float64)), _ptr(t["gw"].astype(np.float64)),
                                   _ptr(t["ixiy"].astype(np.int32)))
        self._keep = [t[k] for k in t]
        lib().orc_set_flux(self._h, FLUX_CODES[flux])
        lib().orc_set_integrator(self._h, INTEGRATORS[integrator])
        lib().orc_set_problem(self._h, int(problem))
        lib().orc_set_method(self._h, METHODS[method])

    def __del__(self):
        # at interpreter exit the module's globals may already be gone
        if getattr(self, "_h", None) and callable(globals().get("lib")):
            lib().orc_destroy(self._h)
            self._h = None

    def full(self, interior):
        """Ghosted array from the interior, ghosts by Boundary_C at the current t_rk."""
        U = np.zeros((self.nf + 8, self.lf + 8, 4), dtype=np.float64)
        U[4:self.nf + 4, 4:self.lf + 4] = interior
        lib().orc_bc(self._h, _ptr(U))
        return U

    def set_t_rk(self, t_rk):
        lib().orc_set_t_rk(self._h, float(t_rk))

    def run(self, U, nsteps, t=0.0, tmax=1e10, masks=False, initial_bc=True):
        """Advance full array U (modified in place) by nsteps steps.  With
        initial_bc=False the first Setdt reads U's ghost cells as given (the
        reference's restart: main.f90:124-151) instead of Boundary_C's."""
        assert U.flags.c_contiguous and U.dtype == np.float64
        if not initial_bc:
            lib().orc_skip_initial_bc(self._h)
        dts = np.zeros(nsteps)
        counts = np.zeros(nsteps, dtype=np.int64)
        passes = np.zeros((nsteps, 3), dtype=np.int32)
        m = np.zeros((nsteps, self.nf, self.lf), dtype=np.int8) if masks else None
        tt = ctypes.c_double(t)
        lib().orc_run(self._h, _ptr(U), nsteps, ctypes.byref(tt), tmax, _ptr(dts), _ptr(counts),
                      _ptr(m), _ptr(passes))
        return dict(dt=dts, counts=counts, masks=m, passes=passes, t=tt.value)

    def set_reference_nn(self, weight0, bias0, weight1, bias1, weight2, bias2):
        """The reference network, matrices with row i = output i (weight0
        (length, L), ..., weight2 (2, length)); stored in Fortran layouts."""
        w0 = np.asarray(weight0, dtype=np.float32)
        length, L = w0.shape
        self._nn = [np.ascontiguousarray(np.asarray(a, dtype=np.float32).reshape(shape).T).ravel()
                    for a, shape in ((w0, (length, L)), (bias0, (length, 1)), (weight1, (length, length)),
                                     (bias1, (length, 1)), (weight2, (2, length)), (bias2, (2, 1)))]
        self.nn_L = L
        if lib().orc_set_refnn(self._h, L, length, *[_ptr(a) for a in self._nn]):
            raise ValueError("network sizes not supported by the oracle")

    def set_niter(self, niter):
        lib().orc_set_niter(self._h, int(niter))

    def pol_table(self):
        """Pol_zT_o3(5, iL:iB, 2) of pol_presim, flattened column-major."""
        out = np.zeros(40)
        lib().orc_pol_table(self._h, _ptr(out))
        return out

    def nn_format_input(self, U, l, n, L):
        """format_input of cell (l, n) (1-based) of full array U: (x[L], cst)."""
        x = np.zeros(L, dtype=np.float32)
        cst = lib().orc_nn_format_input(self._h, _ptr(np.ascontiguousarray(U)), int(l), int(n), int(L), _ptr(x))
        return x, bool(cst)

    def nn_forward(self, x):
        """forward (mod_NN.f90:165-181) of one input vector x[L]."""
        x = np.ascontiguousarray(x, dtype=np.float32)
        r = np.zeros(2, dtype=np.float32)
        lib().orc_nn_forward(self._h, _ptr(x), _ptr(r))
        return r

    def run_nn(self, U, nsteps, t=0.0, tmax=1e10):
        """run() with the reference network's per-step counters and
        CellGPO_priori; each step advances niter first (main.f90:147)."""
        assert U.flags.c_contiguous and U.dtype == np.float64
        dts = np.zeros(nsteps)
        counts = np.zeros(nsteps, dtype=np.int64)
        apriori = np.zeros(nsteps, dtype=np.int64)
        nan_passes = np.zeros(nsteps, dtype=np.int64)
        passes = np.zeros((nsteps, 3), dtype=np.int32)
        m = np.zeros((nsteps, self.nf, self.lf), dtype=np.int8)
        pri = np.zeros((nsteps, self.nf, self.lf), dtype=np.int8)
        tt = ctypes.c_double(t)
        rc = lib().orc_run2(self._h, _ptr(U), nsteps, ctypes.byref(tt), tmax, _ptr(dts), _ptr(counts), _ptr(m),
                            _ptr(passes), _ptr(apriori), _ptr(nan_passes), _ptr(pri))
        if rc:
            raise ValueError("the reference stops on this configuration")
        return dict(dt=dts, counts=counts, masks=m, passes=passes, apriori=apriori, nan_passes=nan_passes,
                    priori=pri, t=tt.value)

    def step(self, U, dt, t=0.0):
        """One time_stepping call from full array U at time t; returns Ur (U's
        ghosts are refilled as by the reference)."""
        Ur = np.zeros_like(U)
        lib().orc_step(self._h, _ptr(U), _ptr(Ur), dt, t)
        return Ur

    def set_initial_orders(self, orders):
        """Fixed a priori CellGPO (int8 [nf][lf]) for every stage, or None."""
        self._orders = None if orders is None else np.ascontiguousarray(orders, dtype=np.int8)
        lib().orc_set_initial_orders(self._h, _ptr(self._orders))

    def setdt(self, U, dtfinal=1e10):
        return lib().orc_setdt(self._h, _ptr(U), dtfinal)

    def step_stages(self, U, dt, t=0.0):
        outs = np.zeros((3, self.nf + 8, self.lf + 8, 4))
        masks = np.zeros((3, self.nf, self.lf), dtype=np.int8)
        lib().orc_step_stages(self._h, _ptr(U), dt, t, _ptr(outs), _ptr(masks))
        p = np.zeros(3, dtype=np.int32)
        lib().orc_passes(self._h, _ptr(p))
        return outs, masks, p


def expf_bits(first, n):
    """The C library's expf of the floats with bit patterns first .. first+n-1."""
    out = np.empty(n, dtype=np.float32)
    lib().orc_expf_bits(int(first), int(n), _ptr(out))
    return out


def riemann(flux, ul, ur, direction):
    """Riemann solver of the oracle (HLLC, LLF, HLL or unsplit) for one face."""
    ul = np.ascontiguousarray(ul, dtype=np.float64)
    ur = np.ascontiguousarray(ur, dtype=np.float64)
    F = np.zeros(4)
    lib().orc_riemann(FLUX_CODES[flux], _ptr(ul), _ptr(ur), int(direction), _ptr(F))
    return F


def hllc(ul, ur, direction):
    """HLLC flux of the oracle (mod_HLLC.f90:11-169) for one face."""
    ul = np.ascontiguousarray(ul, dtype=np.float64)
    ur = np.ascontiguousarray(ur, dtype=np.float64)
    F = np.zeros(4)
    lib().orc_hllc(_ptr(ul), _ptr(ur), int(direction), _ptr(F))
    return F
