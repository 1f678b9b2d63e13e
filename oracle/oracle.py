"""ctypes wrapper of the CPU restatement in oracle/gpm_oracle.c.

TEST INFRASTRUCTURE: only tests/, __graft_entry__.smoke() and bench.py's CPU
baseline leg use this module, as the checker.  It never backs the product path.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
BC_CODES = {"neumann": 0, "periodic": 1, "reflective": 2, "kh": 3}
FLUX_CODES = {"hllc": 0, "hll": 1, "unsplit": 2}
CASCADES = {"reference": 0, "full": 1}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle library not built: run `make -C oracle`")
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        L.orc_create.restype = P
        L.orc_create.argtypes = [ctypes.c_int] * 8 + [ctypes.c_double] * 3 + [P] * 5
        L.orc_destroy.argtypes = [P]
        L.orc_set_flux.argtypes = [P, ctypes.c_int]
        L.orc_riemann.argtypes = [ctypes.c_int, P, P, ctypes.c_int, P]
        L.orc_bc.argtypes = [P, P]
        L.orc_setdt.restype = ctypes.c_double
        L.orc_setdt.argtypes = [P, P, ctypes.c_double]
        L.orc_run.argtypes = [P, P, ctypes.c_int, P, ctypes.c_double, P, P, P, P]
        L.orc_step_stages.argtypes = [P, P, ctypes.c_double, P, P]
        L.orc_count_detected.restype = ctypes.c_long
        L.orc_count_detected.argtypes = [P]
        L.orc_passes.argtypes = [P, P]
        L.orc_hllc.argtypes = [P, P, ctypes.c_int, P]
        L.orc_set_initial_orders.argtypes = [P, P]
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


class Oracle:
    """The reference step on CPU.  ``tables`` is a dict with keys zT, d2x, d2y,
    gw, ixiy in the reference's Fortran layouts (flattened column-major)."""

    def __init__(self, n, mord, ngp, bc, tables, dmp=True, u2=True, u2tol=True,
                 cascade="reference", cfl=0.8, flux="hllc"):
        self.n = n
        self.tables = {k: np.ascontiguousarray(v) for k, v in tables.items()}
        t = self.tables
        dx = 1.0 / n
        self._h = lib().orc_create(n, mord, ngp, BC_CODES[bc], int(dmp), int(u2), int(u2tol),
                                   CASCADES[cascade], cfl, dx, dx,
                                   _ptr(t["zT"].astype(np.float64)), _ptr(t["d2x"].astype(np.float64)),
                                   _ptr(t["d2y"].astype(np.float64)), _ptr(t["gw"].astype(np.float64)),
                                   _ptr(t["ixiy"].astype(np.int32)))
        self._keep = [t[k] for k in t]
        lib().orc_set_flux(self._h, FLUX_CODES[flux])

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_destroy(self._h)
            self._h = None

    def full(self, interior):
        n = self.n
        U = np.zeros((n + 8, n + 8, 4), dtype=np.float64)
        U[4:n + 4, 4:n + 4] = interior
        lib().orc_bc(self._h, _ptr(U))
        return U

    def run(self, U, nsteps, t=0.0, tmax=1e10, masks=False):
        """Advance full array U (modified in place) by nsteps steps."""
        assert U.flags.c_contiguous and U.dtype == np.float64
        dts = np.zeros(nsteps)
        counts = np.zeros(nsteps, dtype=np.int64)
        passes = np.zeros((nsteps, 3), dtype=np.int32)
        m = np.zeros((nsteps, self.n, self.n), dtype=np.int8) if masks else None
        tt = ctypes.c_double(t)
        lib().orc_run(self._h, _ptr(U), nsteps, ctypes.byref(tt), tmax, _ptr(dts), _ptr(counts),
                      _ptr(m), _ptr(passes))
        return dict(dt=dts, counts=counts, masks=m, passes=passes, t=tt.value)

    def set_initial_orders(self, orders):
        """Fixed a priori CellGPO (int8 [N][N]) for every stage, or None."""
        self._orders = None if orders is None else np.ascontiguousarray(orders, dtype=np.int8)
        lib().orc_set_initial_orders(self._h, _ptr(self._orders))

    def setdt(self, U, dtfinal=1e10):
        return lib().orc_setdt(self._h, _ptr(U), dtfinal)

    def step_stages(self, U, dt):
        n = self.n
        outs = np.zeros((3, n + 8, n + 8, 4))
        masks = np.zeros((3, n, n), dtype=np.int8)
        lib().orc_step_stages(self._h, _ptr(U), dt, _ptr(outs), _ptr(masks))
        p = np.zeros(3, dtype=np.int32)
        lib().orc_passes(self._h, _ptr(p))
        return outs, masks, p


def riemann(flux, ul, ur, direction):
    """Riemann solver of the oracle (HLLC, HLL or unsplit) for one face."""
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
