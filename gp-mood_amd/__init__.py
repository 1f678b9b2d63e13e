"""gpmood_amd: MI355X-native GP-MOOD 2D Euler time step.

Python face of libgpmood.so.  The names mirror the reference's interface for
this path (slug-cfd/gp-mood, gp-mood-code/): ``time_stepping`` (the drop-in of
time_stepper.f90:12-23), ``setdt`` (physics.f90:78-129), the compile-time
``module parameters`` as :class:`Config`, and the state array
``U(4, lb:le, nb:ne)`` as a numpy array of shape ``(nf+8, lf+8, 4)``.

The numerics run only in the HIP library; when it is missing or no GPU is
present the calls raise.
"""
import ctypes
from dataclasses import dataclass

import numpy as np

from . import capi, h5io, ic  # noqa: F401
from .nn import MLP  # noqa: F401
from .capi import gp_tables  # noqa: F401

NGC = 4


@dataclass
class Config:
    """Runtime form of the reference's ``module parameters``
    (problems/2DRP3_GPMOOD.f90:8-46)."""

    lf: int
    nf: int = None
    Mord: int = 5
    ngp: int = 3
    BC_type: str = "neumann"
    CFL: float = 0.8
    DMP: bool = True
    U2: bool = True
    U2_tol: bool = True
    cascade: str = "reference"
    Lx: float = 1.0
    Ly: float = 1.0
    ell_o_dx: float = 12.0
    method: int = capi.GP_MOOD
    numFlux: str = "hllc"
    integrator: str = "ssp_rk3"
    problem: int = 0  # the reference's problem id; capi.PROBLEM_RT adds the gravity source

    def to_c(self):
        nf = self.lf if self.nf is None else self.nf
        return capi.GpmConfig(
            lf=self.lf, nf=nf, mord=self.Mord, ngp=self.ngp, method=self.method,
            integrator=capi.INTEGRATOR_IDS[self.integrator], num_flux=capi.FLUX_IDS[self.numFlux],
            bc_type=capi.BC_IDS[self.BC_type], dmp=int(self.DMP), u2=int(self.U2), u2_tol=int(self.U2_tol),
            cascade=capi.CASCADE_IDS[self.cascade], problem=int(self.problem), cfl=self.CFL, lx=self.Lx,
            ly=self.Ly, ell_o_dx=self.ell_o_dx)


class GpmError(RuntimeError):
    pass


def boundary_c(config, U, t_rk=0.0):
    """Boundary_C (BC.f90:18-413) in place on a host array U of shape
    (nf+8, lf+8, 4) for ``config``'s grid and BC_type at stage time t_rk.
    Host code of the library (the rule its kernels apply); needs no GPU."""
    if not (U.flags.c_contiguous and U.dtype == np.float64):
        raise ValueError("U must be C-contiguous float64")
    c = config.to_c()
    if U.shape != (c.nf + 2 * NGC, c.lf + 2 * NGC, 4):
        raise ValueError(f"state shape {U.shape} does not match the grid")
    rc = capi.load().gpm_boundary_c(ctypes.byref(c), float(t_rk), capi.ptr(U))
    if rc != 0:
        raise GpmError(f"gpm_boundary_c failed ({rc})")
    return U


def run_noalloc(config, tables, u_dev_ptr, nsteps, t=0.0, tmax=1e10, stream=None):
    """gpm_run_noalloc: nsteps steps of the driver loop on the device state at
    u_dev_ptr (interior cells, nf*lf*4 float64) without a context and without
    device allocation.  Returns the final time."""
    arr = {k: np.ascontiguousarray(tables[k], dtype=np.int32 if k == "ixiy" else np.float64) for k in
           ("zT", "d2x", "d2y", "gw", "ixiy")}
    c = config.to_c()
    t_end = ctypes.c_double()
    rc = capi.load().gpm_run_noalloc(ctypes.byref(c), *(capi.ptr(arr[k]) for k in ("zT", "d2x", "d2y", "gw", "ixiy")),
                                     ctypes.c_void_p(u_dev_ptr), int(nsteps), float(t), float(tmax),
                                     ctypes.byref(t_end), ctypes.c_void_p(stream) if stream else None)
    if rc != 0:
        raise GpmError(f"gpm_run_noalloc failed ({rc})")
    return t_end.value


class Solver:
    """One solver context on the current HIP device."""

    def __init__(self, config, tables=None):
        self.config = config
        self.lib = capi.load()
        self.n = self.lf = config.lf
        self.nf = config.lf if config.nf is None else config.nf
        self._cfg = config.to_c()
        h = ctypes.c_void_p()
        rc = self.lib.gpm_create(ctypes.byref(self._cfg), ctypes.byref(h))
        if rc != 0:
            raise GpmError(f"gpm_create failed ({rc})")
        self._h = h
        if tables is not None:
            self.set_gp_tables(tables)

    def close(self):
        if getattr(self, "_h", None):
            self.lib.gpm_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def _check(self, rc, what):
        if rc != 0:
            msg = self.lib.gpm_last_error(self._h)
            raise GpmError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

    # -- tables -----------------------------------------------------------
    def set_gp_tables(self, t):
        arr = {k: np.ascontiguousarray(t[k], dtype=np.int32 if k == "ixiy" else np.float64) for k in
               ("zT", "d2x", "d2y", "gw", "ixiy")}
        self._check(self.lib.gpm_set_gp_tables(self._h, *(capi.ptr(arr[k]) for k in ("zT", "d2x", "d2y", "gw", "ixiy"))),
                    "gpm_set_gp_tables")

    # -- state ------------------------------------------------------------
    def full_shape(self):
        return (self.nf + 2 * NGC, self.lf + 2 * NGC, 4)

    def set_state(self, U):
        """U: full array (nf+8, lf+8, 4) or interior (nf, lf, 4), float64.
        An interior array gets its ghost cells from the boundary rule at the
        current time (for copy boundaries the library ignores ghost cells)."""
        U = np.asarray(U, dtype=np.float64)
        if U.shape == (self.nf, self.lf, 4):
            import torch
            dev = torch.from_numpy(np.ascontiguousarray(U)).to("cuda")
            torch.cuda.synchronize()  # the copy runs on torch's stream, the library reads on its own
            self.set_state_device(dev.data_ptr())
            torch.cuda.synchronize()
            return
        U = np.ascontiguousarray(U)
        if U.shape != self.full_shape():
            raise ValueError(f"state shape {U.shape} != {self.full_shape()}")
        self._check(self.lib.gpm_set_state(self._h, capi.ptr(U)), "gpm_set_state")

    def get_state(self, order=False):
        U = np.zeros(self.full_shape())
        g = np.zeros(self.full_shape()[:2], dtype=np.int32) if order else None
        self._check(self.lib.gpm_get_state(self._h, capi.ptr(U), capi.ptr(g)), "gpm_get_state")
        return (U, g) if order else U

    def set_dt_ghosts(self, U):
        """Ghost cells of the full array U (nf+8, lf+8, 4) that the next Setdt
        reads once in place of the resident state's (the reference's restart:
        main.f90:124-151)."""
        U = np.ascontiguousarray(U, dtype=np.float64)
        if U.shape != self.full_shape():
            raise ValueError(f"state shape {U.shape} != {self.full_shape()}")
        self._check(self.lib.gpm_set_dt_ghosts(self._h, capi.ptr(U)), "gpm_set_dt_ghosts")

    def set_state_device(self, u_dev_ptr):
        self._check(self.lib.gpm_set_state_device(self._h, ctypes.c_void_p(u_dev_ptr)), "gpm_set_state_device")

    def get_state_device(self, u_dev_ptr, order_dev_ptr=None):
        self._check(self.lib.gpm_get_state_device(self._h, ctypes.c_void_p(u_dev_ptr) if u_dev_ptr else None,
                                                  ctypes.c_void_p(order_dev_ptr) if order_dev_ptr else None),
                    "gpm_get_state_device")

    # -- the reference's step interface ----------------------------------
    def setdt(self, dtfinal=1e30):
        """Setdt (physics.f90:78-129) on the resident state."""
        dt = ctypes.c_double()
        self._check(self.lib.gpm_compute_dt(self._h, dtfinal, ctypes.byref(dt)), "gpm_compute_dt")
        return dt.value

    def time_stepping(self, Uin, t, dt):
        """Drop-in of time_stepping(Uin, Uout) (time_stepper.f90:12-23).

        Uin (full array) has its ghost cells refilled in place, as the
        reference's Boundary_C(Uin) does.  Returns (Uout, CellGPO, stats)."""
        Uin = np.ascontiguousarray(Uin, dtype=np.float64)
        Uout = np.zeros(self.full_shape())
        g = np.zeros(self.full_shape()[:2], dtype=np.int32)
        st = capi.GpmStats()
        self._check(self.lib.gpm_time_stepping(self._h, capi.ptr(Uin), capi.ptr(Uout), t, dt, capi.ptr(g),
                                               ctypes.byref(st)), "gpm_time_stepping")
        return Uout, g, st.as_dict()

    def run(self, nsteps, tmax=1e10, return_dts=False):
        """Device-resident driver loop of main.f90:145-162 for nsteps steps."""
        dts = np.zeros(nsteps) if return_dts else None
        self._check(self.lib.gpm_run(self._h, nsteps, tmax, capi.ptr(dts)), "gpm_run")
        return dts

    def set_time(self, t):
        self._check(self.lib.gpm_set_time(self._h, t), "gpm_set_time")

    def stats(self):
        st = capi.GpmStats()
        self._check(self.lib.gpm_get_stats(self._h, ctypes.byref(st)), "gpm_get_stats")
        return st.as_dict()

    def reset_stats(self):
        self._check(self.lib.gpm_reset_stats(self._h), "gpm_reset_stats")

    def set_timing(self, on=True, every=1):
        """HIP-event timing of the stage and MOOD launches (kernel_time) of
        every step, or of every `every`-th step (a sample, with less event
        overhead on the stream)."""
        self._check(self.lib.gpm_set_timing(self._h, int(every) if on else 0), "gpm_set_timing")

    def kernel_time(self):
        a, b = ctypes.c_double(), ctypes.c_double()
        na, nb = ctypes.c_int64(), ctypes.c_int64()
        self._check(self.lib.gpm_kernel_time(self._h, ctypes.byref(a), ctypes.byref(na), ctypes.byref(b),
                                             ctypes.byref(nb)), "gpm_kernel_time")
        return {"stage_ms": a.value, "stage_launches": na.value, "mood_ms": b.value, "mood_launches": nb.value}

    def stream(self):
        return self.lib.gpm_stream(self._h)

    def set_mood_launch(self, cooperative=True):
        """MOOD correction passes as one persistent launch per stage (True;
        an ordinary launch, or a cooperative one with GPM_MOOD_ORD=0 at
        creation) or as per-pass launches (False); the results are the same."""
        self._check(self.lib.gpm_set_mood_launch(self._h, int(cooperative)), "gpm_set_mood_launch")

    def mood_launch(self):
        """True while the MOOD passes run as one persistent launch per stage."""
        v = ctypes.c_int32()
        self._check(self.lib.gpm_get_mood_launch(self._h, ctypes.byref(v)), "gpm_get_mood_launch")
        return bool(v.value)

    # -- learned a priori predictor (NN-GP-MOOD) ----------------------------
    def set_predictor(self, mlp, gate=True):
        """Attach an :class:`MLP` (None detaches): every RK stage starts from
        its predicted orders; the a posteriori check still runs."""
        self._predictor = mlp
        self._check(self.lib.gpm_set_predictor(self._h, mlp._h if mlp is not None else None, int(gate)),
                    "gpm_set_predictor")

    def set_apriori_orders(self, orders):
        """Fixed a priori CellGPO (int8 [nf][lf]) for every stage, or None."""
        arr = None if orders is None else np.ascontiguousarray(orders, dtype=np.int8)
        self._check(self.lib.gpm_set_apriori_orders(self._h, capi.ptr(arr)), "gpm_set_apriori_orders")

    # -- the reference's own network (NN_GP_MOOD / NN_GP_MOOD_CC) -----------
    def set_reference_nn(self, weight0, bias0, weight1, bias1, weight2, bias2):
        """The reference network's parameters as matrices, row i = output i
        (PyTorch's nn.Linear layout, neural_network/NN.py): weight0 (length, L),
        bias0 (length,), weight1 (length, length), bias1 (length,), weight2
        (2, length), bias2 (2,).  Steps after the tenth predict every cell's
        order with it (mod_NN.f90:14-58)."""
        w0 = np.asarray(weight0, dtype=np.float32)
        length, L = w0.shape
        fortran = [np.ascontiguousarray(np.asarray(a, dtype=np.float32).reshape(shape).T).ravel()
                   for a, shape in ((w0, (length, L)), (bias0, (length, 1)), (weight1, (length, length)),
                                    (bias1, (length, 1)), (weight2, (2, length)), (bias2, (2, 1)))]
        self._check(self.lib.gpm_set_reference_nn(self._h, int(L), int(length), *[capi.ptr(a) for a in fortran]),
                    "gpm_set_reference_nn")

    def load_reference_nn(self, nn_filename, L, length):
        """load_NN (mod_NN.f90:60-117): reads "<nn_filename>.txt"."""
        self._check(self.lib.gpm_load_reference_nn(self._h, str(nn_filename).encode(), int(L), int(length)),
                    "gpm_load_reference_nn")

    def set_reference_dump(self, inputs_ptr, labels_ptr, cells_ptr, capacity, count_ptr, nr_ptr, seed=0, L=52):
        """Training samples for the reference's trainer after every step's first
        stage (append_to_NN_datatset; see include/gpmood.h).  Device pointers;
        inputs_ptr = 0 switches it off."""
        vp = lambda p: ctypes.c_void_p(p) if p else None  # noqa: E731
        self._check(self.lib.gpm_set_reference_dump(self._h, vp(inputs_ptr), vp(labels_ptr), vp(cells_ptr),
                                                    int(capacity), vp(count_ptr), vp(nr_ptr),
                                                    int(seed) & ((1 << 64) - 1), int(L)), "gpm_set_reference_dump")

    def set_niter(self, niter):
        """niter before the next step (main.f90:147 advances it at each step)."""
        self._check(self.lib.gpm_set_niter(self._h, int(niter)), "gpm_set_niter")

    def cellgpo_priori(self):
        """CellGPO_priori (nf+8, lf+8) int32 after the last step."""
        out = np.zeros((self.nf + 2 * NGC, self.lf + 2 * NGC), dtype=np.int32)
        self._check(self.lib.gpm_get_cellgpo_priori(self._h, capi.ptr(out)), "gpm_get_cellgpo_priori")
        return out

    def nn_expf_bits(self, first, n, out_ptr):
        """The network's FP32 exp of the floats with bit patterns first ..
        first+n-1 into device memory out_ptr."""
        self._check(self.lib.gpm_nn_expf_bits(self._h, int(first), int(n), ctypes.c_void_p(out_ptr)),
                    "gpm_nn_expf_bits")

    def set_dump(self, x_ptr, y_ptr, capacity, count_ptr, keep_mord=0.05, seed=0, gated_only=True, stage_mask=1):
        """Training-data dump into device buffers (see include/gpmood.h)."""
        self._check(self.lib.gpm_set_dump(self._h, ctypes.c_void_p(x_ptr) if x_ptr else None,
                                          ctypes.c_void_p(y_ptr) if y_ptr else None, int(capacity),
                                          ctypes.c_void_p(count_ptr) if count_ptr else None, float(keep_mord),
                                          int(seed) & ((1 << 64) - 1), int(gated_only), int(stage_mask)),
                    "gpm_set_dump")

    def features(self, cells_ptr, n, x_ptr):
        self._check(self.lib.gpm_features(self._h, ctypes.c_void_p(cells_ptr), int(n), ctypes.c_void_p(x_ptr)),
                    "gpm_features")

    # -- differentiable form (frozen masks and dt) ---------------------------
    def record_masks(self, masks_ptr, max_steps):
        """Record the final CellGPO of each RK stage of the next max_steps steps
        into int8 device memory [max_steps][3][nf][lf] (0 stops recording)."""
        self._check(self.lib.gpm_record_masks(self._h, ctypes.c_void_p(masks_ptr) if masks_ptr else None,
                                              int(max_steps)), "gpm_record_masks")

    def recorded_steps(self):
        k = ctypes.c_int32()
        self._check(self.lib.gpm_recorded_steps(self._h, ctypes.byref(k)), "gpm_recorded_steps")
        return k.value

    def frozen_run(self, masks_ptr, dts, stage_states_ptr=None):
        """Advance the resident state len(dts) steps with the given per-stage
        orders and time steps (no detection); optionally store stage inputs."""
        dts = np.ascontiguousarray(dts, dtype=np.float64)
        self._check(self.lib.gpm_frozen_run(self._h, ctypes.c_void_p(masks_ptr), capi.ptr(dts), int(dts.size),
                                            ctypes.c_void_p(stage_states_ptr) if stage_states_ptr else None),
                    "gpm_frozen_run")

    def adjoint(self, stage_states_ptr, masks_ptr, dts, ubar_ptr):
        """In place on ubar (device, interior layout): dJ/dU after len(dts)
        frozen steps -> dJ/dU before them."""
        dts = np.ascontiguousarray(dts, dtype=np.float64)
        self._check(self.lib.gpm_adjoint(self._h, ctypes.c_void_p(stage_states_ptr), ctypes.c_void_p(masks_ptr),
                                         capi.ptr(dts), int(dts.size), ctypes.c_void_p(ubar_ptr)), "gpm_adjoint")


from . import adjoint, driver, refnn  # noqa: E402,F401
