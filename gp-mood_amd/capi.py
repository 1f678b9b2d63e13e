"""ctypes binding of libgpmood.so (the C ABI in include/gpmood.h).

torch is imported before the library is loaded so that the process holds one
HIP runtime: the library then binds to the runtime torch already loaded, and
device pointers of torch tensors are valid inside the library's kernels.
"""
import ctypes
import os

import numpy as np

try:  # one HIP runtime per process (see module docstring)
    import torch  # noqa: F401
except Exception:  # pragma: no cover - torch is part of the image
    torch = None

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("GPM_LIB_PATH") or os.path.join(HERE, "libgpmood.so")

FOG = 1
GP_MOOD = 2
POL_MOOD = 200
UNLIM_POL = 201
GP_SE = 202
NN_GP_MOOD = 203
NN_GP_MOOD_CC = 204
FE, SSP_RK2, SSP_RK3, SSP_RK4 = 8, 9, 10, 12
INTEGRATOR_IDS = {"fe": FE, "ssp_rk2": SSP_RK2, "ssp_rk3": SSP_RK3, "ssp_rk4": SSP_RK4}
HLLC, LLF, HLL, UNSPLIT = 100, 101, 102, 103
FLUX_IDS = {"hllc": HLLC, "llf": LLF, "hll": HLL, "unsplit": UNSPLIT}
NEUMANN, PERIODIC, DIRICHLET, REFLECTIVE, KH_BC = 17, 18, 19, 20, 403
MACH800_BC, DOUBLEMACH800_BC, RMI_BC, DMR_BC, RT_BC = 201, 301, 401, 35, 402
BC_IDS = {"neumann": NEUMANN, "periodic": PERIODIC, "reflective": REFLECTIVE, "kh": KH_BC, "dirichlet": DIRICHLET,
          "mach800": MACH800_BC, "doublemach800": DOUBLEMACH800_BC, "rmi": RMI_BC, "dmr": DMR_BC, "rt": RT_BC}
PROBLEM_RT = 51
CASCADE_IDS = {"reference": 0, "full": 1}

# Every symbol include/gpmood.h declares.
EXPORTS = [
    "gpm_create", "gpm_destroy", "gpm_last_error", "gpm_set_gp_tables", "gpm_gp_tables", "gpm_boundary_c",
    "gpm_set_state", "gpm_get_state", "gpm_set_dt_ghosts", "gpm_set_state_device", "gpm_get_state_device",
    "gpm_compute_dt", "gpm_time_stepping", "gpm_run", "gpm_run_noalloc", "gpm_get_stats", "gpm_reset_stats",
    "gpm_set_time", "gpm_stream", "gpm_set_timing", "gpm_kernel_time", "gpm_set_mood_launch", "gpm_get_mood_launch",
    "gpm_mlp_create", "gpm_mlp_destroy", "gpm_mlp_last_error", "gpm_mlp_num_params", "gpm_mlp_set_params",
    "gpm_mlp_get_params", "gpm_mlp_get_grads", "gpm_mlp_forward_backward", "gpm_mlp_adamw_step",
    "gpm_mlp_train_step", "gpm_mlp_logits", "gpm_set_predictor", "gpm_set_dump", "gpm_features",
    "gpm_set_apriori_orders", "gpm_get_apriori_orders",
    "gpm_set_reference_nn", "gpm_load_reference_nn", "gpm_set_niter", "gpm_get_cellgpo_priori", "gpm_nn_expf_bits",
    "gpm_read_reference_nn", "gpm_nn_expf_bits_host", "gpm_ddiv", "gpm_qdiv", "gpm_set_reference_dump", "gpm_set_reference_dataset",
    "gpm_append_reference_dataset",
    "gpm_record_masks", "gpm_recorded_steps", "gpm_frozen_run", "gpm_adjoint", "gpm_hllc_vjp",
]


class GpmConfig(ctypes.Structure):
    _fields_ = [
        ("lf", ctypes.c_int32), ("nf", ctypes.c_int32), ("mord", ctypes.c_int32), ("ngp", ctypes.c_int32),
        ("method", ctypes.c_int32), ("integrator", ctypes.c_int32), ("num_flux", ctypes.c_int32),
        ("bc_type", ctypes.c_int32), ("dmp", ctypes.c_int32), ("u2", ctypes.c_int32), ("u2_tol", ctypes.c_int32),
        ("cascade", ctypes.c_int32), ("problem", ctypes.c_int32), ("cfl", ctypes.c_double), ("lx", ctypes.c_double),
        ("ly", ctypes.c_double),
        ("ell_o_dx", ctypes.c_double),
    ]


class GpmStats(ctypes.Structure):
    _fields_ = [
        ("steps", ctypes.c_int64), ("count_detected_cell", ctypes.c_int64), ("detected_total", ctypes.c_int64),
        ("passes", ctypes.c_int32 * 3), ("max_passes", ctypes.c_int32), ("second_pass_cells", ctypes.c_int64),
        ("a_priori_lowered", ctypes.c_int64), ("t", ctypes.c_double), ("dt", ctypes.c_double),
        ("count_detected_cell_a_priori", ctypes.c_int64), ("nn_nan_passes", ctypes.c_int64),
        ("nn_nan_passes_total", ctypes.c_int64), ("niter", ctypes.c_int64),
    ]

    def as_dict(self):
        return {
            "steps": self.steps, "count_detected_cell": self.count_detected_cell,
            "detected_total": self.detected_total, "passes": list(self.passes), "max_passes": self.max_passes,
            "second_pass_cells": self.second_pass_cells, "a_priori_lowered": self.a_priori_lowered,
            "t": self.t, "dt": self.dt, "count_detected_cell_a_priori": self.count_detected_cell_a_priori,
            "nn_nan_passes": self.nn_nan_passes, "nn_nan_passes_total": self.nn_nan_passes_total, "niter": self.niter,
        }


_lib = None


def load():
    """Load libgpmood.so; raises if it has not been built (there is no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python gp-mood_amd/build.py` "
                           "(or __graft_entry__.build())")
    L = ctypes.CDLL(LIB_PATH)
    P, I, D = ctypes.c_void_p, ctypes.c_int32, ctypes.c_double
    sig = {
        "gpm_create": ([ctypes.POINTER(GpmConfig), ctypes.POINTER(P)], I),
        "gpm_destroy": ([P], I),
        "gpm_last_error": ([P], ctypes.c_char_p),
        "gpm_set_gp_tables": ([P, P, P, P, P, P], I),
        "gpm_gp_tables": ([I, I, D, D, D, I, I, P, P, P, P, P], I),
        "gpm_boundary_c": ([ctypes.POINTER(GpmConfig), D, P], I),
        "gpm_set_state": ([P, P], I),
        "gpm_get_state": ([P, P, P], I),
        "gpm_set_dt_ghosts": ([P, P], I),
        "gpm_set_state_device": ([P, P], I),
        "gpm_get_state_device": ([P, P, P], I),
        "gpm_compute_dt": ([P, D, ctypes.POINTER(D)], I),
        "gpm_time_stepping": ([P, P, P, D, D, P, ctypes.POINTER(GpmStats)], I),
        "gpm_run": ([P, I, D, P], I),
        "gpm_run_noalloc": ([ctypes.POINTER(GpmConfig), P, P, P, P, P, P, I, D, D, ctypes.POINTER(D), P], I),
        "gpm_get_stats": ([P, ctypes.POINTER(GpmStats)], I),
        "gpm_reset_stats": ([P], I),
        "gpm_set_time": ([P, D], I),
        "gpm_stream": ([P], P),
        "gpm_set_timing": ([P, I], I),
        "gpm_set_mood_launch": ([P, I], I),
        "gpm_get_mood_launch": ([P, ctypes.POINTER(I)], I),
        "gpm_kernel_time": ([P, ctypes.POINTER(D), ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(D),
                             ctypes.POINTER(ctypes.c_int64)], I),
        "gpm_mlp_create": ([I, I, I, ctypes.POINTER(P)], I),
        "gpm_mlp_destroy": ([P], I),
        "gpm_mlp_last_error": ([P], ctypes.c_char_p),
        "gpm_mlp_num_params": ([P], ctypes.c_int64),
        "gpm_mlp_set_params": ([P, P], I),
        "gpm_mlp_get_params": ([P, P], I),
        "gpm_mlp_get_grads": ([P, P], I),
        "gpm_mlp_forward_backward": ([P, P, P, I, ctypes.POINTER(D), ctypes.POINTER(D)], I),
        "gpm_mlp_adamw_step": ([P, D, D, D, D, D], I),
        "gpm_mlp_train_step": ([P, P, P, I, D, D, ctypes.POINTER(D), ctypes.POINTER(D)], I),
        "gpm_mlp_logits": ([P, P, I, P], I),
        "gpm_set_predictor": ([P, P, I], I),
        "gpm_set_dump": ([P, P, P, ctypes.c_int64, P, D, ctypes.c_uint64, I, I], I),
        "gpm_features": ([P, P, I, P], I),
        "gpm_set_apriori_orders": ([P, P], I),
        "gpm_get_apriori_orders": ([P, P], I),
        "gpm_set_reference_nn": ([P, I, I, P, P, P, P, P, P], I),
        "gpm_load_reference_nn": ([P, ctypes.c_char_p, I, I], I),
        "gpm_set_niter": ([P, ctypes.c_int64], I),
        "gpm_get_cellgpo_priori": ([P, P], I),
        "gpm_nn_expf_bits": ([P, ctypes.c_uint32, ctypes.c_int64, P], I),
        "gpm_read_reference_nn": ([ctypes.c_char_p, I, I, P, P, P, P, P, P], I),
        "gpm_nn_expf_bits_host": ([ctypes.c_uint32, ctypes.c_int64, P], I),
        "gpm_ddiv": ([ctypes.c_int64, P, P, P], I),
        "gpm_qdiv": ([ctypes.c_int64, P, P, P], I),
        "gpm_set_reference_dump": ([P, P, P, P, ctypes.c_int64, P, P, ctypes.c_uint64, I], I),
        "gpm_set_reference_dataset": ([P, I, I, I], I),
        "gpm_append_reference_dataset": ([P, P, P, I, P, P, P, P], I),
        "gpm_record_masks": ([P, P, I], I),
        "gpm_recorded_steps": ([P, P], I),
        "gpm_frozen_run": ([P, P, P, I, P], I),
        "gpm_adjoint": ([P, P, P, P, I, P], I),
        "gpm_hllc_vjp": ([I, I, P, P, P, P, P], I),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


def gp_tables(mord, ngp, lf, nf=None, ell_o_dx=12.0, lx=1.0, ly=1.0):
    """GP tables of GP_presim_sphere (GP_init.f90:26-357) in the reference layouts."""
    nf = lf if nf is None else nf
    sz = {1: 1, 3: 5, 5: 13, 7: 25}[mord]
    zT = np.zeros(mord * sz * 4 * ngp)
    d2x = np.zeros(mord * sz)
    d2y = np.zeros(mord * sz)
    gw = np.zeros(25)
    ixiy = np.zeros(7 * 25 * 2, dtype=np.int32)
    rc = load().gpm_gp_tables(mord, ngp, ell_o_dx, lx, ly, lf, nf, ptr(zT), ptr(d2x), ptr(d2y), ptr(gw), ptr(ixiy))
    if rc != 0:
        raise ValueError(f"gpm_gp_tables failed ({rc})")
    return {"zT": zT, "d2x": d2x, "d2y": d2y, "gw": gw, "ixiy": ixiy}
