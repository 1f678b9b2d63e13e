"""The reference's HDF5 file formats (ctypes binding of libgpmood_io.so,
include/gpmood_io.h): snapshots and restarts (output.f90:74-143,
reader.f90:12-120), per-step diagnostics (output.f90:145-214) and NN training
datasets (output.f90:216-300), with the reference's file names
(compute_metadata, output.f90:13-72).  Host code; needs no GPU.

Arrays follow the reference's memory layout seen from C: a state is
``(nf+8, lf+8, 4)`` float64 (ghost cells included, as gpm_get_state returns
it), an order mask ``(nf+8, lf+8)`` or ``(nf, lf)`` int32.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("GPM_IO_LIB_PATH") or os.path.join(HERE, "libgpmood_io.so")

# the reference's method ids (constants.f90:17) and problem ids (:27-34)
FOG, GP_MOOD, POL_MOOD, NN_GP_MOOD, NN_GP_MOOD_CC = 1, 2, 200, 203, 204
PROBLEMS = {"2DRP3": 28, "2DRP4": 49, "2DRP6": 38, "2DRP12": 30, "2DRP15": 31, "2DRP17": 50, "implosion": 29,
            "sedov": 39, "shu_osher": 44, "DMR": 35, "explosion": 36, "Mach800": 45, "RT": 51, "KH": 41}


class Meta(ctypes.Structure):
    _fields_ = [("lf", ctypes.c_int32), ("nf", ctypes.c_int32), ("problem", ctypes.c_int32),
                ("method", ctypes.c_int32), ("dmp", ctypes.c_int32), ("cfl", ctypes.c_double),
                ("nn_filename", ctypes.c_char_p)]


_lib = None


def load():
    """Load libgpmood_io.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python gp-mood_amd/build.py` (needs HDF5)")
    L = ctypes.CDLL(LIB_PATH)
    P, I, D, M = ctypes.c_void_p, ctypes.c_int32, ctypes.c_double, ctypes.POINTER(Meta)
    L.gpm_io_file_name.argtypes = [M, I, ctypes.c_char_p, I]
    L.gpm_io_write_output.argtypes = [ctypes.c_char_p, M, P, P, P, I, I, D]
    L.gpm_io_read_restart.argtypes = [ctypes.c_char_p, I, I, P, ctypes.POINTER(D), ctypes.POINTER(I)]
    L.gpm_io_write_diagnostic.argtypes = [ctypes.c_char_p, M, I, P, P, P, P, P, P, I, I, I]
    L.gpm_io_write_nn_dataset.argtypes = [ctypes.c_char_p, M, I, I, P, P, I, I, I]
    for f in ("gpm_io_file_name", "gpm_io_write_output", "gpm_io_read_restart", "gpm_io_write_diagnostic",
              "gpm_io_write_nn_dataset"):
        getattr(L, f).restype = I
    _lib = L
    return L


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


def _check(rc, what):
    if rc != 0:
        raise OSError(f"{what} failed ({rc})")


def meta(lf, nf, problem, method=GP_MOOD, dmp=True, cfl=0.8, nn_filename=""):
    """The compute_metadata inputs: grid, problem (name or id), method id, DMP
    switch, CFL and NN_filename."""
    pid = PROBLEMS[problem] if isinstance(problem, str) else int(problem)
    m = Meta(int(lf), int(nf), pid, int(method), int(bool(dmp)), float(cfl), nn_filename.encode())
    m._keep = m.nn_filename
    return m


def file_name(m, niter=None):
    """output_<problem>_<method>_CFL_<cfl>_<lf>_<nf>, with _<niter+100000>.h5
    for a snapshot."""
    buf = ctypes.create_string_buffer(512)
    _check(load().gpm_io_file_name(ctypes.byref(m), -1 if niter is None else int(niter), buf, 512), "gpm_io_file_name")
    return buf.value.decode()


def _interior(a, nf, lf):
    a = np.asarray(a)
    if a.shape[:2] == (nf + 8, lf + 8):
        a = a[4:-4, 4:-4]
    return np.ascontiguousarray(a, dtype=np.int32)


def write_output(path, m, U, ordr, niter, t, ordr_priori=None, mord=None):
    """A snapshot of U (nf+8, lf+8, 4) with order masks; ordr_priori None
    writes Mord (``mord``) everywhere, as CellGPO_priori of a method without a
    predictor."""
    U = np.ascontiguousarray(U, dtype=np.float64)
    if U.shape != (m.nf + 8, m.lf + 8, 4):
        raise ValueError(f"state shape {U.shape} does not match the grid")
    o = _interior(ordr, m.nf, m.lf)
    op = None if ordr_priori is None else _interior(ordr_priori, m.nf, m.lf)
    if op is None and mord is None:
        raise ValueError("mord is needed when ordr_priori is None")
    _check(load().gpm_io_write_output(os.fsencode(path), ctypes.byref(m), _ptr(U), _ptr(o), _ptr(op),
                                      int(mord or 0), int(niter), float(t)), "gpm_io_write_output")


def read_restart(path, lf, nf):
    """The restart read: (U (nf+8, lf+8, 4) with zero ghost cells, t, niter)."""
    U = np.zeros((nf + 8, lf + 8, 4))
    t, it = ctypes.c_double(), ctypes.c_int32()
    _check(load().gpm_io_read_restart(os.fsencode(path), int(lf), int(nf), _ptr(U), ctypes.byref(t), ctypes.byref(it)),
           "gpm_io_read_restart")
    return U, t.value, it.value


class Diagnostics:
    """The per-step series main.f90:160-167 records, for write_diagnostic:
    time and the detection percentages as float32, from the counters of each
    step (count / (rk_factor lf nf) * 100 in FP64, then rounded to float32)."""

    def __init__(self, m, rk_factor, iter_0=1):
        self.m, self.rk, self.iter_0 = m, int(rk_factor), int(iter_0)
        self.time, self.pct, self.pct_priori, self.pct_post = [], [], [], []
        self.nn_nan = []                 # steps_NN_produced_NAN of each step (0/1)
        self.count_steps_nn_nan = 0      # count_steps_NN_produced_NAN
        self.niter = iter_0 - 1

    def record(self, t, count, count_priori=0, count_posteriori=None, nn_nan_passes=0):
        """One step's counters; nn_nan_passes: the passes whose NN_DETECTION
        lowered a cell (mod_FE.f90:154-157)."""
        if count_posteriori is None:
            count_posteriori = count - count_priori
        self.nn_nan.append(1 if nn_nan_passes > 0 else 0)
        self.count_steps_nn_nan += int(nn_nan_passes)
        den = float(self.rk * self.m.nf * self.m.lf)
        self.niter += 1
        self.time.append(np.float32(t))
        self.pct.append(np.float32(float(count) * 100.0 / den))
        self.pct_priori.append(np.float32(float(count_priori) * 100.0 / den))
        self.pct_post.append(np.float32(float(count_posteriori) * 100.0 / den))

    def write(self, path, count_steps_nn_nan=None):
        """Writes the series as write_diagnostic does: the reference stores
        the first size - iter_0 of its recorded entries (output.f90:151-158),
        which is every step but the last."""
        n = max(0, len(self.time) - 1)
        arr = [np.ascontiguousarray(np.array(v[:n], dtype=np.float32)) for v in
               (self.time, self.pct, self.pct_priori, self.pct_post)]
        nan = np.ascontiguousarray(np.array(self.nn_nan[:n], dtype=np.int32))
        total = self.count_steps_nn_nan if count_steps_nn_nan is None else int(count_steps_nn_nan)
        _check(load().gpm_io_write_diagnostic(os.fsencode(path), ctypes.byref(self.m), n, *(_ptr(a) for a in arr),
                                              _ptr(nan), None, int(total), self.iter_0, self.niter),
               "gpm_io_write_diagnostic")


def write_nn_dataset(path, m, inputs, labels, n_overwrite=0, nr0=0, nr1=0):
    """A training dataset: inputs (size, L) and labels (size, 2), float32."""
    inputs = np.ascontiguousarray(inputs, dtype=np.float32)
    labels = np.ascontiguousarray(labels, dtype=np.float32)
    if inputs.ndim != 2 or labels.shape != (inputs.shape[0], 2):
        raise ValueError("inputs must be (size, L) and labels (size, 2)")
    _check(load().gpm_io_write_nn_dataset(os.fsencode(path), ctypes.byref(m), inputs.shape[0], inputs.shape[1],
                                          _ptr(inputs), _ptr(labels), int(n_overwrite), int(nr0), int(nr1)),
           "gpm_io_write_nn_dataset")
