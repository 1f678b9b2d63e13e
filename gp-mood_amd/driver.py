"""The reference's driver loop (main.f90:112-215) over a Solver, with its
outputs in the reference's HDF5 formats (h5io): snapshots at IO_freqStep /
IO_freqTime, the per-step diagnostic series, and restart from a snapshot.

Each step is gpm_run for one step (dtfinal = tmax - t, Setdt, the step,
t = t + dt, all on the device), followed by reading the step's counters,
as main.f90:145-167 records them.
"""
import os

import numpy as np

from . import capi, h5io

RK_FACTOR = {"fe": 1, "ssp_rk2": 2, "ssp_rk3": 3, "ssp_rk4": 5}  # main.f90:102-110


def run_problem(solver, problem, tmax, nmax, outdir=".", IO_freqStep=-100, IO_freqTime=-1.0, restart=None,
                nn_filename="", write_diagnostic=True, log=None, nn_size=None, restart_ic_ghosts=True):
    """Runs `solver` from its current state as main.f90 does: while t < tmax
    and niter <= nmax.  Writes output_<...>_<niter+100000>.h5 snapshots when
    niter is a multiple of IO_freqStep (> 0) or when t crosses a multiple of
    IO_freqTime (> 0, main.f90:185-192), and the diagnostic file at the end.
    With `restart` (a snapshot path) the state, t and niter come from the file
    (reader.f90) and the series start at iter_0 = niter + 1.  The reference
    fills U with the initial condition and Boundary_C before the read, which
    overwrites the interior only (main.f90:124-136, reader.f90:12-120), so its
    first Setdt after a restart also takes the minimum over the initial
    condition's ghost cells: the solver's state when run_problem is called
    plays the initial condition, and with restart_ic_ghosts (the default)
    its ghost cells go to that first Setdt (gpm_set_dt_ghosts); with
    restart_ic_ghosts=False they come from Boundary_C of the restart state.
    With method NN_GP_MOOD or NN_GP_MOOD_CC and nn_size = (L, length), the
    network is read from "<nn_filename>.txt" first (load_NN, main.f90:116);
    a network set on the solver beforehand is used otherwise.
    Returns (t, niter, list of snapshot paths)."""
    cfg = solver.config
    nf = cfg.lf if cfg.nf is None else cfg.nf
    m = h5io.meta(cfg.lf, nf, problem, cfg.method, dmp=cfg.DMP, cfl=cfg.CFL, nn_filename=nn_filename)
    t, niter = 0.0, 0
    if restart is not None:
        U_ic = solver.get_state() if restart_ic_ghosts else None  # InitialCdt + Boundary_C (main.f90:124-125)
        U, t, niter = h5io.read_restart(restart, cfg.lf, nf)
        solver.set_time(t)
        solver.set_state(np.ascontiguousarray(U[4:-4, 4:-4]))
        if U_ic is not None:
            solver.set_dt_ghosts(U_ic)
    solver.set_time(t)
    solver.set_niter(niter)  # the reference network predicts from niter 11 on
    if cfg.method in (capi.NN_GP_MOOD, capi.NN_GP_MOOD_CC) and nn_filename and nn_size:
        solver.load_reference_nn(nn_filename, *nn_size)  # load_NN, main.f90:116
    iter_0 = niter + 1
    diag = h5io.Diagnostics(m, RK_FACTOR[cfg.integrator], iter_0=iter_0)
    counter = 1
    written = []
    os.makedirs(outdir, exist_ok=True)

    def snapshot():
        U, g = solver.get_state(order=True)
        path = os.path.join(outdir, h5io.file_name(m, niter))
        # CellGPO_priori: the network's a priori orders (Mord without a network)
        pri = solver.cellgpo_priori() if cfg.method in (capi.NN_GP_MOOD, capi.NN_GP_MOOD_CC) else None
        h5io.write_output(path, m, U, g, niter, t, ordr_priori=pri, mord=cfg.Mord)
        written.append(path)
        if log:
            log(f"snapshot {path}")

    while t < tmax and niter <= nmax:
        niter += 1
        solver.run(1, tmax=tmax)
        st = solver.stats()
        dt = st["dt"]
        t = st["t"]
        diag.record(t, st["count_detected_cell"], st["count_detected_cell_a_priori"],
                    nn_nan_passes=st["nn_nan_passes"])
        if IO_freqStep > 0 and niter % IO_freqStep == 0:
            snapshot()
        if IO_freqTime > 0.0:
            if (t - float(counter) * IO_freqTime < 0.0) and (t + dt - float(counter) * IO_freqTime > 0.0):
                counter += 1
                snapshot()
    if write_diagnostic:
        diag.write(os.path.join(outdir, "diagnostic_" + h5io.file_name(m) + ".h5"))
    return t, niter, written
