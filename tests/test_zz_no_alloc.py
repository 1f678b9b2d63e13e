"""The allocation-free entry point gpm_run_noalloc against gpm_run on a
context (the existing entry point it is a variant of): the same final state
bit for bit and the same final time, on two calls in a row, the second on a
smaller grid (the scratch memory keeps the first call's contents)."""
import numpy as np
import pytest

from conftest import has_gpu

pytestmark = pytest.mark.gpu

# (n, Mord, ngp, BC_type, IC kind, seed, steps, cascade): the second call is smaller
CALLS = [(256, 5, 3, "neumann", "four_quadrant", 3, 20, "reference"),
         (128, 7, 4, "reflective", "blast", 5, 15, "full")]


def _run_context(gpm, cfg, tables, U0, nsteps):
    import torch
    s = gpm.Solver(cfg, tables)
    torch.cuda.synchronize()
    s.set_state_device(U0.data_ptr())
    s.set_time(0.0)
    s.run(nsteps)
    U = torch.empty_like(U0)
    s.get_state_device(U.data_ptr())
    t = s.stats()["t"]
    torch.cuda.synchronize()
    s.close()
    return U, t


def test_run_noalloc_matches_run(gpm):
    if not has_gpu():
        pytest.skip("no GPU")
    import torch
    for n, mord, ngp, bc, kind, seed, nsteps, cascade in CALLS:
        print(f"[case no-alloc {n}^2 Mord {mord} {bc} {kind} seed {seed}, {nsteps} steps]", flush=True)
        cfg = gpm.Config(lf=n, Mord=mord, ngp=ngp, BC_type=bc, cascade=cascade)
        tables = gpm.gp_tables(mord, ngp, n)
        U0 = torch.from_numpy(gpm.ic.make(kind, n, seed)).cuda()
        Uref, tref = _run_context(gpm, cfg, tables, U0, nsteps)
        U = U0.clone()
        torch.cuda.synchronize()
        t_end = gpm.run_noalloc(cfg, tables, U.data_ptr(), nsteps, 0.0, 1e10)
        assert t_end == tref
        assert bool(torch.equal(U, Uref)), f"max |diff| {float((U - Uref).abs().max())}"
        assert not bool(torch.equal(U, U0))
