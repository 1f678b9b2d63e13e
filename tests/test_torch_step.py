"""The FP64 PyTorch re-expression of the frozen-mask step (oracle/torch_step.py),
which checks the adjoint through torch.autograd, pinned to the C oracle: HLLC
pointwise on every branch, and whole SSP-RK3 stages with the oracle's own
final MOOD masks (the frozen step with those masks is the MOOD step)."""
import os
import sys

import numpy as np
import pytest
import torch

from conftest import ROOT, load_oracle

sys.path.insert(0, os.path.join(ROOT, "oracle"))
import torch_step as TS  # noqa: E402

sys.path.pop(0)


def random_pairs(rng, n):
    """Admissible state pairs spanning the four HLLC branches."""
    r = rng.uniform(0.1, 2.0, (n, 2))
    u = rng.uniform(-3.0, 3.0, (n, 2, 2))
    p = rng.uniform(0.05, 2.0, (n, 2))
    E = p / 0.4 + 0.5 * r * (u ** 2).sum(-1)
    U = np.stack([r, r * u[..., 0], r * u[..., 1], E], -1)
    return U[:, 0], U[:, 1]


@pytest.mark.parametrize("direction", [0, 1])
def test_torch_hllc_equals_oracle(direction):
    O = load_oracle()
    rng = np.random.default_rng(7 + direction)
    ul, ur = random_pairs(rng, 400)
    F = TS.hllc(torch.from_numpy(ul), torch.from_numpy(ur), direction).numpy()
    ref = np.array([O.hllc(ul[i], ur[i], direction) for i in range(len(ul))])
    assert np.allclose(F, ref, rtol=1e-14, atol=1e-14)
    # all four branches were exercised
    vx = ul[:, 1 + direction] / ul[:, 0]
    assert (np.abs(vx) > 2.0).any()


def test_torch_hllc_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    ul, ur = random_pairs(rng, 64)
    a = torch.from_numpy(ul).requires_grad_(True)
    b = torch.from_numpy(ur).requires_grad_(True)
    w = torch.from_numpy(rng.standard_normal((64, 4)))
    (TS.hllc(a, b, 0) * w).sum().backward()
    eps = 1e-6
    for k in range(4):
        d = np.zeros_like(ul)
        d[:, k] = eps
        fp = TS.hllc(torch.from_numpy(ul + d), torch.from_numpy(ur), 0)
        fm = TS.hllc(torch.from_numpy(ul - d), torch.from_numpy(ur), 0)
        fd = ((fp - fm) * w).sum(-1).numpy() / (2 * eps)
        assert np.allclose(a.grad.numpy()[:, k], fd, rtol=1e-6, atol=1e-7)


@pytest.mark.parametrize("case", [("four_quadrant", 5, 3, "neumann", 3, True), ("blast", 7, 4, "reflective", 1, True),
                                  ("shear_layer", 5, 3, "periodic", 2, False),
                                  ("four_quadrant", 3, 2, "periodic", 4, True)])
def test_frozen_torch_step_equals_oracle_step(gpm, case):
    kind, mord, ngp, bc, seed, lowered = case
    n = 32
    O = load_oracle()
    t = gpm.gp_tables(mord, ngp, n)
    o = O.Oracle(n, mord, ngp, bc, t)
    U = o.full(gpm.ic.make(kind, n, seed))
    o.run(U, 2)  # a few steps in, so that MOOD has lowered some orders
    dt = o.setdt(U)
    outs, masks, _ = o.step_stages(U.copy(), dt)
    T = TS.Tables(t, mord, ngp)
    dx = dy = 1.0 / n
    V = torch.from_numpy(U[4:-4, 4:-4].copy())
    got = TS.ssprk3(V, masks, dt, dx, dy, T, bc).numpy()
    ref = outs[2][4:-4, 4:-4]
    if lowered:
        assert (masks < mord).any(), "case should exercise lowered orders"
    assert np.max(np.abs(got - ref) / (np.abs(ref) + 1e-300)) < 1e-12
