"""Differentiable step on the GPU: the hand-written adjoint of K SSP-RK3 steps
with the MOOD masks and dts frozen from the forward run.

Checks: the HLLC reverse mode against torch.autograd of the FP64 PyTorch
re-expression (itself pinned to the C oracle); the frozen-mask forward against
the MOOD run it was recorded from (bit for bit); the full gradient at 64^2
against torch.autograd through the re-expressed steps; and at 512^2, K = 200
steps with a checkpoint every 20, central finite differences for J = sum
over a seeded 64x64 subregion of rho (u^2 + v^2): along 32 seeded box
directions (eps = 1e-6) and along 32 seeded Euclidean-unit dense directions
(fourth-order stencil, eps = 1e-3), relative error at most 1e-6 for each."""
import math
import os
import sys

import numpy as np
import pytest
import torch

from conftest import ROOT, has_gpu

sys.path.insert(0, os.path.join(ROOT, "oracle"))
import torch_step as TS  # noqa: E402

sys.path.pop(0)

pytestmark = pytest.mark.gpu


def _need_gpu():
    if not has_gpu():
        pytest.skip("no GPU")


def _pairs(rng, n):
    r = rng.uniform(0.1, 2.0, (n, 2))
    u = rng.uniform(-3.0, 3.0, (n, 2, 2))
    p = rng.uniform(0.05, 2.0, (n, 2))
    E = p / 0.4 + 0.5 * r * (u ** 2).sum(-1)
    U = np.stack([r, r * u[..., 0], r * u[..., 1], E], -1)
    return U[:, 0].copy(), U[:, 1].copy()


@pytest.mark.parametrize("direction", [0, 1])
def test_hllc_vjp_matches_autograd(gpm, direction):
    _need_gpu()
    print(f"[case hllc_vjp dir {direction}]", flush=True)
    rng = np.random.default_rng(11 + direction)
    ul, ur = _pairs(rng, 2000)
    fb = rng.standard_normal((2000, 4))
    a = torch.from_numpy(ul).requires_grad_(True)
    b = torch.from_numpy(ur).requires_grad_(True)
    (TS.hllc(a, b, direction) * torch.from_numpy(fb)).sum().backward()
    dev = [torch.from_numpy(x).cuda() for x in (ul, ur, fb)]
    ga, gb = gpm.adjoint.hllc_vjp(*dev, direction)
    for got, ref in ((ga.cpu().numpy(), a.grad.numpy()), (gb.cpu().numpy(), b.grad.numpy())):
        scale = np.abs(ref).max(axis=1, keepdims=True) + 1e-300
        assert np.max(np.abs(got - ref) / scale) < 1e-12


CASES = [("four_quadrant", 5, 3, "neumann", 8), ("blast", 7, 4, "reflective", 6), ("shear_layer", 3, 2, "periodic", 2),
         ("four_quadrant", 7, 4, "periodic", 5)]


def _solver(gpm, n, mord, ngp, bc):
    return gpm.Solver(gpm.Config(lf=n, Mord=mord, ngp=ngp, BC_type=bc))


@pytest.mark.parametrize("case", CASES)
def test_frozen_run_reproduces_mood_run(gpm, case):
    _need_gpu()
    kind, mord, ngp, bc, seed = case
    n, K = 64, 12
    print(f"[case frozen {case} n {n} K {K}]", flush=True)
    s = _solver(gpm, n, mord, ngp, bc)
    U0 = torch.from_numpy(gpm.ic.make(kind, n, seed)).cuda()
    UK, masks, dts = gpm.adjoint.record(s, U0, K)
    assert s.recorded_steps() == K
    if kind == "four_quadrant":
        assert (masks < mord).any(), "MOOD should have lowered orders"
    UF = gpm.adjoint.frozen(s, U0, masks, dts)
    assert torch.equal(UF, UK)
    s.close()


@pytest.mark.parametrize("case", CASES)
def test_adjoint_matches_autograd_64(gpm, case):
    _need_gpu()
    kind, mord, ngp, bc, seed = case
    n, K, w = 64, 3, 16
    print(f"[case adjoint-vs-autograd {case} n {n} K {K}]", flush=True)
    rng = np.random.default_rng(100 + seed)
    i0, j0 = (int(v) for v in rng.integers(0, n - w, 2))
    s = _solver(gpm, n, mord, ngp, bc)
    U0 = torch.from_numpy(gpm.ic.make(kind, n, seed)).cuda().requires_grad_(True)
    UK = gpm.adjoint.steps(s, U0, K, checkpoint_every=2)
    J = gpm.adjoint.kinetic_energy(UK, i0, j0, w)
    J.backward()
    g_gpu = U0.grad.cpu().numpy()
    _, masks, dts = gpm.adjoint.record(s, U0.detach(), K)
    T = TS.Tables(gpm.gp_tables(mord, ngp, n), mord, ngp)
    V0 = U0.detach().cpu().clone().requires_grad_(True)
    VK = TS.run(V0, masks.cpu().numpy(), dts, 1.0 / n, 1.0 / n, T, bc)
    assert np.max(np.abs(VK.detach().numpy() - UK.detach().cpu().numpy())) < 1e-12 * float(VK.abs().max())
    Jt = TS.kinetic(VK, i0, j0, w)
    Jt.backward()
    g_ref = V0.grad.numpy()
    assert abs(float(J) - float(Jt)) < 1e-12 * abs(float(Jt))
    err = np.max(np.abs(g_gpu - g_ref)) / np.max(np.abs(g_ref))
    print(f"  max |g - g_autograd| / max |g_autograd| = {err:.3e}", flush=True)
    assert err < 1e-10
    s.close()


def _fd_setup(gpm):
    n, K, m, w = 512, 200, 20, 64
    rng = np.random.default_rng(2024)
    i0, j0 = (int(v) for v in rng.integers(0, n - w, 2))
    s = _solver(gpm, n, 5, 3, "neumann")
    U0 = torch.from_numpy(gpm.ic.make("four_quadrant", n, 3)).cuda().requires_grad_(True)
    UK = gpm.adjoint.steps(s, U0, K, checkpoint_every=m)
    J = gpm.adjoint.kinetic_energy(UK, i0, j0, w)
    J.backward()
    _, masks, dts = gpm.adjoint.record(s, U0.detach(), K)
    assert (masks < 5).any()

    def cells(U):
        S = U[j0:j0 + w, i0:i0 + w]
        return ((S[..., 1] * S[..., 1] + S[..., 2] * S[..., 2]) / S[..., 0]).cpu().numpy().ravel()

    def run(d, e):
        return cells(gpm.adjoint.frozen(s, U0.detach() + e * d, masks, dts))

    def fd(d, eps=1e-6, order=2):
        """Central difference quotient of J along d: second order
        (J(e) - J(-e)) / 2e or fourth order (8[J(e) - J(-e)] - [J(2e) - J(-2e)]) / 12e.
        J's differences are taken cell by cell and summed exactly (math.fsum),
        so the quotient carries no cancellation of the two sums."""
        d1 = run(d, eps) - run(d, -eps)
        if order == 2:
            return math.fsum(d1) / (2 * eps)
        d2 = run(d, 2 * eps) - run(d, -2 * eps)
        return math.fsum(8.0 * d1 - d2) / (12 * eps)

    return s, U0, U0.grad.detach(), fd, (i0, j0, w)


def test_adjoint_finite_differences_512(gpm):
    """512^2 GP-R2, four-quadrant IC, K = 200 steps, checkpoints every 20:
    32 seeded unit directions (max norm 1: every component of a seeded
    128x128-cell box moves by +-eps) against central differences, eps = 1e-6."""
    _need_gpu()
    print("[case adjoint FD n 512 K 200 m 20, box directions]", flush=True)
    s, U0, g, fd, (i0, j0, w) = _fd_setup(gpm)
    n = U0.shape[0]
    gen = torch.Generator(device="cuda")
    gen.manual_seed(7)
    rng = np.random.default_rng(99)
    errs = []
    for q in range(32):
        d = torch.where(torch.rand(U0.shape, device="cuda", generator=gen, dtype=torch.float64) < 0.5, -1.0, 1.0)
        d = d.to(torch.float64)
        a0 = int(np.clip(i0 + w // 2 - 64 + rng.integers(-32, 33), 0, n - 128))
        b0 = int(np.clip(j0 + w // 2 - 64 + rng.integers(-32, 33), 0, n - 128))
        box = torch.zeros_like(d)
        box[b0:b0 + 128, a0:a0 + 128] = 1.0
        d = d * box
        f = fd(d)
        ad = float((g * d).sum())
        errs.append(abs(f - ad) / abs(f))
    print(f"  FD relative errors: max {max(errs):.3e}, median {np.median(errs):.3e}", flush=True)
    assert max(errs) < 1e-6
    s.close()


def test_adjoint_finite_differences_512_dense(gpm):
    """Same set-up, 32 seeded Euclidean-unit Gaussian directions over the whole
    state (2^20 components of ~1e-3 each), against the fourth-order central
    difference with eps = 1e-3, i.e. each component moves by ~1e-6 as in the
    box-direction test.  At eps = 1e-6 (components moved by ~1e-9) the FP64
    round-off of the two 200-step runs dominates the quotient (errors ~1e-6 to
    1e-5); the sweep over eps on the first directions, printed, shows that
    floor falling as 1/eps until truncation takes over."""
    _need_gpu()
    print("[case adjoint FD n 512 K 200 m 20, dense directions]", flush=True)
    s, U0, g, fd, _ = _fd_setup(gpm)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(8)
    errs = []
    for q in range(32):
        d = torch.randn(U0.shape, dtype=torch.float64, device="cuda", generator=gen)
        d /= torch.linalg.vector_norm(d)
        ad = float((g * d).sum())
        if q < 3:
            sweep = {f"{o}:{e:g}": abs(fd(d, e, o) - ad) / abs(ad) for o in (2, 4) for e in (1e-6, 1e-4, 1e-3, 1e-2)}
            print(f"  direction {q} relative error by order:eps " +
                  ", ".join(f"{k} {v:.1e}" for k, v in sweep.items()), flush=True)
        errs.append(abs(fd(d, 1e-3, 4) - ad) / abs(ad))
    print(f"  FD relative errors (order 4, eps 1e-3): max {max(errs):.3e}, median {np.median(errs):.3e}", flush=True)
    assert max(errs) <= 1e-6
    s.close()
