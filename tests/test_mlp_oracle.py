"""The MLP restatement (oracle/mlp_oracle.py) pinned against PyTorch FP64
autograd rounding to BF16 at the same points (tests/bf16_torch.py): logits,
loss and every weight and bias gradient per tensor, and a planted mistake in
the oracle that must fail the same comparison.  AdamW against
torch.optim.AdamW."""
import os
import sys

import numpy as np
import pytest

from conftest import ROOT

import bf16_torch as T


def mlp_oracle():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    try:
        import mlp_oracle as M
    finally:
        sys.path.pop(0)
    return M


def rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def init_params(in_f, C, seed):
    """torch.nn.Linear's default range, U(-1/sqrt(fan_in), 1/sqrt(fan_in))."""
    rng = np.random.default_rng(seed)
    dims = (in_f,) + T.HIDDEN + (C,)
    out = []
    for a, b in zip(dims[:-1], dims[1:]):
        r = 1.0 / np.sqrt(a)
        out += [rng.uniform(-r, r, size=a * b), rng.uniform(-r, r, size=b)]
    return np.concatenate(out).astype(np.float32)


def case(in_f, C, B, seed):
    rng = np.random.default_rng(seed)
    p = init_params(in_f, C, seed + 100)
    x = rng.uniform(-1, 1, size=(B, in_f)).astype(np.float32)
    y = rng.integers(0, C, size=B).astype(np.int32)
    return p, x, y


def compare(M, p, x, y, in_f, C):
    z_o, _ = M.forward(p, in_f, C, x)
    loss_o, _, g_o = M.forward_backward(p, in_f, C, x, y)
    z_t, loss_t, g_t = T.forward_backward(p, in_f, C, x, y)
    errs = {"logits": rel(z_o, z_t), "loss": abs(loss_o - loss_t) / abs(loss_t)}
    for li, ((gw, gb), (tw, tb)) in enumerate(zip(T.split_flat(g_o, in_f, C), g_t)):
        errs[f"W{li}"] = rel(gw, tw)
        errs[f"b{li}"] = rel(gb, tb)
    return errs


@pytest.mark.parametrize("in_f,C,B,seed", [(100, 3, 512, 1), (196, 4, 512, 2), (52, 2, 300, 3)])
def test_oracle_vs_torch_bf16_autograd(in_f, C, B, seed):
    M = mlp_oracle()
    p, x, y = case(in_f, C, B, seed)
    errs = compare(M, p, x, y, in_f, C)
    bad = {k: v for k, v in errs.items() if not v <= 1e-5}
    assert not bad, errs


def test_planted_rounding_mistake_fails(monkeypatch):
    """With the oracle's BF16 rounding removed (a plausible slip: rounding at
    none of the points) the same per-tensor comparison must fail."""
    M = mlp_oracle()
    p, x, y = case(100, 3, 512, 1)
    monkeypatch.setattr(M, "bf16", lambda v: np.asarray(v, dtype=np.float64).astype(np.float32))
    errs = compare(M, p, x, y, 100, 3)
    assert max(errs.values()) > 1e-3, errs


def test_planted_mask_mistake_fails(monkeypatch):
    """A backward pass that takes the wrong layer's activation (an off-by-one
    index: layer 1's in place of layer 2's, both 128 wide) for one weight
    gradient and mask must fail the per-tensor comparison."""
    M = mlp_oracle()
    p, x, y = case(100, 3, 512, 4)
    real = M.forward

    def shifted(params, in_f, C, xx):
        z, acts = real(params, in_f, C, xx)
        return z, [acts[0], acts[1], acts[1], acts[3]]

    monkeypatch.setattr(M, "forward", shifted)
    loss_o, _, g_o = M.forward_backward(p, 100, 3, x, y)
    _, _, g_t = T.forward_backward(p, 100, 3, x, y)
    errs = [rel(gw, tw) for (gw, _), (tw, _) in zip(T.split_flat(g_o, 100, 3), g_t)]
    assert max(errs) > 1e-2, errs


def test_adamw_vs_torch():
    import torch
    M = mlp_oracle()
    rng = np.random.default_rng(5)
    p = rng.standard_normal(1000).astype(np.float32)
    pt = torch.nn.Parameter(torch.from_numpy(p.copy()))
    opt = torch.optim.AdamW([pt], lr=1e-3, betas=(0.9, 0.999), eps=1e-8, weight_decay=1e-2)
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    for step in range(1, 6):
        g = rng.standard_normal(1000).astype(np.float32)
        p, m, v = M.adamw(p, g, m, v, step, 1e-3)
        pt.grad = torch.from_numpy(g)
        opt.step()
        assert np.allclose(p, pt.detach().numpy(), rtol=1e-6, atol=1e-8), step
