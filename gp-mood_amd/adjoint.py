"""Initial-state sensitivities of K GP-MOOD SSP-RK3 steps.

The forward pass is the plain GPU solver (the reference's step, bit for bit);
it records the final order mask of every RK stage and every dt, and keeps the
state every ``checkpoint_every`` steps.  The backward pass freezes those masks
and dts: per checkpoint segment, from the last one back, it recomputes the
stage inputs with the frozen-mask forward (bit-identical to the recorded run)
and applies the hand-written adjoint kernels step by step.  No counterpart in
the reference.

    U0 = torch.tensor(..., dtype=torch.float64, device="cuda", requires_grad=True)
    UK = gpmood_amd.adjoint.steps(solver, U0, K=200, checkpoint_every=20)
    J = gpmood_amd.adjoint.kinetic_energy(UK, i0, j0, 64)
    J.backward()          # U0.grad = dJ/dU0

States are interior arrays of shape (nf, lf, 4), float64, on the solver's
device.
"""
import numpy as np
import torch


def _handoff(device):
    """The solver enqueues on its own (non-blocking) HIP stream: finish torch's
    work on buffers before the solver reads them."""
    torch.cuda.synchronize(device)


class _Steps(torch.autograd.Function):
    @staticmethod
    def forward(ctx, U0, solver, K, m):
        n = solver.n
        U0c = U0.detach().contiguous()
        masks = torch.empty((max(K, 1), 3, n, n), dtype=torch.int8, device=U0.device)
        ckpts = [U0c.clone()]
        dts = []
        _handoff(U0.device)
        solver.set_state_device(U0c.data_ptr())
        solver.set_time(0.0)
        solver.record_masks(masks.data_ptr(), K)
        done = 0
        while done < K:
            k = min(m, K - done)
            dts.append(solver.run(k, return_dts=True))
            done += k
            if done < K:
                c = torch.empty_like(U0c)
                solver.get_state_device(c.data_ptr())
                ckpts.append(c)
        solver.record_masks(0, 0)
        UK = torch.empty_like(U0c)
        solver.get_state_device(UK.data_ptr())
        torch.cuda.synchronize(U0.device)
        ctx.solver, ctx.K, ctx.m = solver, K, m
        ctx.masks, ctx.ckpts = masks, ckpts
        ctx.dts = np.concatenate(dts) if dts else np.zeros(0)
        return UK

    @staticmethod
    def backward(ctx, gK):
        solver, K, m = ctx.solver, ctx.K, ctx.m
        n = solver.n
        ub = gK.detach().to(torch.float64).contiguous().clone()
        stages = torch.empty((min(m, max(K, 1)), 3, n, n, 4), dtype=torch.float64, device=ub.device)
        nseg = (K + m - 1) // m
        _handoff(ub.device)
        for seg in range(nseg - 1, -1, -1):
            s0 = seg * m
            ns = min(m, K - s0)
            solver.set_state_device(ctx.ckpts[seg].data_ptr())
            seg_masks = ctx.masks[s0:s0 + ns]
            solver.frozen_run(seg_masks.data_ptr(), ctx.dts[s0:s0 + ns], stages.data_ptr())
            solver.adjoint(stages.data_ptr(), seg_masks.data_ptr(), ctx.dts[s0:s0 + ns], ub.data_ptr())
        return ub, None, None, None


def steps(solver, U0, K, checkpoint_every=20):
    """U_K as a differentiable function of U0 (masks and dts frozen at U0's run)."""
    if K < 0 or checkpoint_every < 1:
        raise ValueError("K >= 0 and checkpoint_every >= 1 required")
    if U0.dtype != torch.float64 or U0.shape != (solver.n, solver.n, 4) or not U0.is_cuda:
        raise ValueError("U0 must be a float64 CUDA tensor of shape (n, n, 4)")
    return _Steps.apply(U0, solver, int(K), int(checkpoint_every))


def record(solver, U0, K):
    """Run K steps of the plain solver from U0, returning (U_K, masks, dts)
    with masks int8 [K][3][n][n] (device) and dts float64 [K] (host)."""
    n = solver.n
    masks = torch.empty((max(K, 1), 3, n, n), dtype=torch.int8, device=U0.device)
    U0c = U0.detach().contiguous()
    _handoff(U0.device)
    solver.set_state_device(U0c.data_ptr())
    solver.set_time(0.0)
    solver.record_masks(masks.data_ptr(), K)
    dts = solver.run(K, return_dts=True)
    solver.record_masks(0, 0)
    UK = torch.empty_like(U0)
    solver.get_state_device(UK.data_ptr())
    torch.cuda.synchronize(U0.device)
    return UK, masks[:K], dts


def frozen(solver, U0, masks, dts):
    """U_K of the frozen-mask, frozen-dt map (no autograd)."""
    U0c = U0.detach().contiguous()
    _handoff(U0.device)
    solver.set_state_device(U0c.data_ptr())
    solver.frozen_run(masks.data_ptr(), dts)
    UK = torch.empty_like(U0)
    solver.get_state_device(UK.data_ptr())
    torch.cuda.synchronize(U0.device)
    return UK


def kinetic_energy(U, i0, j0, w):
    """J = sum over cells [j0, j0+w) x [i0, i0+w) of rho (u^2 + v^2)."""
    S = U[j0:j0 + w, i0:i0 + w]
    return ((S[..., 1] * S[..., 1] + S[..., 2] * S[..., 2]) / S[..., 0]).sum()


def hllc_vjp(ul, ur, fbar, direction):
    """Reverse mode of the HLLC flux on [n, 4] float64 CUDA tensors."""
    from . import capi
    lib = capi.load()
    ul, ur, fbar = (t.detach().contiguous() for t in (ul, ur, fbar))
    ulb, urb = torch.empty_like(ul), torch.empty_like(ur)
    _handoff(ul.device)
    rc = lib.gpm_hllc_vjp(int(direction), int(ul.shape[0]), ul.data_ptr(), ur.data_ptr(), fbar.data_ptr(),
                          ulb.data_ptr(), urb.data_ptr())
    if rc != 0:
        raise RuntimeError(f"gpm_hllc_vjp failed ({rc})")
    return ulb, urb
