"""The predictor network as a PyTorch FP64 autograd model that rounds to BF16
(from FP32, round to nearest even) at the points where the GPU kernels round:
inputs and weights, each hidden activation in the forward pass; the logit
gradient and each back-propagated pre-activation gradient in the backward
pass.  Rounding in the forward pass is transparent to the gradient (the
kernels differentiate through it as the identity).  The backward pass is
torch's own autograd, so this derives every gradient independently of the
hand-written backward it checks."""
import numpy as np
import torch

HIDDEN = (128, 128, 64)


def _bf16(t):
    return t.to(torch.float32).to(torch.bfloat16).to(torch.float64)


class _RoundFwd(torch.autograd.Function):
    """Value rounded to BF16 (or FP32 when fp32=True), gradient passed as is."""

    @staticmethod
    def forward(ctx, x, fp32):
        return x.to(torch.float32).to(torch.float64) if fp32 else _bf16(x)

    @staticmethod
    def backward(ctx, g):
        return g, None


class _RoundGrad(torch.autograd.Function):
    """Identity whose incoming gradient is rounded to BF16."""

    @staticmethod
    def forward(ctx, x):
        return x.view_as(x)

    @staticmethod
    def backward(ctx, g):
        return _bf16(g)


def param_tensors(params, in_f, C):
    dims = (in_f,) + HIDDEN + (C,)
    out, off = [], 0
    for a, b in zip(dims[:-1], dims[1:]):
        W = torch.tensor(params[off:off + a * b].reshape(b, a), dtype=torch.float64, requires_grad=True)
        off += a * b
        bb = torch.tensor(params[off:off + b], dtype=torch.float64, requires_grad=True)
        off += b
        out.append((W, bb))
    return out


def logits(layers, x):
    h = _RoundFwd.apply(torch.as_tensor(x, dtype=torch.float64), False)
    for li, (W, b) in enumerate(layers):
        z = torch.nn.functional.linear(h, _RoundFwd.apply(W, False), b)
        if li < 3:
            h = _RoundFwd.apply(torch.relu(_RoundGrad.apply(z)), False)
        else:
            return _RoundFwd.apply(z, True)


def forward_backward(params, in_f, C, x, y):
    """(FP32 logits, loss, per-layer [(dW, db)] as float64 arrays)."""
    layers = param_tensors(params, in_f, C)
    z = logits(layers, x)
    loss = torch.nn.functional.cross_entropy(_RoundGrad.apply(z), torch.as_tensor(y, dtype=torch.long))
    loss.backward()
    grads = [(W.grad.numpy(), b.grad.numpy()) for W, b in layers]
    return z.detach().numpy(), float(loss.detach()), grads


def split_flat(flat, in_f, C):
    """Flat parameter-order vector -> [(W, b)] per layer."""
    dims = (in_f,) + HIDDEN + (C,)
    out, off = [], 0
    for a, b in zip(dims[:-1], dims[1:]):
        W = np.asarray(flat[off:off + a * b]).reshape(b, a)
        off += a * b
        out.append((W, np.asarray(flat[off:off + b])))
        off += b
    return out
