"""CPU restatement of the a priori predictor's arithmetic -- TEST INFRASTRUCTURE.

The network (in -> 128 -> 128 -> 64 -> C, ReLU) in numpy with BF16 rounding
(round to nearest even, from FP32) at exactly the points where the GPU kernels
round: inputs and weights, each hidden activation, the logit gradient and each
back-propagated activation gradient; logits and gradients are FP32.  Sums are
taken in FP64 and rounded at those points, i.e. this is the correctly rounded
value of each FP32 quantity; the matrix cores accumulate in FP32 in another
order, so the GPU agrees to FP32 accumulation error plus rare one-ulp BF16
flips.  AdamW follows torch.optim.AdamW's update formula.

Pinned (tests/test_mlp_oracle.py) against an independent PyTorch FP64
autograd model that rounds at the same points: per-tensor relative error of
the logits, loss and every weight and bias gradient, and a planted mistake
(rounding removed here) that must fail it.

The reference's own network (neural_network/NN.py:3-67, a 52 -> L -> L -> 2
sigmoid MLP trained with Adam) is superseded by this architecture; there is no
reference arithmetic to restate, so the check is against this definition and
against PyTorch autograd (tests).
"""
import numpy as np

HIDDEN = (128, 128, 64)


def bf16(x):
    x = np.ascontiguousarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    r = np.where(nan, 0x7FC00000, r).astype(np.uint32)
    return r.view(np.float32)


def unpack(params, in_f, C):
    dims = (in_f,) + HIDDEN + (C,)
    out, off = [], 0
    for a, b in zip(dims[:-1], dims[1:]):
        W = params[off:off + a * b].reshape(b, a)
        off += a * b
        bb = params[off:off + b]
        off += b
        out.append((W.astype(np.float32), bb.astype(np.float32)))
    return out


def f32(x):
    return np.asarray(x, dtype=np.float64).astype(np.float32)


def forward(params, in_f, C, x):
    """FP32 logits [B][C] and the BF16 activations (input and 3 hidden)."""
    layers = unpack(params, in_f, C)
    h = bf16(x)
    acts = [h]
    for li, (W, b) in enumerate(layers):
        z = h.astype(np.float64) @ bf16(W).astype(np.float64).T + b
        if li < 3:
            h = bf16(np.maximum(z, 0.0))
            acts.append(h)
        else:
            return f32(z), acts


def forward_backward(params, in_f, C, x, y):
    """Mean cross-entropy loss, accuracy and the flat FP32 gradient."""
    layers = unpack(params, in_f, C)
    z, acts = forward(params, in_f, C, x)
    B = x.shape[0]
    z = z.astype(np.float64)
    zmax = z.max(axis=1, keepdims=True)
    e = np.exp(z - zmax)
    p = e / e.sum(axis=1, keepdims=True)
    loss = float(np.mean(np.log(e.sum(axis=1)) + zmax[:, 0] - z[np.arange(B), y]))
    acc = float(np.mean(np.argmax(z, axis=1) == y))
    onehot = np.zeros_like(z)
    onehot[np.arange(B), y] = 1
    d = bf16((p - onehot) / B)
    grads = [None] * 4
    for li in range(3, -1, -1):
        W, b = layers[li]
        d64 = d.astype(np.float64)
        grads[li] = (f32(d64.T @ acts[li].astype(np.float64)), f32(d64.sum(axis=0)))
        if li > 0:
            d = bf16((d64 @ bf16(W).astype(np.float64)) * (acts[li] > 0))
    flat = np.concatenate([np.concatenate([gw.ravel(), gb]) for gw, gb in grads]).astype(np.float32)
    return loss, acc, flat


def adamw(p, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8, wd=1e-2):
    p = p * (1 - lr * wd)
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    bc1 = 1 - b1 ** step
    bc2 = 1 - b2 ** step
    p = p - (lr / bc1) * m / (np.sqrt(v) / np.sqrt(bc2) + eps)
    return p.astype(np.float32), m, v
