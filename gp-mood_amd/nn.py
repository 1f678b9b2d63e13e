"""The learned a priori order predictor: an MLP  in -> 128 -> 128 -> 64 -> C
(ReLU; BF16 weights and activations, FP32 accumulation) whose forward,
backward and AdamW kernels run on the GPU's matrix cores (libgpmood).

Parameters are exchanged as one float32 vector in PyTorch Linear order
(W0 [128][in], b0, W1 [128][128], b1, W2 [64][128], b2, W3 [C][64], b3).
Training batches are device tensors (torch is plumbing for device memory).
"""
import ctypes

import numpy as np

from . import capi

HIDDEN = (128, 128, 64)


def param_shapes(in_features, n_classes):
    dims = (in_features,) + HIDDEN + (n_classes,)
    shapes = []
    for a, b in zip(dims[:-1], dims[1:]):
        shapes += [(b, a), (b,)]
    return shapes


def default_init(in_features, n_classes, seed=0):
    """PyTorch's nn.Linear default initialisation (uniform in +-1/sqrt(fan_in)
    for weights and biases), from a seeded numpy generator."""
    rng = np.random.default_rng(seed)
    out = []
    for shape in param_shapes(in_features, n_classes):
        fan_in = shape[1] if len(shape) == 2 else None
        if fan_in is None:
            fan_in = prev_fan
        bound = 1.0 / np.sqrt(fan_in)
        out.append(rng.uniform(-bound, bound, size=shape).astype(np.float32).ravel())
        if len(shape) == 2:
            prev_fan = shape[1]
    return np.concatenate(out)


class MLP:
    """Device-resident MLP (gpm_mlp).  in_features is 4*S*S for an S x S
    stencil (100 for the R2 solver, 196 for R3); n_classes = (Mord+1)/2."""

    def __init__(self, in_features, n_classes, max_batch=8192, params=None, seed=0):
        self.lib = capi.load()
        self.in_features = in_features
        self.n_classes = n_classes
        self.max_batch = max_batch
        h = ctypes.c_void_p()
        rc = self.lib.gpm_mlp_create(in_features, n_classes, max_batch, ctypes.byref(h))
        if rc != 0:
            raise RuntimeError(f"gpm_mlp_create failed ({rc})")
        self._h = h
        self.set_params(default_init(in_features, n_classes, seed) if params is None else params)

    def close(self):
        if getattr(self, "_h", None):
            self.lib.gpm_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def _check(self, rc, what):
        if rc != 0:
            msg = self.lib.gpm_mlp_last_error(self._h)
            raise RuntimeError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

    @property
    def num_params(self):
        return int(self.lib.gpm_mlp_num_params(self._h))

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.num_params
        self._check(self.lib.gpm_mlp_set_params(self._h, capi.ptr(p)), "gpm_mlp_set_params")

    def get_params(self):
        p = np.zeros(self.num_params, dtype=np.float32)
        self._check(self.lib.gpm_mlp_get_params(self._h, capi.ptr(p)), "gpm_mlp_get_params")
        return p

    def get_grads(self):
        g = np.zeros(self.num_params, dtype=np.float32)
        self._check(self.lib.gpm_mlp_get_grads(self._h, capi.ptr(g)), "gpm_mlp_get_grads")
        return g

    def forward_backward(self, x, y):
        """x: float32 device tensor [B][in]; y: int32 device tensor [B]."""
        loss, acc = ctypes.c_double(), ctypes.c_double()
        self._check(self.lib.gpm_mlp_forward_backward(self._h, ctypes.c_void_p(x.data_ptr()),
                                                      ctypes.c_void_p(y.data_ptr()), int(x.shape[0]),
                                                      ctypes.byref(loss), ctypes.byref(acc)),
                    "gpm_mlp_forward_backward")
        return loss.value, acc.value

    def adamw_step(self, lr=1e-3, betas=(0.9, 0.999), eps=1e-8, weight_decay=1e-2):
        self._check(self.lib.gpm_mlp_adamw_step(self._h, lr, betas[0], betas[1], eps, weight_decay),
                    "gpm_mlp_adamw_step")

    def train_step(self, x, y, lr=1e-3, weight_decay=1e-2, report=True):
        """One AdamW step on a batch.  With report=False the loss and accuracy
        are not read back, so the call only enqueues work (returns None, None)."""
        if not report:
            self._check(self.lib.gpm_mlp_train_step(self._h, ctypes.c_void_p(x.data_ptr()),
                                                    ctypes.c_void_p(y.data_ptr()), int(x.shape[0]), lr,
                                                    weight_decay, None, None), "gpm_mlp_train_step")
            return None, None
        loss, acc = ctypes.c_double(), ctypes.c_double()
        self._check(self.lib.gpm_mlp_train_step(self._h, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(y.data_ptr()),
                                                int(x.shape[0]), lr, weight_decay, ctypes.byref(loss),
                                                ctypes.byref(acc)), "gpm_mlp_train_step")
        return loss.value, acc.value

    def logits(self, x):
        import torch
        out = torch.empty((x.shape[0], self.n_classes), dtype=torch.float32, device=x.device)
        self._check(self.lib.gpm_mlp_logits(self._h, ctypes.c_void_p(x.data_ptr()), int(x.shape[0]),
                                            ctypes.c_void_p(out.data_ptr())), "gpm_mlp_logits")
        return out

    def save(self, path):
        np.savez(path, params=self.get_params(), in_features=self.in_features, n_classes=self.n_classes)

    @classmethod
    def load(cls, path, max_batch=8192):
        z = np.load(path, allow_pickle=False)
        return cls(int(z["in_features"]), int(z["n_classes"]), max_batch=max_batch, params=z["params"])
