"""The reference's own network file (neural_network/save_NN_for_fortran.py,
read by load_NN, mod_NN.f90:60-117) from Python.

A network is a dict of float32 matrices with row i = output i, as PyTorch's
nn.Linear holds them (neural_network/NN.py:31-43): weight0 (length, L),
bias0 (length,), weight1 (length, length), bias1 (length,), weight2
(2, length), bias2 (2,).  ``Solver.set_reference_nn(**net)`` runs it."""
import numpy as np

from . import capi

KEYS = ("weight0", "bias0", "weight1", "bias1", "weight2", "bias2")


def read_text(nn_filename, L, length):
    """load_NN: "<nn_filename>.txt", one number per record, each matrix row by
    row; L and length as the parameters module sets them."""
    shapes = ((length, L), (length,), (length, length), (length,), (2, length), (2,))
    bufs = [np.zeros(int(np.prod(s)), np.float32) for s in shapes]
    rc = capi.load().gpm_read_reference_nn(str(nn_filename).encode(), int(L), int(length),
                                           *[capi.ptr(b) for b in bufs])
    if rc != 0:
        raise ValueError(f"cannot read {nn_filename}.txt as a network of L={L}, length={length}")
    # the reader fills the Fortran arrays (element (i, j) at i + j*rows)
    return {k: b.reshape(tuple(reversed(s))).T.reshape(s).copy() if len(s) == 2 else b
            for k, b, s in zip(KEYS, bufs, shapes)}


def write_text(path, net):
    """The text of save_NN_for_fortran.py:12-25: every parameter in PyTorch
    order, flattened row-major, one number per line."""
    with open(path, "w") as f:
        for k in KEYS:
            for v in np.asarray(net[k], dtype=np.float32).ravel():
                f.write(repr(float(v)) + "\n")


def seeded_network(L=57, length=60, seed=0, stencil=13, lowered_quantile=0.75, U=None):
    """A seeded, untrained network of the reference's shapes, for throughput
    runs: weights ~ N(0, 2/sqrt(fan-in)) (the -666 inputs past 4*stencil get
    1e-3), and, given an interior state U (nf, lf, 4), bias2 set so that the
    cells above `lowered_quantile` of r(1) - r(2) predict order 1."""
    rng = np.random.default_rng(seed)
    w0 = rng.normal(0.0, 2.0 / np.sqrt(4 * stencil), (length, L))
    w0[:, 4 * stencil:] = rng.normal(0.0, 1e-3, (length, L - 4 * stencil))
    net = dict(weight0=w0, bias0=rng.normal(0.0, 0.5, length),
               weight1=rng.normal(0.0, 2.0 / np.sqrt(length), (length, length)),
               bias1=rng.normal(0.0, 0.5, length), weight2=rng.normal(0.0, 2.0 / np.sqrt(length), (2, length)),
               bias2=np.zeros(2))
    if U is not None:
        step = max(1, int(np.sqrt(U.shape[0] * U.shape[1] / 65536)))
        U = np.ascontiguousarray(U[::step, ::step])
        x = np.full((U.shape[0] * U.shape[1], L), -666.0)
        x[:, :4 * stencil] = np.tile(U.reshape(-1, 4), (1, stencil))  # the centre cell: enough for a threshold
        with np.errstate(over="ignore"):
            sig = lambda z: 1.0 / (1.0 + np.exp(-z))  # noqa: E731
            r = sig(sig(x @ net["weight0"].T + net["bias0"]) @ net["weight1"].T + net["bias1"]) @ net["weight2"].T
        net["bias2"] = np.array([-np.quantile(r[:, 0] - r[:, 1], lowered_quantile), 0.0])
    return {k: np.asarray(v, np.float32) for k, v in net.items()}


def collect_dataset(solver, nsteps, capacity, L=52, seed=0, tmax=1e10):
    """Runs `nsteps` steps of `solver` and collects the reference trainer's
    samples after each step's first stage (append_to_NN_datatset): returns
    (inputs (size, L), labels (size, 2), cells (size,), NR0, NR1), float32
    inputs and labels as gpm.h5io.write_nn_dataset writes them."""
    import torch
    dev = "cuda"
    inputs = torch.zeros((capacity, L), dtype=torch.float32, device=dev)
    labels = torch.zeros((capacity, 2), dtype=torch.float32, device=dev)
    cells = torch.zeros(capacity, dtype=torch.int64, device=dev)
    count = torch.zeros(1, dtype=torch.int64, device=dev)
    nr = torch.zeros(2, dtype=torch.int64, device=dev)
    torch.cuda.synchronize()
    solver.set_reference_dump(inputs.data_ptr(), labels.data_ptr(), cells.data_ptr(), capacity, count.data_ptr(),
                              nr.data_ptr(), seed, L)
    try:
        solver.run(nsteps, tmax=tmax)
    finally:
        solver.set_reference_dump(0, 0, 0, 0, 0, 0)
    torch.cuda.synchronize()
    n = min(int(count.item()), capacity)
    n0, n1 = (int(v) for v in nr.cpu())
    return (inputs[:n].cpu().numpy(), labels[:n].cpu().numpy(), cells[:n].cpu().numpy(), n0, n1)
