"""The reference's own network (NN_GP_MOOD, NN_GP_MOOD_CC with the text-format
network of neural_network/save_NN_for_fortran.py) on the GPU, through the
C-ABI: bit for bit against runs of the reference build, and the network's FP32
exp against the host C library on every float of its range."""
import os

import numpy as np
import pytest

from conftest import GOLDEN, case_options, has_gpu, tables_from_npz
from test_refnn import REFNN_CASES, WEIGHTS, refnn_oracle, write_text

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not has_gpu():
        pytest.skip("no GPU")


def make_solver(gpm, z):
    o = case_options(z)
    method = gpm.capi.NN_GP_MOOD_CC if o["method"] == "nn_gp_mood_cc" else gpm.capi.NN_GP_MOOD
    cfg = gpm.Config(lf=int(z["n"]), nf=o["nf"], Mord=int(z["mord"]), ngp=int(z["ngp"]), BC_type=str(z["bc"]),
                     DMP=bool(z["dmp"]), U2=bool(z["u2"]), U2_tol=bool(z["u2tol"]), numFlux=o["flux"],
                     integrator=o["integrator"], Lx=o["lx"], Ly=o["ly"], method=method)
    return gpm.Solver(cfg, tables_from_npz(z))


def run_steps(s, nsteps):
    """Step by step: per-step stats, CellGPO and CellGPO_priori of the interior."""
    out = dict(dt=[], counts=[], apriori=[], nan=[], masks=[], priori=[])
    for _ in range(nsteps):
        out["dt"].append(s.run(1, return_dts=True)[0])
        st = s.stats()
        out["counts"].append(st["count_detected_cell"])
        out["apriori"].append(st["count_detected_cell_a_priori"])
        out["nan"].append(st["nn_nan_passes"])
        _, g = s.get_state(order=True)
        out["masks"].append(g[4:-4, 4:-4].astype(np.int8))
        out["priori"].append(s.cellgpo_priori()[4:-4, 4:-4].astype(np.int8))
    return {k: np.array(v) for k, v in out.items()}


@pytest.mark.parametrize("case", REFNN_CASES)
def test_reference_network_matches_reference_run(gpm, case):
    """Every step's dt, count_detected_cell, a priori count,
    count_steps_NN_produced_NAN increments, CellGPO and CellGPO_priori, and the
    final state, bit for bit against the reference build's run."""
    z = np.load(os.path.join(GOLDEN, case))
    nsteps = int(z["nsteps"])
    print(f"[case {case}] {nsteps} steps, network L={int(z['L'])} length={int(z['length'])}", flush=True)
    s = make_solver(gpm, z)
    s.set_reference_nn(*(z[k] for k in WEIGHTS))
    s.set_state(z["U0"])
    s.set_time(0.0)
    r = run_steps(s, nsteps)
    np.testing.assert_array_equal(r["dt"], z["dt"])
    np.testing.assert_array_equal(r["apriori"], z["apriori"])
    np.testing.assert_array_equal(r["counts"], z["counts"])
    np.testing.assert_array_equal(np.cumsum(r["nan"]), z["nan_cum"])
    np.testing.assert_array_equal(r["masks"], z["masks"])
    np.testing.assert_array_equal(r["priori"], z["priori"])
    U = s.get_state()
    assert np.array_equal(U.view(np.uint64), z["U"].view(np.uint64)), f"max |diff| {np.abs(U - z['U']).max()}"
    st = s.stats()
    assert st["niter"] == nsteps and st["nn_nan_passes_total"] == int(z["nan_cum"][-1])


def test_load_reference_nn_text_and_niter(gpm, orc, tmp_path):
    """gpm_load_reference_nn on the text of save_NN_for_fortran.py gives the
    network of gpm_set_reference_nn; with niter set past nsteps_with_no_NN the
    network predicts from the first step, as the oracle does."""
    case = [c for c in REFNN_CASES if c.startswith("refnn_nn_r1_n64_neu")][0]
    z = np.load(os.path.join(GOLDEN, case))
    write_text(str(tmp_path / "refnn.txt"), z)
    print(f"[case {case}] text weights, niter 40, 6 steps", flush=True)
    s = make_solver(gpm, z)
    s.load_reference_nn(str(tmp_path / "refnn"), int(z["L"]), int(z["length"]))
    s.set_niter(40)
    s.set_state(z["U0"])
    s.set_time(0.0)
    r = run_steps(s, 6)
    oc = refnn_oracle(orc, z)
    oc.set_niter(40)
    Uo = oc.full(z["U0"])
    ro = oc.run_nn(Uo, 6)
    assert r["apriori"][0] > 0
    np.testing.assert_array_equal(r["dt"], ro["dt"])
    np.testing.assert_array_equal(r["apriori"], ro["apriori"])
    np.testing.assert_array_equal(r["masks"], ro["masks"])
    np.testing.assert_array_equal(r["priori"], ro["priori"])
    assert np.array_equal(s.get_state().view(np.uint64), Uo.view(np.uint64))


def test_network_expf_every_float(gpm, orc):
    """The network kernel's FP32 exp on the device against the host C
    library's expf, for every float in [-0x1.9fe368p6, 0x1.62e42ep6] (the
    range where neither overflow nor underflow applies) and the special
    values."""
    import torch
    z = np.load(os.path.join(GOLDEN, REFNN_CASES[0]))
    s = make_solver(gpm, z)
    chunk = 1 << 27
    buf = torch.empty(chunk, dtype=torch.float32, device="cuda")
    total = 0
    for first, last in ((0x00000000, 0x42B17217), (0x80000000, 0xC2CFF1B4)):
        for b in range(first, last + 1, chunk):
            n = min(chunk, last - b + 1)
            torch.cuda.synchronize()
            s.nn_expf_bits(b, n, buf.data_ptr())
            mine = buf[:n].cpu().numpy()
            ref = orc.expf_bits(b, n)
            bad = np.nonzero(mine.view(np.uint32) != ref.view(np.uint32))[0]
            assert bad.size == 0, f"{bad.size} differ from {hex(b)}, first at bits {hex(b + int(bad[0]))}"
            total += n
    print(f"[expf] {total} floats equal", flush=True)
    special = np.array([np.inf, -np.inf, np.nan, 88.8, 200.0, -104.0, -1e4, 0.0, -0.0, 1e-45],
                       np.float32).view(np.uint32)
    for u in special:
        s.nn_expf_bits(int(u), 1, buf.data_ptr())
        mine, ref = buf[:1].cpu().numpy(), orc.expf_bits(int(u), 1)
        assert (np.isnan(mine[0]) and np.isnan(ref[0])) or mine.view(np.uint32)[0] == ref.view(np.uint32)[0]


def test_reference_dataset_samples(gpm, orc, tmp_path):
    """The reference trainer's samples after the first RK stage
    (append_to_NN_datatset): each kept row holds format_input of its cell's
    R+1 stencil of the stage input, bit for bit as the oracle forms it, and the
    stage's final order as the label; every order-1 cell with a non-constant
    stencil is kept; NR0 and NR1 count the kept labels; the rows write as the
    reference's dataset file."""
    from conftest import golden
    z = golden("step_r1_n32_neu_fq5.npz")
    n, L = int(z["n"]), 57
    print(f"[case reference dataset: {n}^2 R1, one step]", flush=True)
    cfg = gpm.Config(lf=n, Mord=3, ngp=2, BC_type=str(z["bc"]))
    s = gpm.Solver(cfg, tables_from_npz(z))
    s.set_state(z["U0"])
    s.set_time(0.0)
    X, Y, cells, n0, n1 = gpm.refnn.collect_dataset(s, 1, n * n, L=L, seed=7)
    o = case_options(z)
    dx, dy = (float(v) for v in z["dxdy"])
    oc = orc.Oracle(n, 3, 2, str(z["bc"]), tables_from_npz(z), flux=o["flux"], dx=dx, dy=dy)
    U = oc.full(z["U0"])
    _, masks, _ = oc.step_stages(U.copy(), float(z["dt"][0]), 0.0)
    m0 = masks[0]
    assert np.array_equal(m0, z["stage_masks"][0])
    assert len(cells) == n0 + n1 and len(set(cells.tolist())) == len(cells)
    for row, cell in enumerate(cells):
        j, i = divmod(int(cell), n)
        x, cst = oc.nn_format_input(U, i + 1, j + 1, L)
        assert not cst
        assert np.array_equal(X[row].view(np.uint32), x.view(np.uint32)), f"cell {cell}"
        exp = (0.0, 1.0) if m0[j, i] == 3 else (1.0, 0.0)
        assert tuple(Y[row]) == exp
    # every order-1 cell whose stencil is not constant is kept
    low = {j * n + i for j, i in zip(*np.nonzero(m0 == 1)) if not oc.nn_format_input(U, i + 1, j + 1, L)[1]}
    assert low and low <= set(cells.tolist())
    assert n0 == len(low) and n1 == int((Y[:, 1] == 1.0).sum())
    try:
        gpm.h5io.load()
    except RuntimeError:
        return
    m = gpm.h5io.meta(n, n, "2DRP3", gpm.capi.GP_MOOD)
    gpm.h5io.write_nn_dataset(str(tmp_path / "ds.h5"), m, X, Y, nr0=n0, nr1=n1)
    assert os.path.getsize(str(tmp_path / "ds.h5")) > X.nbytes


def test_reference_dataset_arrays_wrap_as_the_reference(gpm):
    """gpm_set_reference_dataset / gpm_append_reference_dataset with a small
    dataset_size: the arrays act as the reference's ring (index wraps to 1,
    n_overwrite counts the wraps, mod_append_NN_dataset.f90:64-71), and hold
    the last dataset_size rows of an unwrapped run with the same draws."""
    import ctypes
    if not has_gpu():
        pytest.skip("no GPU")
    n, L, steps, small = 64, 52, 6, 100
    print(f"[case reference dataset ring, {n}^2 R1, {steps} steps, dataset_size {small}]", flush=True)
    U0 = gpm.ic.make("four_quadrant", n, 76)
    lib = gpm.capi.load()
    runs = {}
    for size in (100000, small):
        s = gpm.Solver(gpm.Config(lf=n, Mord=3, ngp=2, BC_type="neumann"))
        U = s.full_shape()
        Uf = np.zeros(U)
        Uf[4:-4, 4:-4] = U0
        gpm.boundary_c(s.config, Uf, 0.0)
        s.set_state(Uf)
        inputs = np.full((L, size), -666.0, dtype=np.float32)   # Fortran inputs(size, L), column-major
        labels = np.full((2, size), -666.0, dtype=np.float32)
        cnt = {k: ctypes.c_int32(v) for k, v in (("index", 1), ("n_over", 0), ("nr0", 0), ("nr1", 0))}
        assert lib.gpm_set_reference_dataset(s._h, L, 0, 0) == 0
        t = 0.0
        for _ in range(steps):
            dt = s.setdt()
            Uf, _, _ = s.time_stepping(Uf, t, dt)
            t += dt
            assert lib.gpm_append_reference_dataset(s._h, gpm.capi.ptr(inputs), gpm.capi.ptr(labels), size,
                                                    ctypes.byref(cnt["index"]), ctypes.byref(cnt["n_over"]),
                                                    ctypes.byref(cnt["nr0"]), ctypes.byref(cnt["nr1"])) == 0
        runs[size] = (inputs.T.copy(), labels.T.copy(), {k: v.value for k, v in cnt.items()})
        s.close()
    big_in, big_lab, cb = runs[100000]
    ring_in, ring_lab, cr = runs[small]
    total = cb["nr0"] + cb["nr1"]
    assert total > 3 * small and cb["n_over"] == 0 and cb["index"] == total + 1
    assert (cr["nr0"], cr["nr1"]) == (cb["nr0"], cb["nr1"])
    assert cr["n_over"] == total // small and cr["index"] == total % small + 1
    # row r goes to slot r mod size: slot k holds the last such row before `total`
    for k in range(small):
        r = k + small * ((total - 1 - k) // small)
        assert np.array_equal(ring_in[k], big_in[r]) and np.array_equal(ring_lab[k], big_lab[r])
