"""The reference's own network (methods NN_GP_MOOD and NN_GP_MOOD_CC, mod_NN.f90)
on the CPU: the oracle against runs of the reference build, the library's
text-weight reader and its FP32 exp against the host C library."""
import ctypes
import glob
import os

import numpy as np
import pytest

from conftest import GOLDEN, case_options, load_package, tables_from_npz

REFNN_CASES = sorted(os.path.basename(f) for f in glob.glob(os.path.join(GOLDEN, "refnn_*.npz")))
WEIGHTS = ("weight0", "bias0", "weight1", "bias1", "weight2", "bias2")


def refnn_oracle(orc, z):
    o = case_options(z)
    dx, dy = (float(v) for v in z["dxdy"])
    oc = orc.Oracle(int(z["n"]), int(z["mord"]), int(z["ngp"]), str(z["bc"]), tables_from_npz(z), dmp=bool(z["dmp"]),
                    u2=bool(z["u2"]), u2tol=bool(z["u2tol"]), flux=o["flux"], nf=o["nf"], integrator=o["integrator"],
                    dx=dx, dy=dy, method=o["method"])
    oc.set_reference_nn(*(z[k] for k in WEIGHTS))
    return oc


def write_text(path, z):
    """save_NN_for_fortran.py:12-25: parameters in PyTorch order, each
    flattened row-major, one number per line."""
    with open(path, "w") as f:
        for k in WEIGHTS:
            for v in np.asarray(z[k], dtype=np.float32).ravel():
                f.write(repr(float(v)) + "\n")


def test_cases_present():
    assert len(REFNN_CASES) >= 5
    methods = {str(np.load(os.path.join(GOLDEN, c))["method"]) for c in REFNN_CASES}
    assert methods == {"nn_gp_mood", "nn_gp_mood_cc"}


@pytest.mark.parametrize("case", REFNN_CASES)
def test_oracle_matches_reference_run(orc, case):
    """Every step's dt, count_detected_cell, a priori and a posteriori counts,
    count_steps_NN_produced_NAN, CellGPO and CellGPO_priori, and the final
    state (ghost cells included), bit for bit against the reference build
    (steps 1-10 are plain GP-MOOD, the network predicts from step 11)."""
    z = np.load(os.path.join(GOLDEN, case))
    oc = refnn_oracle(orc, z)
    U = oc.full(z["U0"])
    r = oc.run_nn(U, int(z["nsteps"]))
    np.testing.assert_array_equal(r["dt"], z["dt"])
    np.testing.assert_array_equal(r["counts"], z["counts"])
    np.testing.assert_array_equal(r["apriori"], z["apriori"])
    np.testing.assert_array_equal(r["counts"] - r["apriori"], z["aposteriori"])
    np.testing.assert_array_equal(np.cumsum(r["nan_passes"]), z["nan_cum"])
    np.testing.assert_array_equal(r["masks"], z["masks"])
    np.testing.assert_array_equal(r["priori"], z["priori"])
    assert np.array_equal(U.view(np.uint64), z["U"].view(np.uint64))
    # the network acted: a priori detections only after nsteps_with_no_NN = 10
    assert z["apriori"][:10].sum() == 0 and z["apriori"][10:].sum() > 0


def test_cc_fixture_pins_the_blend(orc):
    """The NN_GP_MOOD_CC run is not the NN_GP_MOOD run of the same network:
    the softmax, the blended order-3 values of column L (values_NN(L, n, :),
    mod_NN.f90:41,44) and the raw order-1 values change the state, so the
    fixture pins them and not only the orders."""
    case = [c for c in REFNN_CASES if c.startswith("refnn_nncc")][0]
    z = np.load(os.path.join(GOLDEN, case))
    assert int(z["L"]) <= int(z["n"]) + 2  # values_NN(L, n) lies in row n
    oc = refnn_oracle(orc, z)
    orc.lib().orc_set_method(oc._h, orc.METHODS["nn_gp_mood"])
    U = oc.full(z["U0"])
    oc.run_nn(U, int(z["nsteps"]))
    assert not np.array_equal(U, z["U"])


def test_read_reference_nn_text(gpm, tmp_path):
    """gpm_read_reference_nn reads the text of save_NN_for_fortran.py into the
    Fortran arrays load_NN fills (weight0(length, L) ... with j fastest in the
    file); blank records are skipped, records past the last weight ignored, and
    a short file is an error (the reference's read stops at end of file)."""
    lib = gpm.capi.load()
    z = np.load(os.path.join(GOLDEN, REFNN_CASES[0]))
    L, length = int(z["L"]), int(z["length"])
    stem = str(tmp_path / "net")
    write_text(stem + ".txt", z)
    with open(stem + ".txt", "a") as f:
        f.write("\n12345.0\n")  # past up5: not read
    bufs = [np.zeros(n, np.float32) for n in (length * L, length, length * length, length, 2 * length, 2)]
    assert lib.gpm_read_reference_nn(stem.encode(), L, length, *[gpm.capi.ptr(b) for b in bufs]) == 0
    for b, k in zip(bufs, WEIGHTS):
        ref = np.asarray(z[k], np.float32)
        ref = ref.reshape(ref.shape[0], -1)
        np.testing.assert_array_equal(b, ref.T.ravel())  # Fortran layout: element (i, j) at i + j*rows
    # a file with blank records between the numbers reads the same
    lines = open(stem + ".txt").read().split("\n")
    with open(stem + "_b.txt", "w") as f:
        f.write("\n\n".join(lines))
    bufs2 = [np.zeros_like(b) for b in bufs]
    assert lib.gpm_read_reference_nn((stem + "_b").encode(), L, length, *[gpm.capi.ptr(b) for b in bufs2]) == 0
    for a, b in zip(bufs, bufs2):
        np.testing.assert_array_equal(a, b)
    # too few records
    with open(stem + "_short.txt", "w") as f:
        f.write("\n".join(lines[:length * L]))
    args = [gpm.capi.ptr(b) for b in bufs2]
    assert lib.gpm_read_reference_nn((stem + "_short").encode(), L, length, *args) == -1
    assert lib.gpm_read_reference_nn(b"/nonexistent/file", L, length, *args) == -1


def test_oracle_forward_against_float32_numpy(orc):
    """forward of the oracle (mod_NN.f90:165-181) against the same network
    written with numpy float32 and the row sums taken in column order: equal
    up to the exp of the two implementations (the oracle uses the C library's
    expf, numpy its own), so within a few float32 ulps."""
    z = np.load(os.path.join(GOLDEN, REFNN_CASES[0]))
    oc = refnn_oracle(orc, z)
    rng = np.random.default_rng(3)
    L = int(z["L"])
    w0, b0, w1, b1, w2, b2 = (np.asarray(z[k], np.float32) for k in WEIGHTS)
    f32 = np.float32

    def seqmv(W, x):
        acc = np.zeros(W.shape[0], f32)
        for k in range(W.shape[1]):
            acc = (acc + (W[:, k] * x[k]).astype(f32)).astype(f32)
        return acc

    for _ in range(20):
        x = np.concatenate([rng.uniform(0.1, 3.0, 52), np.full(L - 52, -666.0)]).astype(f32)
        h0 = (f32(1) / (f32(1) + np.exp(-(seqmv(w0, x) + b0)))).astype(f32)
        h1 = (f32(1) / (f32(1) + np.exp(-(seqmv(w1, h0) + b1)))).astype(f32)
        r = seqmv(w2, h1) + b2
        np.testing.assert_allclose(oc.nn_forward(x), r, rtol=1e-5, atol=1e-5)


def test_expf_matches_host_libm(gpm, orc):
    """The library's FP32 exp (the function its network kernel runs, compiled
    here for the host) against the host C library's expf: 512 runs of 32768
    consecutive floats spread over [-0x1.9fe368p6, 0x1.62e42ep6] and the
    special values; GPM_EXHAUSTIVE=1 checks every float of the range (about a
    minute; it has passed)."""
    lib = gpm.capi.load()
    exhaustive = bool(os.environ.get("GPM_EXHAUSTIVE"))
    for first, last in ((0x00000000, 0x42B17217), (0x80000000, 0xC2CFF1B4)):
        if exhaustive:
            runs = [(b, min(1 << 26, last - b + 1)) for b in range(first, last + 1, 1 << 26)]
        else:
            runs = [(int(b), 32768) for b in np.linspace(first, last - 32768, 256)]
        for b, n in runs:
            mine = np.empty(n, np.float32)
            assert lib.gpm_nn_expf_bits_host(b, n, gpm.capi.ptr(mine)) == 0
            ref = orc.expf_bits(b, n)
            assert np.array_equal(mine.view(np.uint32), ref.view(np.uint32)), hex(b)
    special = np.array([np.inf, -np.inf, np.nan, 88.5, 88.8, 200.0, -103.9, -104.0, -1e4, 0.0, -0.0, 1e-45,
                        -1e-45, float.fromhex("0x1.62e42ep6"), -float.fromhex("0x1.9fe368p6")], np.float32)
    for x in special:
        u = int(x.view(np.uint32))
        mine = np.empty(1, np.float32)
        lib.gpm_nn_expf_bits_host(u, 1, gpm.capi.ptr(mine))
        ref = orc.expf_bits(u, 1)
        assert (np.isnan(mine[0]) and np.isnan(ref[0])) or mine.view(np.uint32)[0] == ref.view(np.uint32)[0], x
