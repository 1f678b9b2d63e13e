"""The library's quad-precision GP precompute (gp-mood_amd/csrc/gp_tables.cpp)
against the reference's GP_presim_sphere output, plus invariants that the
mathematics fixes.  Host code only: runs without a GPU."""
import glob
import os

import numpy as np
import pytest

from conftest import GOLDEN, case_options, golden, load_package

SZ = {1: 1, 3: 5, 5: 13, 7: 25}


def ulps(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    ia = a.view(np.int64).astype(np.float64)
    ib = b.view(np.int64).astype(np.float64)
    return np.abs(ia - ib)


def fixture_cases():
    """Every distinct table configuration among the reference fixtures: the
    step cases, the problem-file and long-run checksums."""
    seen = {}
    for p in sorted(glob.glob(os.path.join(GOLDEN, "*.npz"))):
        z = np.load(p, allow_pickle=False)
        if "zT" not in z.files or "mord" not in z.files:
            continue
        o = case_options(z)
        key = (int(z["mord"]), int(z["ngp"]), int(z["n"]), o["nf"], o["lx"], o["ly"])
        seen.setdefault(key, os.path.basename(p))
    return sorted(seen.items())


def bits(a):
    a = np.asarray(a)
    return a.view(np.int64) if a.dtype == np.float64 else a


def zt_view(z, mord, ngp):
    """zT_sp(Mord, sz, 4, ngp) (Fortran order) as an array [ord-1, s, face, gp]."""
    return np.asarray(z).reshape((ngp, 4, SZ[mord], mord)).transpose(3, 2, 1, 0)


@pytest.mark.parametrize("key,fname", fixture_cases())
def test_tables_match_reference(key, fname):
    """The library's own tables (gpm_create without gpm_set_gp_tables) equal
    the reference build's GP_presim_sphere output bit for bit: prediction
    vectors, second-derivative vectors, Gauss weights and stencil order."""
    mord, ngp, n, nf, lx, ly = key
    gpm = load_package()
    ref = golden(fname)
    t = gpm.gp_tables(mord, ngp, n, nf, lx=lx, ly=ly)
    assert np.array_equal(t["ixiy"], ref["ixiy"])
    assert np.array_equal(bits(t["gw"]), bits(ref["gw"]))
    zr, zm = zt_view(ref["zT"], mord, ngp), zt_view(t["zT"], mord, ngp)
    for ord_ in range(1, mord + 1, 2):
        sz = SZ[ord_]
        d = ulps(zm[ord_ - 1, :sz], zr[ord_ - 1, :sz])
        assert d.max() == 0, f"ord {ord_}: {int((d > 0).sum())} weights differ, max {d.max()} ulp"
    for key2 in ("d2x", "d2y"):
        a = np.asarray(t[key2]).reshape(SZ[mord], mord).T
        b = np.asarray(ref[key2]).reshape(SZ[mord], mord).T
        for ord_ in range(3, mord + 1, 2):
            sz = SZ[ord_]
            assert np.array_equal(bits(a[ord_ - 1, :sz]), bits(b[ord_ - 1, :sz])), f"{key2} ord {ord_}"


@pytest.mark.parametrize("mord,ngp", [(3, 2), (5, 3), (7, 4), (7, 3), (5, 2)])
def test_table_invariants(mord, ngp):
    gpm = load_package()
    t = gpm.gp_tables(mord, ngp, 128)
    z = zt_view(t["zT"], mord, ngp)
    gw = np.asarray(t["gw"]).reshape(5, 5, order="F")
    # Gauss-Legendre on [-1/2, 1/2]: weights sum to one
    assert abs(gw[ngp - 1, :ngp].sum() - 1.0) < 4e-16
    for ord_ in range(1, mord + 1, 2):
        sz = SZ[ord_]
        w = z[ord_ - 1, :sz]
        # normalisation (GP_init.f90:262): every prediction vector sums to 1
        assert np.abs(w.sum(axis=0) - 1.0).max() < 1e-14
        if ord_ == 1:
            assert np.all(w == 1.0)
        # mirror symmetry: face R at point j is face L mirrored in x
        sx = np.array([0, -1, 0, 1, 0, -1, -1, 1, 1, 0, -2, 0, 2, -3, -2, -1, 0, 1, 2, 3, 2, 1, 0, -1, -2])[:sz]
        sy = np.array([0, 0, 1, 0, -1, -1, 1, 1, -1, -2, 0, 2, 0, 0, 1, 2, 3, 2, 1, 0, -1, -2, -3, -2, -1])[:sz]
        idx = {(a, b): k for k, (a, b) in enumerate(zip(sx, sy))}
        perm = np.array([idx[(-a, b)] for a, b in zip(sx, sy)])
        L, R = 0, 2
        assert np.abs(w[:, R, :] - w[perm][:, L, :]).max() < 1e-12
        # and face T is face R rotated: (x, y) -> (y, x)
        permT = np.array([idx[(b, a)] for a, b in zip(sx, sy)])
        T = 1
        assert np.abs(w[:, T, :] - w[permT][:, R, :]).max() < 1e-12
    # second-derivative vectors (order-3 stencil (0,0),(-1,0),(0,1),(1,0),(0,-1)):
    # d2x is even in x and in y, and d2y is d2x with x and y exchanged
    d2x = np.asarray(t["d2x"]).reshape(SZ[mord], mord).T[2, :5]
    d2y = np.asarray(t["d2y"]).reshape(SZ[mord], mord).T[2, :5]
    scale = np.abs(d2x).max()
    assert abs(d2x[1] - d2x[3]) < 1e-12 * scale and abs(d2x[2] - d2x[4]) < 1e-12 * scale
    assert np.abs(d2y[[0, 1, 2, 3, 4]] - d2x[[0, 2, 1, 4, 3]]).max() < 1e-12 * scale
    # the x-curvature weight sits on the x-neighbours: d2x(-1,0) > 0 > d2x(0,0)
    assert d2x[1] > 0 > d2x[0]


def test_tables_predict_smooth_field():
    """Face predictions of a smooth field from its cell averages (x-only field
    sin(2 pi x), prediction at the right face's Gauss points): the error falls
    with the order (R1 > R2 > R3) and with dx in the pre-asymptotic range, to the
    levels the reference's weights reach.  (With ell = 12 dx the covariance is
    ill-conditioned and the FP64 weights saturate near 1e-9, so no asymptotic
    order is tested.)  A transposed face, point or stencil index fails this."""
    gpm = load_package()
    sx = np.array([0, -1, 0, 1, 0, -1, -1, 1, 1, 0, -2, 0, 2, -3, -2, -1, 0, 1, 2, 3, 2, 1, 0, -1, -2])
    errs = {}
    for n in (16, 32):
        t = gpm.gp_tables(7, 4, n)
        dx = 1.0 / n
        xc = 0.3
        avg = lambda x: (np.cos(2 * np.pi * (x - dx / 2)) - np.cos(2 * np.pi * (x + dx / 2))) / (2 * np.pi * dx)
        for ord_ in (3, 5, 7):
            z = zt_view(t["zT"], 7, 4)[ord_ - 1][:SZ[ord_]]
            pred = z[:, 2, :].T @ avg(xc + sx[:SZ[ord_]] * dx)
            errs[(ord_, n)] = np.abs(pred - np.sin(2 * np.pi * (xc + dx / 2))).max()
    assert errs[(3, 32)] < 2e-4 and errs[(5, 32)] < 1e-6 and errs[(7, 32)] < 1e-9
    for n in (16, 32):
        assert errs[(3, n)] > errs[(5, n)] > errs[(7, n)]
    assert errs[(3, 16)] / errs[(3, 32)] > 2 ** 3.5 and errs[(5, 16)] / errs[(5, 32)] > 2 ** 5.5
