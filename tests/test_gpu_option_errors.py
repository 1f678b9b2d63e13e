"""Option combinations the library refuses, and the reference-network entry
points' argument checks, through the C-ABI (each returns its GPM_E* code and
leaves a message; nothing runs)."""
import numpy as np
import pytest

from conftest import has_gpu

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not has_gpu():
        pytest.skip("no GPU")


def create_rc(gpm, **kw):
    c = gpm.Config(**kw).to_c()
    import ctypes
    h = ctypes.c_void_p()
    rc = gpm.capi.load().gpm_create(ctypes.byref(c), ctypes.byref(h))
    if rc == 0:
        gpm.capi.load().gpm_destroy(h)
    return rc


@pytest.mark.parametrize("kw", [
    dict(lf=32, Mord=5, ngp=3, method=200),                              # POL_MOOD above Mord 3
    dict(lf=32, Mord=3, ngp=3, method=201),                              # Unlim_POL with 3 Gauss points
    dict(lf=32, Mord=3, ngp=2, method=1),                                # FOG with ngp 2
    dict(lf=32, Mord=5, ngp=3, method=202, integrator="ssp_rk4"),        # a method without detection, SSP-RK4
    dict(lf=32, Mord=3, ngp=1, method=1, integrator="ssp_rk4", BC_type="dmr"),  # FOG, SSP-RK4
    dict(lf=64, Mord=5, ngp=3, method=204),                              # NN_GP_MOOD_CC above Mord 3
    dict(lf=64, Mord=3, ngp=2, method=203, BC_type="dirichlet"),         # the network with a formula boundary
    dict(lf=32, Mord=9, ngp=4),                                          # no such order
])
def test_unsupported_combinations(gpm, kw):
    print(f"[case create {kw}]", flush=True)
    assert create_rc(gpm, **kw) == -4  # GPM_EUNSUP


def test_supported_combinations(gpm):
    for kw in (dict(lf=32, Mord=3, ngp=2, method=200), dict(lf=32, Mord=7, ngp=4, method=202, integrator="fe"),
               dict(lf=32, Mord=3, ngp=1, method=1, integrator="ssp_rk3"), dict(lf=64, Mord=3, ngp=2, method=204),
               dict(lf=32, nf=48, Mord=5, ngp=3, method=203, integrator="ssp_rk2", BC_type="kh"),
               dict(lf=32, Mord=3, ngp=2, BC_type="dmr", integrator="ssp_rk4"),
               dict(lf=32, Mord=5, ngp=3, BC_type="rt", problem=gpm.capi.PROBLEM_RT, integrator="ssp_rk4")):
        assert create_rc(gpm, **kw) == 0, kw


def net(L, length):
    rng = np.random.default_rng(0)
    return dict(weight0=rng.normal(size=(length, L)), bias0=np.zeros(length), weight1=rng.normal(size=(length, length)),
                bias1=np.zeros(length), weight2=rng.normal(size=(2, length)), bias2=np.zeros(2))


def test_reference_network_argument_checks(gpm, tmp_path):
    s = gpm.Solver(gpm.Config(lf=32, Mord=3, ngp=2, method=gpm.capi.NN_GP_MOOD))
    with pytest.raises(gpm.GpmError, match=r"\(-1\)"):
        s.set_reference_nn(**net(40, 60))          # L below the 52 stencil values
    with pytest.raises(gpm.GpmError, match=r"\(-4\)"):
        s.set_reference_nn(**net(57, 65))          # more hidden units than the kernel holds
    with pytest.raises(gpm.GpmError, match=r"\(-1\)"):
        s.set_reference_nn(**net(129, 60))         # more inputs than the kernel holds
    with pytest.raises(gpm.GpmError, match="cannot read"):
        s.load_reference_nn(str(tmp_path / "missing"), 57, 60)
    with open(str(tmp_path / "short.txt"), "w") as f:
        f.write("\n".join(["0.5"] * 100))
    with pytest.raises(gpm.GpmError, match="cannot read"):
        s.load_reference_nn(str(tmp_path / "short"), 57, 60)
    s.set_reference_nn(**net(57, 60))
    p = gpm.Solver(gpm.Config(lf=32, Mord=3, ngp=2))
    with pytest.raises(gpm.GpmError, match=r"\(-5\)"):
        p.set_reference_nn(**net(57, 60))          # GP_MOOD has no network
    cc = gpm.Solver(gpm.Config(lf=16, Mord=3, ngp=2, method=gpm.capi.NN_GP_MOOD_CC))
    with pytest.raises(gpm.GpmError, match="values_NN"):
        cc.set_reference_nn(**net(60, 60))         # values_NN(L, n) past the array at lf = 16
    import torch
    x = torch.zeros((10, 57), dtype=torch.float32, device="cuda")
    y = torch.zeros((10, 2), dtype=torch.float32, device="cuda")
    cnt = torch.zeros(1, dtype=torch.int64, device="cuda")
    nr = torch.zeros(2, dtype=torch.int64, device="cuda")
    with pytest.raises(gpm.GpmError, match=r"\(-1\)"):
        s.set_reference_dump(x.data_ptr(), y.data_ptr(), 0, 10, cnt.data_ptr(), nr.data_ptr(), 0, 40)  # L < 52
    with pytest.raises(gpm.GpmError, match=r"\(-1\)"):
        s.set_niter(-1)
