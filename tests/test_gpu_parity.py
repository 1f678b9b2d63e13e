"""GPU parity of the HIP step against the reference (golden fixtures written by
the reference Fortran build) -- bit-exact state, dt sequence and order masks.

The fixtures come from tests/golden/make_fixtures.py; the reference's own GP
tables are loaded into the solver so that both sides use identical weights.
"""
import glob
import os

import numpy as np
import pytest

from conftest import GOLDEN, case_options, golden, has_gpu, load_oracle, tables_from_npz

pytestmark = pytest.mark.gpu

STEP_CASES = sorted(os.path.basename(p)[5:-4] for p in glob.glob(os.path.join(GOLDEN, "step_*.npz")))


def make_config(gpm, z):
    o = case_options(z)
    return gpm.Config(lf=int(z["n"]), nf=o["nf"], Mord=int(z["mord"]), ngp=int(z["ngp"]), BC_type=str(z["bc"]),
                      DMP=bool(z["dmp"]), U2=bool(z["u2"]), U2_tol=bool(z["u2tol"]), numFlux=o["flux"],
                      integrator=o["integrator"], problem=o["problem"], Lx=o["lx"], Ly=o["ly"], cascade=o["cascade"],
                      method={"fog": gpm.capi.FOG, "pol_mood": gpm.capi.POL_MOOD, "gp_se": gpm.capi.GP_SE,
                              "unlim_pol": gpm.capi.UNLIM_POL}.get(o["method"], gpm.capi.GP_MOOD))


def make_solver(gpm, z, tables=True):
    return gpm.Solver(make_config(gpm, z), tables_from_npz(z) if tables else None)


def oracle_first_step(z, U):
    """The oracle's time_stepping from full array U at t = 0 (pinned to the
    reference by the golden fixtures)."""
    O = load_oracle()
    o = case_options(z)
    dx, dy = (float(v) for v in z["dxdy"])
    orc = O.Oracle(int(z["n"]), int(z["mord"]), int(z["ngp"]), str(z["bc"]), tables_from_npz(z),
                   dmp=bool(z["dmp"]), u2=bool(z["u2"]), u2tol=bool(z["u2tol"]), flux=o["flux"], nf=o["nf"],
                   integrator=o["integrator"], problem=o["problem"], dx=dx, dy=dy, method=o["method"],
                   cascade=o["cascade"])
    Uo = U.copy()
    return orc.step(Uo, float(z["dt"][0]), 0.0), Uo


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not has_gpu():
        pytest.skip("no GPU")


@pytest.mark.parametrize("case", STEP_CASES)
def test_first_step_stages(gpm, case):
    """One drop-in time_stepping call: final state (ghost cells included),
    refilled input ghosts and mask bit-identical."""
    z = golden(f"step_{case}.npz")
    print(f"[case {case}] first step via gpm_time_stepping", flush=True)
    cfg = make_config(gpm, z)
    s = gpm.Solver(cfg, tables_from_npz(z))
    U = np.zeros(s.full_shape())
    U[4:-4, 4:-4] = z["U0"]
    gpm.boundary_c(cfg, U, 0.0)  # main.f90:125
    s.set_state(U)
    Uin = U.copy()
    Uout, g, st = s.time_stepping(Uin, 0.0, float(z["dt"][0]))
    ref, ref_in = oracle_first_step(z, U)
    if "stage_U" in z.files:
        assert np.array_equal(ref, z["stage_U"][2]), "oracle differs from the reference"
        assert np.array_equal(g[4:-4, 4:-4], z["stage_masks"][2]), "order mask of the last stage differs"
    assert np.array_equal(Uout[4:-4, 4:-4], ref[4:-4, 4:-4]), \
        f"max |diff| {np.abs(Uout[4:-4, 4:-4] - ref[4:-4, 4:-4]).max()}"
    assert np.array_equal(Uout, ref), f"ghost cells differ (max {np.abs(Uout - ref).max()})"
    assert np.array_equal(Uin, ref_in), "refilled input ghost cells differ"
    assert st["count_detected_cell"] == int(z["counts"][0])


@pytest.mark.parametrize("case", STEP_CASES)
def test_run_matches_reference(gpm, case):
    """Device-resident loop (dt on device): dt sequence, per-step detection
    counts, last order mask and final state all bit-identical."""
    z = golden(f"step_{case}.npz")
    nsteps = int(z["nsteps"])
    print(f"[case {case}] {nsteps} steps via gpm_run", flush=True)
    s = make_solver(gpm, z)
    s.set_state(z["U0"])
    s.set_time(0.0)
    dts = s.run(nsteps, return_dts=True)
    U, g = s.get_state(order=True)
    assert np.array_equal(dts, z["dt"]), f"dt differs at steps {np.nonzero(dts != z['dt'])[0][:5]}"
    assert np.array_equal(g[4:-4, 4:-4], z["masks"][-1])
    assert np.array_equal(U, z["U"]), f"max |diff| {np.abs(U - z['U']).max()}"
    st = s.stats()
    assert st["steps"] == nsteps
    assert st["detected_total"] == int(np.sum(z["counts"]))


@pytest.mark.parametrize("case", STEP_CASES)
def test_run_per_pass_launches(gpm, case):
    """The MOOD correction passes as per-pass launches instead of one
    persistent launch per stage (set_mood_launch(False)): the same dt
    sequence, counts, mask and state as the reference run."""
    z = golden(f"step_{case}.npz")
    nsteps = int(z["nsteps"])
    print(f"[case {case}] {nsteps} steps via gpm_run, per-pass MOOD launches", flush=True)
    s = make_solver(gpm, z)
    s.set_mood_launch(False)
    s.set_state(z["U0"])
    s.set_time(0.0)
    dts = s.run(nsteps, return_dts=True)
    U, g = s.get_state(order=True)
    assert not s.mood_launch()
    assert np.array_equal(dts, z["dt"]), f"dt differs at steps {np.nonzero(dts != z['dt'])[0][:5]}"
    assert np.array_equal(g[4:-4, 4:-4], z["masks"][-1])
    assert np.array_equal(U, z["U"]), f"max |diff| {np.abs(U - z['U']).max()}"
    assert s.stats()["detected_total"] == int(np.sum(z["counts"]))


@pytest.mark.parametrize("case", STEP_CASES)
def test_run_wide_launches(gpm, case, monkeypatch):
    """Every stage's first two MOOD correction passes through the wide
    launches (GPM_MOOD_WIDE_MIN=1: any list of more than one cell; by default
    only lists longer than one trip of the persistent grid go there), the
    rest in the persistent kernel: the same dt sequence, counts, mask and
    state as the reference run."""
    z = golden(f"step_{case}.npz")
    nsteps = int(z["nsteps"])
    print(f"[case {case}] {nsteps} steps via gpm_run, wide MOOD launches for every list", flush=True)
    monkeypatch.setenv("GPM_MOOD_WIDE", "2")
    monkeypatch.setenv("GPM_MOOD_WIDE_MIN", "1")
    s = make_solver(gpm, z)
    s.set_state(z["U0"])
    s.set_time(0.0)
    dts = s.run(nsteps, return_dts=True)
    U, g = s.get_state(order=True)
    assert np.array_equal(dts, z["dt"]), f"dt differs at steps {np.nonzero(dts != z['dt'])[0][:5]}"
    assert np.array_equal(g[4:-4, 4:-4], z["masks"][-1])
    assert np.array_equal(U, z["U"]), f"max |diff| {np.abs(U - z['U']).max()}"
    st = s.stats()
    assert st["steps"] == nsteps
    assert st["detected_total"] == int(np.sum(z["counts"]))


@pytest.mark.parametrize("form", ["cooperative", "ordinary", "alone"])
@pytest.mark.parametrize("case", STEP_CASES)
def test_run_persistent_launch_forms(gpm, case, form, monkeypatch, capfd):
    """The persistent MOOD kernel in each of its launch forms: a cooperative
    launch with the hardware's grid barrier (the default, GPM_MOOD_ORD=0); an
    ordinary launch whose workgroups meet at a rendezvous and then use grid
    barriers of their own (GPM_MOOD_ORD=1, the form taken when the device
    refuses a cooperative launch); and the ordinary launch with
    GPM_MOOD_RV_ALONE=1, where the first workgroup to arrive runs every pass
    alone and the others leave (the path taken when they are not all
    resident together).  The same dt sequence, counts, mask and state as the
    reference run."""
    z = golden(f"step_{case}.npz")
    nsteps = int(z["nsteps"])
    print(f"[case {case}] {nsteps} steps via gpm_run, {form} persistent MOOD launch", flush=True)
    monkeypatch.setenv("GPM_MOOD_ORD", "0" if form == "cooperative" else "1")
    monkeypatch.setenv("GPM_MOOD_RV_ALONE", "1" if form == "alone" else "0")
    s = make_solver(gpm, z)
    s.set_state(z["U0"])
    s.set_time(0.0)
    dts = s.run(nsteps, return_dts=True)
    U, g = s.get_state(order=True)
    assert np.array_equal(dts, z["dt"]), f"dt differs at steps {np.nonzero(dts != z['dt'])[0][:5]}"
    assert np.array_equal(g[4:-4, 4:-4], z["masks"][-1])
    assert np.array_equal(U, z["U"]), f"max |diff| {np.abs(U - z['U']).max()}"
    st = s.stats()
    assert st["detected_total"] == int(np.sum(z["counts"]))
    # the persistent kernel ran (every stage with detection launches it) and,
    # when asked, on one workgroup
    said = "ran on one workgroup" in capfd.readouterr().err
    assert said == (form == "alone" and st["max_passes"] > 0)


def test_ordinary_launch_stall_is_reported(gpm, monkeypatch):
    """A grid-barrier wait of the ordinary persistent launch that reaches its
    bound (forced from the host: GPM_MOOD_BAR_SPINS=0, no poll before giving
    up) ends the launch and fails the call with an error instead of waiting
    on; a solver created afterwards without the switch runs the case exactly."""
    z = golden("step_r3_n64_ref_bl6.npz")
    print("[case r3_n64_ref_bl6] 3 steps, ordinary persistent MOOD launch, forced barrier stall", flush=True)
    monkeypatch.setenv("GPM_MOOD_ORD", "1")
    monkeypatch.setenv("GPM_MOOD_BAR_SPINS", "0")
    s = make_solver(gpm, z)
    s.set_state(z["U0"])
    s.set_time(0.0)
    with pytest.raises(gpm.GpmError, match="grid barrier"):
        s.run(3)
    s.close()
    monkeypatch.delenv("GPM_MOOD_BAR_SPINS")
    s = make_solver(gpm, z)
    s.set_state(z["U0"])
    s.set_time(0.0)
    dts = s.run(int(z["nsteps"]), return_dts=True)
    assert np.array_equal(dts, z["dt"])
    assert np.array_equal(s.get_state(), z["U"])


def test_config0_256_300_steps_checksum(gpm):
    """Config-0 workload (256^2, R2/ngp3, Neumann, 300 steps, seeded
    four-quadrant IC): final state and mask checksums equal the reference's."""
    import hashlib
    z = golden("hash_r2_n256_neu_fq0_300.npz")
    print("[case config0 256^2 x 300 steps]", flush=True)
    gpm_ic = gpm.ic
    U0 = gpm_ic.make(str(z["kind"]), int(z["n"]), int(z["seed"]))
    assert hashlib.sha256(U0.tobytes()).hexdigest() == str(z["ic_sha"])
    s = make_solver(gpm, z)
    s.set_state(U0)
    dts = s.run(int(z["nsteps"]), return_dts=True)
    U, g = s.get_state(order=True)
    Ui = np.ascontiguousarray(U[4:-4, 4:-4])
    assert np.array_equal(dts, z["dt"])
    assert hashlib.sha256(np.ascontiguousarray(g[4:-4, 4:-4].astype(np.int8)).tobytes()).hexdigest() == str(z["mask_sha"])
    assert hashlib.sha256(Ui.tobytes()).hexdigest() == str(z["U_sha"]), \
        f"sums {Ui.sum(axis=(0, 1))} vs {z['U_sum']}"
