"""GPU step against the CPU oracle on configurations the reference fixtures do
not cover (partial tiles, cascade modes, switches, forced failures), at the
full 2048^2 size through invariants and the pinned oracle run, and the API
paths (drop-in vs device-resident, device tensors, determinism)."""
import glob
import hashlib
import os

import numpy as np
import pytest

from conftest import GOLDEN, has_gpu, load_oracle

pytestmark = pytest.mark.gpu
GAMMA = 1.4


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not has_gpu():
        pytest.skip("no GPU")


def both(gpm, n, mord, ngp, bc, U0, steps, **kw):
    """Run the GPU solver and the oracle on the same input and tables."""
    O = load_oracle()
    t = gpm.gp_tables(mord, ngp, n)
    cfg = gpm.Config(lf=n, Mord=mord, ngp=ngp, BC_type=bc, DMP=kw.get("dmp", True), U2=kw.get("u2", True),
                     U2_tol=kw.get("u2tol", True), cascade=kw.get("cascade", "reference"))
    s = gpm.Solver(cfg, t)
    s.set_state(U0)
    dts = s.run(steps, return_dts=True)
    U, g = s.get_state(order=True)
    o = O.Oracle(n, mord, ngp, bc, t, dmp=kw.get("dmp", True), u2=kw.get("u2", True), u2tol=kw.get("u2tol", True),
                 cascade=kw.get("cascade", "reference"))
    Uo = o.full(U0)
    r = o.run(Uo, steps, masks=True)
    return (dts, U, g, s.stats()), (r["dt"], Uo, r["masks"][-1], r)


def assert_same(gpu, orc_, equal_nan=False):
    (dts, U, g, st), (dto, Uo, go, r) = gpu, orc_
    assert np.array_equal(dts, dto, equal_nan=equal_nan), "dt sequence"
    assert np.array_equal(g[4:-4, 4:-4], go), f"order mask: {np.sum(g[4:-4, 4:-4] != go)} cells differ"
    assert np.array_equal(U, Uo, equal_nan=equal_nan), f"state: max diff {np.nanmax(np.abs(U - Uo))}"
    assert st["detected_total"] == int(r["counts"].sum())


@pytest.mark.parametrize("n,mord,ngp,bc,kind", [
    (40, 7, 4, "reflective", "blast"), (100, 5, 3, "periodic", "four_quadrant"),
    (72, 3, 2, "neumann", "four_quadrant"), (48, 7, 4, "neumann", "four_quadrant")])
def test_partial_tiles_vs_oracle(gpm, n, mord, ngp, bc, kind):
    print(f"[case n={n} R{(mord - 1) // 2} {bc} {kind}]", flush=True)
    assert_same(*both(gpm, n, mord, ngp, bc, gpm.ic.make(kind, n, n), 12))


@pytest.mark.parametrize("lf,nf,mord,ngp,bc,flux", [
    (8, 8, 7, 4, "reflective", "hllc"), (8, 8, 3, 2, "periodic", "hll"), (9, 8, 5, 3, "neumann", "hllc"),
    (8, 21, 3, 2, "kh", "unsplit"), (23, 8, 7, 4, "periodic", "llf"), (17, 13, 5, 3, "reflective", "hllc")])
def test_smallest_and_ragged_grids_vs_oracle(gpm, lf, nf, mord, ngp, bc, flux):
    """Grids of 8 cells a side (the smallest the library accepts) and ragged
    shapes below one stage tile, where the ghost rule, the halo and the MOOD
    neighbourhoods reach across the whole grid."""
    print(f"[case {lf}x{nf} R{(mord - 1) // 2} {bc} {flux}]", flush=True)
    O = load_oracle()
    lx, ly = 1.0, nf / lf
    t = gpm.gp_tables(mord, ngp, lf, nf, lx=lx, ly=ly)
    U0 = gpm.ic.problem("quadrants", lf, nf, lx, ly, seed=lf * nf, noise=0.05)
    cfg = gpm.Config(lf=lf, nf=nf, Lx=lx, Ly=ly, Mord=mord, ngp=ngp, BC_type=bc, numFlux=flux)
    s = gpm.Solver(cfg, t)
    s.set_state(U0)
    dts = s.run(15, return_dts=True)
    U, g = s.get_state(order=True)
    o = O.Oracle(lf, mord, ngp, bc, t, flux=flux, nf=nf, dx=lx / lf, dy=ly / nf)
    Uo = o.full(U0)
    r = o.run(Uo, 15, masks=True)
    assert np.array_equal(dts, r["dt"]), "dt sequence"
    assert np.array_equal(g[4:-4, 4:-4], r["masks"][-1])
    assert np.array_equal(U.view(np.uint64), Uo.view(np.uint64)), f"max diff {np.abs(U - Uo).max()}"
    assert s.stats()["detected_total"] == int(r["counts"].sum())


@pytest.mark.parametrize("mord,ngp", [(7, 4), (5, 3)])
def test_full_cascade_vs_oracle(gpm, mord, ngp):
    print(f"[case full cascade Mord {mord}]", flush=True)
    assert_same(*both(gpm, 64, mord, ngp, "reflective", gpm.ic.blast(64, 11), 15, cascade="full"))


@pytest.mark.parametrize("dmp,u2,u2tol", [(False, False, False), (True, True, False), (True, False, False)])
def test_switches_vs_oracle(gpm, dmp, u2, u2tol):
    print(f"[case DMP={dmp} U2={u2} U2_tol={u2tol}]", flush=True)
    assert_same(*both(gpm, 64, 5, 3, "neumann", gpm.ic.four_quadrant(64, 12), 15, dmp=dmp, u2=u2, u2tol=u2tol))


def test_positivity_cascade_vs_oracle(gpm):
    """Near-vacuum pocket: positivity failures cascade cells to first order."""
    U0 = gpm.ic.four_quadrant(32, 0)
    U0[14:18, 14:18] = gpm.ic.prim_to_cons(1e-3, 0.0, 0.0, 1e-9)
    gpu, orc_ = both(gpm, 32, 5, 3, "neumann", U0, 3)
    assert_same(gpu, orc_)
    assert np.all(np.isfinite(gpu[1]))


def test_inadmissible_state_propagates_like_the_reference(gpm):
    """A negative density is not an error: the face fluxes become Inf/NaN as
    in mod_HLLC.f90:166, MOOD lowers the cells it reaches to first order, and
    the GPU reproduces the oracle, NaNs included."""
    U0 = gpm.ic.four_quadrant(32, 1)
    U0[10, 20, 0] = -0.5
    gpu, orc_ = both(gpm, 32, 5, 3, "neumann", U0, 1)
    assert_same(gpu, orc_, equal_nan=True)
    assert gpu[2][4 + 10, 4 + 20] == 1


def test_time_stepping_matches_run(gpm):
    """The host-array drop-in (time_stepping with Setdt's dt) and the device
    loop produce the same states."""
    n = 64
    U0 = gpm.ic.four_quadrant(n, 5)
    cfg = gpm.Config(lf=n, Mord=7, ngp=4, BC_type="neumann")
    a = gpm.Solver(cfg)
    a.set_state(U0)
    a.run(4)
    Ua, ga = a.get_state(order=True)
    b = gpm.Solver(cfg)
    b.set_state(U0)
    U = b.get_state()
    t = 0.0
    for _ in range(4):
        b.set_state(U)
        dt = b.setdt(1e10 - t)
        U, g, st = b.time_stepping(U, t, dt)
        t += dt
    assert np.array_equal(U, Ua) and np.array_equal(g, ga)


def test_device_tensors_and_determinism(gpm):
    import torch
    n = 128
    U0 = gpm.ic.blast(n, 4)
    cfg = gpm.Config(lf=n, Mord=7, ngp=4, BC_type="reflective")
    outs = []
    for _ in range(2):
        s = gpm.Solver(cfg)
        dev = torch.from_numpy(U0.copy()).cuda()
        s.set_state_device(dev.data_ptr())
        s.run(10)
        out = torch.empty_like(dev)
        order = torch.empty((n, n), dtype=torch.int8, device="cuda")
        s.get_state_device(out.data_ptr(), order.data_ptr())
        torch.cuda.synchronize()
        U, g = s.get_state(order=True)
        assert np.array_equal(out.cpu().numpy(), U[4:-4, 4:-4])
        assert np.array_equal(order.cpu().numpy(), g[4:-4, 4:-4])
        outs.append(U)
    assert np.array_equal(outs[0], outs[1])


def test_setdt_matches_oracle(gpm):
    O = load_oracle()
    n = 96
    U0 = gpm.ic.four_quadrant(n, 8)
    s = gpm.Solver(gpm.Config(lf=n, Mord=5, ngp=3))
    s.set_state(U0)
    o = O.Oracle(n, 5, 3, "neumann", gpm.gp_tables(5, 3, n))
    assert s.setdt() == o.setdt(o.full(U0))
    assert s.setdt(1e-6) == 1e-6


def test_full_size_conservation(gpm):
    """2048^2, R3, reflective walls, blast: mass and energy are conserved to
    round-off (wall fluxes of mirrored states vanish); periodic smooth wave at
    2048^2, R2: all four totals conserved."""
    n = 2048
    print("[case 2048^2 reflective blast conservation]", flush=True)
    U0 = gpm.ic.blast(n, 0)
    s = gpm.Solver(gpm.Config(lf=n, Mord=7, ngp=4, BC_type="reflective"))
    s.set_state(U0)
    s.run(60)
    U = s.get_state()[4:-4, 4:-4]
    assert s.stats()["detected_total"] > 0
    for k in (0, 3):
        assert abs(U[..., k].sum() - U0[..., k].sum()) <= 1e-12 * abs(U0[..., k].sum())
    print("[case 2048^2 periodic smooth wave conservation]", flush=True)
    # advected smooth density wave: no cell is lowered, so every periodic face
    # flux is shared by its two cells and all four totals are conserved
    x = (np.arange(n) + 0.5) / n
    X, Y = np.meshgrid(x, x, indexing="xy")
    U0 = gpm.ic.prim_to_cons(1.0 + 0.2 * np.sin(2 * np.pi * X) * np.sin(2 * np.pi * Y), 0.5, 0.3, 1.0)
    s = gpm.Solver(gpm.Config(lf=n, Mord=5, ngp=3, BC_type="periodic"))
    s.set_state(U0)
    s.run(20)
    U, g = s.get_state(order=True)
    assert np.all(g == 5)
    tot0, tot1 = U0.sum(axis=(0, 1)), U[4:-4, 4:-4].sum(axis=(0, 1))
    assert np.all(np.abs(tot1 - tot0) <= 1e-13 * np.abs(U0).sum(axis=(0, 1)))


def _long_fixture(cascade):
    tag = "ref" if cascade == "reference" else "ref_full"
    paths = glob.glob(os.path.join(GOLDEN, f"hash_r3_n2048_{tag}_bl*_*.npz"))
    if not paths:
        pytest.skip("2048^2 fixture not generated")
    return np.load(paths[0], allow_pickle=False)


@pytest.mark.parametrize("cascade", ["reference", "full"])
def test_config1_2048_first_steps_vs_reference(gpm, cascade):
    """Config-1 workload at full size: the first steps equal the reference's
    (dt, order mask and state checksum pinned from the reference run; for the
    cascade 7 -> 5 -> 3 -> 1 from the one-line patched reference build)."""
    z = _long_fixture(cascade)
    n = int(z["n"])
    print(f"[case config1 2048^2 first steps, cascade {cascade}]", flush=True)
    U0 = gpm.ic.make(str(z["kind"]), n, int(z["seed"]))
    assert hashlib.sha256(U0.tobytes()).hexdigest() == str(z["ic_sha"])
    s = gpm.Solver(gpm.Config(lf=n, Mord=7, ngp=4, BC_type="reflective", cascade=cascade))
    s.set_state(U0)
    dts = s.run(int(z["pin_steps"]), return_dts=True)
    U, g = s.get_state(order=True)
    assert np.array_equal(dts, z["dt_pin"])
    assert np.array_equal(g[4:-4, 4:-4], z["mask_pin"])
    assert hashlib.sha256(np.ascontiguousarray(U[4:-4, 4:-4]).tobytes()).hexdigest() == str(z["U_pin_sha"])


@pytest.mark.parametrize("cascade", ["reference", "full"])
def test_config1_2048_500_steps_vs_oracle(gpm, cascade):
    """The full config-1 run (2048^2, R3, reflective blast, 500 steps) against
    the oracle's run (the oracle agrees with the reference, or with the
    patched reference for the cascade 7 -> 5 -> 3 -> 1, bit for bit on the
    first steps at this size): dt sequence, detection counts and checksum."""
    z = _long_fixture(cascade)
    n = int(z["n"])
    print(f"[case config1 2048^2 x {int(z['nsteps'])} steps, cascade {cascade}]", flush=True)
    U0 = gpm.ic.make(str(z["kind"]), n, int(z["seed"]))
    s = gpm.Solver(gpm.Config(lf=n, Mord=7, ngp=4, BC_type="reflective", cascade=cascade))
    s.set_state(U0)
    dts = s.run(int(z["nsteps"]), return_dts=True)
    U = s.get_state()
    assert np.array_equal(dts, z["dt"])
    assert s.stats()["detected_total"] == int(np.sum(z["counts"]))
    assert hashlib.sha256(np.ascontiguousarray(U[4:-4, 4:-4]).tobytes()).hexdigest() == str(z["U_sha"])


@pytest.mark.parametrize("mord,ngp,cascade,bc", [(5, 3, "reference", "neumann"), (7, 4, "full", "periodic")])
def test_face_flux_cache_changes_no_bit(gpm, mord, ngp, cascade, bc):
    """Workgroup 0's MOOD passes reuse cached face fluxes, compute the likely
    next ones ahead, and read the device-wide cache that the other
    workgroups of the persistent kernel fill meanwhile (switched off at
    gpm_create by GPM_MOOD_FC=0 and GPM_MOOD_HELP=0).  1024^2 four-quadrant
    problems, whose failure chains run long runs of small passes on
    workgroup 0, give the same dt sequence, counts, order mask and state in
    every combination and in both MOOD launch forms."""
    n, steps = 1024, 12
    print(f"[case face-flux caches on/off, {n}^2 R{(mord - 1) // 2} {cascade} cascade {bc}, {steps} steps]", flush=True)
    U0 = gpm.ic.four_quadrant(n, 0)
    t = gpm.gp_tables(mord, ngp, n)
    out = {}
    variants = {"both": {"GPM_MOOD_FC": "1", "GPM_MOOD_HELP": "1"}, "lds": {"GPM_MOOD_FC": "1", "GPM_MOOD_HELP": "0"},
                "none": {"GPM_MOOD_FC": "0", "GPM_MOOD_HELP": "0"}}
    for name, env in variants.items():
        for coop in (True, False):
            old = {k: os.environ.get(k) for k in env}
            os.environ.update(env)
            try:
                s = gpm.Solver(gpm.Config(lf=n, Mord=mord, ngp=ngp, BC_type=bc, cascade=cascade), t)
            finally:
                for k, v in old.items():
                    if v is None:
                        del os.environ[k]
                    else:
                        os.environ[k] = v
            s.set_mood_launch(coop)
            s.set_state(U0)
            dts = s.run(steps, return_dts=True)
            U, g = s.get_state(order=True)
            st = s.stats()
            s.close()
            out[(name, coop)] = (dts, U, g, st)
            print(f"  caches={name} cooperative={coop}: max_passes {st['max_passes']}, "
                  f"second-pass cells {st['second_pass_cells']}", flush=True)
    ref = out[("none", True)]
    assert ref[3]["max_passes"] > 20
    for key, (dts, U, g, st) in out.items():
        assert np.array_equal(dts, ref[0]), key
        assert np.array_equal(U, ref[1]), key
        assert np.array_equal(g, ref[2]), key
        for k in ("detected_total", "second_pass_cells", "max_passes"):
            assert st[k] == ref[3][k], (key, k)


def test_kernel_timing_every_kth_step(gpm):
    """gpm_set_timing: events around the stage and MOOD launches of every step
    (1) or of every k-th step (k), none when off; the averages are positive
    and the results do not depend on the timing."""
    n = 128
    print(f"[case kernel timing {n}^2]", flush=True)
    U0 = gpm.ic.four_quadrant(n, 4)
    out = {}
    for every in (0, 1, 3):
        s = gpm.Solver(gpm.Config(lf=n, Mord=5, ngp=3))
        s.set_state(U0)
        s.set_timing(every > 0, every=max(every, 1))
        dts = s.run(7, return_dts=True)
        kt = s.kernel_time()
        out[every] = (dts, s.get_state(), kt)
        s.close()
    assert out[0][2]["stage_launches"] == 0 and out[0][2]["mood_launches"] == 0
    assert out[1][2]["stage_launches"] == 21 and out[1][2]["mood_launches"] == 21
    assert out[3][2]["stage_launches"] == 9 and out[3][2]["mood_launches"] == 9  # steps 0, 3, 6
    for k in (1, 3):
        assert out[k][2]["stage_ms"] > 0 and out[k][2]["mood_ms"] > 0
        assert np.array_equal(out[k][0], out[0][0]) and np.array_equal(out[k][1], out[0][1])
