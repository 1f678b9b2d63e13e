"""GPU parity of the HIP step against the reference (golden fixtures written by
the reference Fortran build) -- bit-exact state, dt sequence and order masks.

The fixtures come from tests/golden/make_fixtures.py; the reference's own GP
tables are loaded into the solver so that both sides use identical weights.
"""
import glob
import os

import numpy as np
import pytest

from conftest import GOLDEN, golden, has_gpu, tables_from_npz

pytestmark = pytest.mark.gpu

STEP_CASES = sorted(os.path.basename(p)[5:-4] for p in glob.glob(os.path.join(GOLDEN, "step_*.npz")))
BC_NAME = {"neumann": "neumann", "periodic": "periodic", "reflective": "reflective"}


def make_solver(gpm, z, tables=True):
    cfg = gpm.Config(lf=int(z["n"]), Mord=int(z["mord"]), ngp=int(z["ngp"]), BC_type=str(z["bc"]),
                     DMP=bool(z["dmp"]), U2=bool(z["u2"]), U2_tol=bool(z["u2tol"]),
                     numFlux=str(z["flux"]) if "flux" in z else "hllc")
    return gpm.Solver(cfg, tables_from_npz(z) if tables else None)


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not has_gpu():
        pytest.skip("no GPU")


@pytest.mark.parametrize("case", STEP_CASES)
def test_first_step_stages(gpm, case):
    """One drop-in time_stepping call: final state and mask bit-identical."""
    z = golden(f"step_{case}.npz")
    print(f"[case {case}] first step via gpm_time_stepping", flush=True)
    s = make_solver(gpm, z)
    n = int(z["n"])
    U = np.zeros((n + 8, n + 8, 4))
    U[4:-4, 4:-4] = z["U0"]
    s.set_state(U)
    U = s.get_state()  # ghosts filled by the boundary rule
    Uout, g, st = s.time_stepping(U, 0.0, float(z["dt"][0]))
    ref = z["stage_U"][2]
    assert np.array_equal(g[4:-4, 4:-4], z["stage_masks"][2]), "order mask of stage 3 differs"
    assert np.array_equal(Uout[4:-4, 4:-4], ref[4:-4, 4:-4]), \
        f"max |diff| {np.abs(Uout[4:-4, 4:-4] - ref[4:-4, 4:-4]).max()}"
    assert np.array_equal(Uout, ref), "ghost cells differ"
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
