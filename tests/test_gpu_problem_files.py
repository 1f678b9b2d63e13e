"""The reference's problem files at their own configurations and sizes
(problems/*.f90: grid, domain, method, integrator, numFlux, BC_type, Mord,
ngp, MOOD switches; tests/golden/hash_pf_*.npz written by the reference
Fortran, tests/golden/make_fixtures.py), run on the GPU through gpm_run: dt
sequence, detection counts and checksums of the final state (ghost cells
included) and order mask equal the reference's, bit for bit.  The initial
conditions are the seeded synthetic ones of gp-mood_amd/ic.py (no
transcendental functions, so they regenerate identically on any host)."""
import glob
import hashlib
import os

import numpy as np
import pytest

from conftest import GOLDEN, case_options, golden, has_gpu, load_oracle, tables_from_npz

CASES = sorted(os.path.basename(p)[5:-4] for p in glob.glob(os.path.join(GOLDEN, "hash_pf_*.npz")))


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def initial_state(gpm, z):
    o = case_options(z)
    kind, noise = str(z["kind"]).split(":")
    U0 = gpm.ic.problem(kind, int(z["n"]), o["nf"], o["lx"], o["ly"], seed=int(z["seed"]), noise=float(noise))
    assert sha(U0) == str(z["ic_sha"]), "initial condition does not regenerate bit for bit"
    return U0, o


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES)
def test_problem_file_run_matches_reference(gpm, case):
    if not has_gpu():
        pytest.skip("no GPU")
    z = golden(f"hash_{case}.npz")
    U0, o = initial_state(gpm, z)
    nsteps = int(z["nsteps"])
    print(f"[case {case}: {z['problem_file']} {z['n']}x{o['nf']}, {nsteps} steps]", flush=True)
    cfg = gpm.Config(lf=int(z["n"]), nf=o["nf"], Mord=int(z["mord"]), ngp=int(z["ngp"]), BC_type=str(z["bc"]),
                     DMP=bool(z["dmp"]), U2=bool(z["u2"]), U2_tol=bool(z["u2tol"]), numFlux=o["flux"],
                     integrator=o["integrator"], problem=o["problem"], Lx=o["lx"], Ly=o["ly"],
                     method={"fog": gpm.capi.FOG, "pol_mood": gpm.capi.POL_MOOD, "gp_se": gpm.capi.GP_SE,
                              "unlim_pol": gpm.capi.UNLIM_POL}.get(o["method"], gpm.capi.GP_MOOD))
    s = gpm.Solver(cfg, tables_from_npz(z))
    s.set_state(U0)
    dts = s.run(nsteps, return_dts=True)
    U, g = s.get_state(order=True)
    assert np.array_equal(dts, z["dt"]), f"dt differs from step {np.nonzero(dts != z['dt'])[0][:1]}"
    assert s.stats()["detected_total"] == int(np.sum(z["counts"]))
    assert sha(g[4:-4, 4:-4].astype(np.int8)) == str(z["mask_sha"])
    assert sha(U) == str(z["U_sha"]), f"sums {U.sum(axis=(0, 1))} vs {z['U_sum']}"


@pytest.mark.parametrize("case", CASES)
def test_problem_file_oracle_matches_reference(gpm, case):
    """The oracle on the same runs (its pin at the problem files' sizes; about
    15 s of CPU for all cases on 8 threads)."""
    O = load_oracle()
    z = golden(f"hash_{case}.npz")
    U0, o = initial_state(gpm, z)
    dx, dy = (float(v) for v in z["dxdy"])
    orc = O.Oracle(int(z["n"]), int(z["mord"]), int(z["ngp"]), str(z["bc"]), tables_from_npz(z),
                   dmp=bool(z["dmp"]), u2=bool(z["u2"]), u2tol=bool(z["u2tol"]), flux=o["flux"], nf=o["nf"],
                   integrator=o["integrator"], problem=o["problem"], dx=dx, dy=dy, method=o["method"])
    U = orc.full(U0)
    r = orc.run(U, int(z["nsteps"]), masks=True)
    assert np.array_equal(r["dt"], z["dt"])
    assert np.array_equal(r["counts"], z["counts"])
    assert sha(r["masks"][-1]) == str(z["mask_sha"])
    assert sha(U) == str(z["U_sha"])
