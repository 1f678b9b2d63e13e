"""POL_MOOD (mod_pol_MOOD.f90) on the CPU: the oracle's polynomial face
weights against the reference's own Pol_zT_o3, and the step cases' use of
the order-1 and U2 paths that set it apart from GP-MOOD."""
import glob
import os

import numpy as np
import pytest

from conftest import GOLDEN, case_options, tables_from_npz

POL_CASES = sorted(os.path.basename(f) for f in glob.glob(os.path.join(GOLDEN, "step_pol_*.npz")))


def pol_oracle(orc, z):
    o = case_options(z)
    dx, dy = (float(v) for v in z["dxdy"])
    return orc.Oracle(int(z["n"]), int(z["mord"]), int(z["ngp"]), str(z["bc"]), tables_from_npz(z), dmp=bool(z["dmp"]),
                      u2=bool(z["u2"]), u2tol=bool(z["u2tol"]), flux=o["flux"], nf=o["nf"], integrator=o["integrator"],
                      dx=dx, dy=dy, method=o["method"])


def test_cases_present():
    assert len(POL_CASES) >= 4


@pytest.mark.parametrize("case", POL_CASES)
def test_pol_table_matches_reference(orc, case):
    """pol_presim (GP_init.f90:359-411) restated: Pol_zT_o3 bit for bit as the
    reference build holds it, and consistent: each face's weights sum to 1
    (a constant is reproduced) up to rounding."""
    z = np.load(os.path.join(GOLDEN, case))
    oc = pol_oracle(orc, z)
    mine = oc.pol_table()
    assert np.array_equal(mine.view(np.uint64), z["pol"].view(np.uint64))
    w = mine.reshape(2, 4, 5)  # [gp][face][s]
    assert np.allclose(w.sum(axis=2), 1.0, atol=1e-15)


def test_pol_cases_exercise_lowering_and_differ_from_gp(orc):
    """The POL cases lower cells (so order-1 cell values and the D2O2 U2
    test run), and POL_MOOD's result is not GP-MOOD's on the same input."""
    lowered = 0
    for case in POL_CASES:
        z = np.load(os.path.join(GOLDEN, case))
        lowered += int((z["masks"] < 3).sum())
    assert lowered > 1000
    z = np.load(os.path.join(GOLDEN, POL_CASES[0]))
    oc = pol_oracle(orc, z)
    orc.lib().orc_set_method(oc._h, orc.METHODS["gp_mood"])
    U = oc.full(z["U0"])
    oc.run(U, 5)
    oc2 = pol_oracle(orc, z)
    U2 = oc2.full(z["U0"])
    oc2.run(U2, 5)
    assert not np.array_equal(U, U2)
