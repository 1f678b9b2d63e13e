"""The library's boundary rule (gpm_boundary_c: the host instance of the
template its kernels use) against the oracle's Boundary_C (oracle/gpm_oracle.c,
itself pinned to the reference by the golden step fixtures, whose final
states include the ghost cells), for every BC_type, on rectangular grids, at
several stage times.  Host code only: runs without a GPU."""
import zlib

import numpy as np
import pytest

from conftest import load_package

BCS = ["neumann", "periodic", "reflective", "kh", "dirichlet", "mach800", "doublemach800", "rmi", "dmr", "rt"]
# (lf, nf, Lx, Ly): square and rectangular grids, square cells
GRIDS = [(16, 16, 1.0, 1.0), (24, 8, 1.5, 0.5), (16, 48, 0.5, 1.5), (64, 16, 4.0, 1.0), (64, 8, 9.0, 1.125)]


def random_interior(rng, lf, nf):
    r = rng.uniform(0.2, 2.0, size=(nf, lf))
    u, v = rng.uniform(-1, 1, size=(2, nf, lf))
    p = rng.uniform(0.1, 3.0, size=(nf, lf))
    return np.stack([r, r * u, r * v, p / 0.4 + 0.5 * r * (u * u + v * v)], axis=-1)


@pytest.mark.parametrize("bc", BCS)
@pytest.mark.parametrize("grid", GRIDS)
def test_boundary_rule_matches_oracle(orc, bc, grid):
    gpm = load_package()
    lf, nf, lx, ly = grid
    rng = np.random.default_rng(zlib.crc32(f"{bc}{grid}".encode()))
    interior = random_interior(rng, lf, nf)
    t = gpm.gp_tables(3, 2, lf, nf, lx=lx, ly=ly)
    cfg = gpm.Config(lf=lf, nf=nf, Mord=3, ngp=2, BC_type=bc, Lx=lx, Ly=ly)
    o = orc.Oracle(lf, 3, 2, bc, t, nf=nf, lx=lx, ly=ly)
    for t_rk in (0.0, 0.013, 0.1):
        o.set_t_rk(t_rk)
        expected = o.full(interior)
        got = np.zeros_like(expected)
        got[4:-4, 4:-4] = interior
        gpm.boundary_c(cfg, got, t_rk)
        assert np.array_equal(got, expected), f"{bc} {grid} t_rk={t_rk}"


def test_dmr_top_boundary_moves_with_t_rk(orc):
    """The DMR top ghost rows switch from the post- to the pre-shock state at
    x = 1/6 + (1 + 20 t) / sqrt(3) (BC.f90:301): the switch column moves right
    with t_rk by 20/sqrt(3) per unit time."""
    gpm = load_package()
    lf, nf = 64, 16
    cfg = gpm.Config(lf=lf, nf=nf, Mord=3, ngp=2, BC_type="dmr", Lx=4.0, Ly=1.0)
    interior = random_interior(np.random.default_rng(1), lf, nf)
    x = (np.arange(-4, lf + 4) + 1) * (4.0 / lf) - (4.0 / lf) / 2
    for t_rk in (0.0, 0.05, 0.2):
        U = np.zeros((nf + 8, lf + 8, 4))
        U[4:-4, 4:-4] = interior
        gpm.boundary_c(cfg, U, t_rk)
        post = np.isclose(U[-1, :, 0], 8.0)
        xs = 1 / 6 + (1 + 20 * t_rk) / np.sqrt(3)
        assert np.array_equal(post, x < xs)
