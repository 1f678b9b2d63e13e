"""Seeded initial conditions: shapes, determinism and the stated properties."""
import numpy as np

from conftest import load_package


def test_four_quadrant_ranges_and_determinism():
    ic = load_package().ic
    a, b = ic.four_quadrant(64, 3), ic.four_quadrant(64, 3)
    assert a.shape == (64, 64, 4) and a.flags.c_contiguous
    assert np.array_equal(a, b)
    assert not np.array_equal(a, ic.four_quadrant(64, 4))
    for (rho, u, v, p) in ic.four_quadrant_states(3):
        assert 0.1 <= rho <= 1.5 and -1 <= u <= 1 and -1 <= v <= 1 and 0.03 <= p <= 1.5
    # four constant quadrants split at (0.5, 0.5)
    assert len(np.unique(a[:32, :32].reshape(-1, 4), axis=0)) == 1
    assert len(np.unique(a.reshape(-1, 4), axis=0)) == 4


def test_blast_energy_and_noise():
    ic = load_package().ic
    n = 128
    U = ic.blast(n, 0)
    rho = U[..., 0]
    assert np.all(np.abs(rho - 1.0) <= 0.01 + 1e-15)
    assert np.all(U[..., 1] == 0) and np.all(U[..., 2] == 0)
    e_bg = 1e-5 / (1.4 - 1.0)
    added = (U[..., 3] - e_bg).sum() / n ** 2
    assert abs(added - 1.0) < 1e-12
    # the blast is centred in [0.3, 0.7]^2
    j, i = np.unravel_index(np.argmax(U[..., 3]), (n, n))
    assert 0.29 <= (i + 0.5) / n <= 0.71 and 0.29 <= (j + 0.5) / n <= 0.71


def test_shear_layer():
    ic = load_package().ic
    n = 64
    U = ic.shear_layer(n, 1)
    x = (np.arange(n) + 0.5) / n
    inner = np.abs(x - 0.5) < 0.25
    assert np.allclose(U[inner, :, 0], 2.0) and np.allclose(U[~inner, :, 0], 1.0)
    u = U[..., 1] / U[..., 0]
    assert np.allclose(u[inner], 0.5) and np.allclose(u[~inner], -0.5)
    v = U[..., 2] / U[..., 0]
    assert 0 < np.abs(v).max() < 0.08
    p = (1.4 - 1) * (U[..., 3] - 0.5 * (U[..., 1] ** 2 + U[..., 2] ** 2) / U[..., 0])
    assert np.allclose(p, 2.5)
