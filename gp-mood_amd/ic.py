"""Seeded synthetic initial conditions for the GP-MOOD 2D Euler step.

Every generator returns the conserved interior state as a C-contiguous float64
array of shape ``(N, N, 4)`` indexed ``[n-1, l-1, k]``: ``n`` is the cell row
(y), ``l`` the cell column (x), ``k`` the variable ``rho, rho*u, rho*v, E``.
This is byte-identical to the reference's interior block ``U(4, 1:lf, 1:nf)``
(global_variables.f90:26, column-major, variable fastest).

Cell centres follow ``mesh_init.f90:14-20``: ``x_l = l*dx - dx/2`` on [0, 1].
The primitive-to-conserved map follows ``physics.f90:146-160``:
``E = p/(gamma-1) + 0.5*rho*(u**2+v**2)`` with ``gamma - 1`` evaluated in FP64
from ``gamma = 1.4`` (constants.f90:9).
"""
import numpy as np

GAMMA = 1.4
GM1 = GAMMA - 1.0


def cell_centres(n):
    dx = 1.0 / n
    idx = np.arange(1, n + 1, dtype=np.float64)
    return idx * dx - dx / 2


def prim_to_cons(rho, u, v, p):
    """Conserved state from primitive arrays (physics.f90:155-159 evaluation order)."""
    rho, u, v, p = np.broadcast_arrays(*(np.asarray(a, dtype=np.float64) for a in (rho, u, v, p)))
    out = np.empty(rho.shape + (4,), dtype=np.float64)
    out[..., 0] = rho
    out[..., 1] = rho * u
    out[..., 2] = rho * v
    out[..., 3] = (p / GM1) + 0.5 * rho * (u ** 2 + v ** 2)
    return np.ascontiguousarray(out)


def four_quadrant_states(seed):
    """The four constant primitive states (BL, BR, TL, TR) of a seeded quadrant IC.

    rho ~ U[0.1, 1.5], u, v ~ U[-1, 1], p ~ U[0.03, 1.5], drawn in that order per
    quadrant from numpy's PCG64 generator seeded with ``seed``.
    """
    rng = np.random.default_rng(seed)
    states = []
    for _ in range(4):
        rho = rng.uniform(0.1, 1.5)
        u = rng.uniform(-1.0, 1.0)
        v = rng.uniform(-1.0, 1.0)
        p = rng.uniform(0.03, 1.5)
        states.append((rho, u, v, p))
    return states


def four_quadrant(n, seed):
    """Four quadrants split at (0.5, 0.5), each a seeded constant state."""
    x = cell_centres(n)
    X, Y = np.meshgrid(x, x, indexing="xy")  # X[n, l] = x_l, Y[n, l] = y_n
    states = four_quadrant_states(seed)
    q = np.where(Y < 0.5, np.where(X < 0.5, 0, 1), np.where(X < 0.5, 2, 3))
    prim = np.array(states, dtype=np.float64)[q]
    return prim_to_cons(prim[..., 0], prim[..., 1], prim[..., 2], prim[..., 3])


def blast(n, seed, p_background=1e-5, energy=1.0, width_cells=2.0, noise=0.01):
    """Gaussian energy blast in a quiescent medium.

    rho = 1 + noise*xi with xi ~ U[-1, 1] per cell; u = v = 0; p = p_background.
    A total energy ``energy`` is added as a Gaussian of standard deviation
    ``width_cells`` cells centred at a seeded point in [0.3, 0.7]^2, normalised so
    that the cell sum times dx*dy equals ``energy`` exactly in real arithmetic.
    """
    rng = np.random.default_rng(seed)
    xc, yc = rng.uniform(0.3, 0.7, size=2)
    xi = rng.uniform(-1.0, 1.0, size=(n, n))
    x = cell_centres(n)
    X, Y = np.meshgrid(x, x, indexing="xy")
    dx = 1.0 / n
    sig = width_cells * dx
    g = np.exp(-((X - xc) ** 2 + (Y - yc) ** 2) / (2.0 * sig * sig))
    g *= energy / (g.sum() * dx * dx)
    rho = 1.0 + noise * xi
    U = prim_to_cons(rho, 0.0, 0.0, p_background)
    U[..., 3] += g
    return U


def shear_layer(n, seed, nmodes=8, amp=0.01):
    """Periodic shear layer with a seeded multi-mode transverse perturbation.

    rho = 2, u = +0.5 for |y-0.5| < 0.25, else rho = 1, u = -0.5; p = 2.5;
    v = amp * sum_{k=1..nmodes} sin(2 pi k x + phi_k) * exp(-((|y-0.5|-0.25)/0.05)^2)
    with phi_k ~ U[0, 2 pi).
    """
    rng = np.random.default_rng(seed)
    phi = rng.uniform(0.0, 2.0 * np.pi, size=nmodes)
    x = cell_centres(n)
    X, Y = np.meshgrid(x, x, indexing="xy")
    inner = np.abs(Y - 0.5) < 0.25
    rho = np.where(inner, 2.0, 1.0)
    u = np.where(inner, 0.5, -0.5)
    env = np.exp(-(((np.abs(Y - 0.5) - 0.25) / 0.05) ** 2))
    v = np.zeros_like(X)
    for k in range(1, nmodes + 1):
        v += np.sin(2.0 * np.pi * k * X + phi[k - 1])
    v *= amp * env
    return prim_to_cons(rho, u, v, 2.5)


GENERATORS = {"four_quadrant": four_quadrant, "blast": blast, "shear_layer": shear_layer}


def make(kind, n, seed):
    return GENERATORS[kind](n, seed)


# --- problem set-ups of the formula boundary conditions, on lf x nf grids ---
# Each returns the interior (nf, lf, 4) on [0, lx] x [0, ly]; ``seed`` adds a
# +-``noise`` relative density perturbation per cell (0 disables it).

def centres(m, length):
    d = length / m
    return np.arange(1, m + 1, dtype=np.float64) * d - d / 2


def _grid(lf, nf, lx, ly):
    return np.meshgrid(centres(lf, lx), centres(nf, ly), indexing="xy")


def _noisy(rho, seed, noise):
    if not noise:
        return rho
    return rho * (1.0 + noise * np.random.default_rng(seed).uniform(-1.0, 1.0, size=rho.shape))


def dmr(lf, nf, lx=4.0, ly=1.0, seed=0, noise=0.0):
    """Double Mach reflection: the Mach-10 shock through (1/6, 0) at 60 degrees,
    post-shock state (8, 7.1447096, -4.125, 116.5) above the line
    y = (x - 1/6) sqrt(3), (1.4, 0, 0, 1) below (the states of IC.f90:336-355)."""
    X, Y = _grid(lf, nf, lx, ly)
    post = Y > (X - 1.0 / 6) * np.sqrt(3.0)
    rho = _noisy(np.where(post, 8.0, 1.4), seed, noise)
    return prim_to_cons(rho, np.where(post, 7.1447096, 0.0), np.where(post, -4.125, 0.0), np.where(post, 116.5, 1.0))


def mach800(lf, nf, lx=1.5, ly=1.5, seed=0, noise=0.0):
    """Mach-800 jet: ambient (14, 0, 0, 1), and the jet state (1.4, 0, 800, 1)
    in the first row under the inflow window |x - lx/2| <= 0.05."""
    X, Y = _grid(lf, nf, lx, ly)
    dy = ly / nf
    jet = (np.abs(X - 0.5 * lx) <= 0.05) & (Y < dy)
    rho = np.where(jet, 1.4, _noisy(np.full(X.shape, 14.0), seed, noise))
    return prim_to_cons(rho, 0.0, np.where(jet, 800.0, 0.0), 1.0)


def rayleigh_taylor(lf, nf, lx=0.5, ly=1.5, seed=0, noise=0.0, g=0.1):
    """Rayleigh-Taylor: density 1 below y = ly/2 and 2 above, p = 2.5 + g y,
    v = 0.01 (1 + cos(2 pi (x - lx/2)/lx)) (1 + cos(2 pi (y - ly/2)/ly)) / 4."""
    X, Y = _grid(lf, nf, lx, ly)
    rho = _noisy(np.where(Y < ly / 2, 1.0, 2.0), seed, noise)
    v = 0.01 * (1 + np.cos(2 * np.pi * (X - lx / 2) / lx)) * (1 + np.cos(2 * np.pi * (Y - ly / 2) / ly)) / 4
    return prim_to_cons(rho, 0.0, v, 2.5 + g * Y)


def shock_tube_x(lf, nf, lx=1.0, ly=1.0, seed=0, noise=0.0, left=(3.857143, 2.629369, 0.0, 10.33333),
                 right=(1.0, 0.0, 0.0, 1.0), x0=0.25):
    """A planar shock in x: primitive state ``left`` for x < x0 * lx, ``right``
    beyond it, with a seeded density perturbation."""
    X, _ = _grid(lf, nf, lx, ly)
    L = X < x0 * lx
    prim = [np.where(L, a, b) for a, b in zip(left, right)]
    return prim_to_cons(_noisy(prim[0], seed, noise), prim[1], prim[2], prim[3])


def quadrants(lf, nf, lx=1.0, ly=1.0, seed=0, noise=0.0):
    """Four seeded constant states (four_quadrant_states) split at the domain centre."""
    X, Y = _grid(lf, nf, lx, ly)
    q = np.where(Y < 0.5 * ly, np.where(X < 0.5 * lx, 0, 1), np.where(X < 0.5 * lx, 2, 3))
    prim = np.array(four_quadrant_states(seed), dtype=np.float64)[q]
    return prim_to_cons(_noisy(prim[..., 0], seed, noise), prim[..., 1], prim[..., 2], prim[..., 3])


PROBLEMS = {"quadrants": quadrants, "dmr": dmr, "mach800": mach800, "rayleigh_taylor": rayleigh_taylor, "shock_tube_x": shock_tube_x}


def problem(kind, lf, nf, lx, ly, seed=0, noise=0.0):
    return PROBLEMS[kind](lf, nf, lx, ly, seed=seed, noise=noise)
