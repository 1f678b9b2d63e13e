"""FP64 PyTorch re-expression of the frozen-mask SSP-RK3 step (test infrastructure).

The differentiable form of the step has no counterpart in the reference; its
checker is this plain PyTorch restatement of the same arithmetic, run through
torch.autograd on the CPU, plus central finite differences.  The restatement
follows the reference step with the order of every RK stage given instead of
found by MOOD:

  * ghost cells: Boundary_C, BC.f90:18-59 (Neumann, periodic) and :89-135
    (reflective: mirror, normal momentum negated);
  * reconstruction: mod_GP_MOOD.f90:46-65, Uh = dot_product(zT_sp(ord, :, face, j), q)
    with ord = min(CellGPO(self), CellGPO(neighbour)) (ghost cells at Mord);
  * HLLC: mod_HLLC.f90:11-169 with c2p_local and Flux (physics.f90:9-75);
  * quadrature and update: mod_FE.f90:75-145;
  * SSP-RK3: mod_SSP_RK3.f90:12-39.

Only tests import this module.  It is pinned against the C oracle
(gpm_oracle.c): the HLLC flux pointwise and whole RK stages with the
oracle's own final masks.
"""
import numpy as np
import torch

GAMMA = 1.4
SX = [0, -1, 0, 1, 0, -1, -1, 1, 1, 0, -2, 0, 2, -3, -2, -1, 0, 1, 2, 3, 2, 1, 0, -1, -2]
SY = [0, 0, 1, 0, -1, -1, 1, 1, -1, -2, 0, 2, 0, 0, 1, 2, 3, 2, 1, 0, -1, -2, -3, -2, -1]
SZ = {1: 1, 3: 5, 5: 13, 7: 25}
IL, IT, IR, IB = 0, 1, 2, 3
NG = 4


def _fmin(a, b):
    return torch.where(a < b, a, b)


def _fmax(a, b):
    return torch.where(a > b, a, b)


def hllc_x(ul, ur):
    """x-direction HLLC flux on [..., 4] state tensors (admissible states)."""
    rL, mxL, myL, EL = ul.unbind(-1)
    rR, mxR, myR, ER = ur.unbind(-1)
    vxl, vyl = mxL / rL, myL / rL
    vxr, vyr = mxR / rR, myR / rR
    pl = (GAMMA - 1.0) * (EL - (0.5 * (mxL * mxL + myL * myL)) / rL)
    pr = (GAMMA - 1.0) * (ER - (0.5 * (mxR * mxR + myR * myR)) / rR)
    cl = torch.sqrt(GAMMA * pl / rL)
    cr = torch.sqrt(GAMMA * pr / rR)
    Fl = torch.stack([mxL, rL * (vxl * vxl) + pl, myL * vxl, vxl * (EL + pl)], -1)
    Fr = torch.stack([mxR, rR * (vxr * vxr) + pr, myR * vxr, vxr * (ER + pr)], -1)
    SL = _fmin(vxl - cl, vxr - cr)
    SR = _fmax(vxl + cl, vxr + cr)
    nL, nR = SL - vxl, SR - vxr
    uS = ((pr - pl) + (rL * vxl * nL - rR * vxr * nR)) / (rL * nL - rR * nR)
    one = torch.ones_like(rL)
    dL = rL * nL / (SL - uS)
    UsL = torch.stack([one, uS, vyl, EL / rL + (uS - vxl) * (uS + pl / (rL * nL))], -1)
    FsL = Fl + SL[..., None] * (UsL * dL[..., None] - ul)
    dR = rR * nR / (SR - uS)
    UsR = torch.stack([one, uS, vyr, ER / rR + (uS - vxr) * (uS + pr / (rR * nR))], -1)
    FsR = Fr + SR[..., None] * (UsR * dR[..., None] - ur)
    b1 = (SL >= 0)[..., None]
    b2 = ((SL < 0) & (uS >= 0))[..., None]
    b3 = ((uS < 0) & (SR >= 0))[..., None]
    return torch.where(b1, Fl, torch.where(b2, FsL, torch.where(b3, FsR, Fr)))


def _swap(u):
    return u[..., [0, 2, 1, 3]]


def hllc(ul, ur, direction):
    """HLLC for x-faces (0) or y-faces (1: the x-solver on momentum-swapped states)."""
    if direction == 0:
        return hllc_x(ul, ur)
    return _swap(hllc_x(_swap(ul), _swap(ur)))


def _bc_index(n, bc):
    """Source index and normal-momentum flip of positions -NG..n+NG-1."""
    idx, flip = [], []
    for i in range(-NG, n + NG):
        f = False
        if i < 0:
            s = 0 if bc == "neumann" else (i + n if bc == "periodic" else -1 - i)
            f = bc == "reflective"
        elif i >= n:
            s = n - 1 if bc == "neumann" else (i - n if bc == "periodic" else 2 * n - 1 - i)
            f = bc == "reflective"
        else:
            s = i
        idx.append(s)
        flip.append(f)
    return torch.tensor(idx), torch.tensor(flip)


def ghosted(U, bc):
    """Interior [n, n, 4] (index [j, i]) -> [n+8, n+8, 4] with the ghost rule."""
    n = U.shape[0]
    idx, flip = _bc_index(n, bc)
    G = U.index_select(0, idx).index_select(1, idx)
    sx = torch.where(flip, -1.0, 1.0).to(U.dtype)
    sign = torch.ones(n + 2 * NG, n + 2 * NG, 4, dtype=U.dtype)
    sign[:, :, 1] = sx[None, :]
    sign[:, :, 2] = sx[:, None]
    return G * sign


class Tables:
    """zT_sp, gauss_weight from the reference layouts (flat Fortran order)."""

    def __init__(self, t, mord, ngp):
        self.mord, self.ngp = mord, ngp
        sz = SZ[mord]
        zT = np.asarray(t["zT"], dtype=np.float64)
        self.z = {}
        for o in range(1, mord + 1, 2):
            for f in range(4):
                for g in range(ngp):
                    self.z[o, f, g] = [float(zT[(o - 1) + mord * (s + sz * (f + 4 * g))]) for s in range(SZ[o])]
        gw = np.asarray(t["gw"], dtype=np.float64)
        self.gw = [float(gw[(ngp - 1) + 5 * g]) for g in range(ngp)]


def _recon(Ug, T, ords, face, g, ci, cj):
    """Reconstruction at (face, point g) of cells with ghosted indices (ci, cj)
    (broadcast LongTensors) and per-face orders `ords` (same shape)."""
    out = torch.zeros(ci.shape + (4,), dtype=Ug.dtype)
    for o in range(1, T.mord + 1, 2):
        w = T.z[o, face, g]
        acc = torch.zeros_like(out)
        for s in range(SZ[o]):
            acc = acc + w[s] * Ug[cj + SY[s], ci + SX[s]]
        out = torch.where((ords == o)[..., None], acc, out)
    return out


def stage(V, masks, dtdx, dtdy, T, bc):
    """Forward Euler stage with frozen orders: V + (flux differences), interior [n, n, 4]."""
    n = V.shape[0]
    Ug = ghosted(V, bc)
    M = torch.full((n + 2 * NG, n + 2 * NG), T.mord, dtype=torch.int64)
    M[NG:NG + n, NG:NG + n] = torch.as_tensor(masks, dtype=torch.int64)
    j = torch.arange(NG, NG + n)
    # x-faces between ghosted columns a = NG-1 .. NG+n-1 and a+1, rows NG..NG+n-1
    a = torch.arange(NG - 1, NG + n)
    AI, AJ = torch.meshgrid(a, j, indexing="xy")
    ox = torch.minimum(M[AJ, AI], M[AJ, AI + 1])
    Fx = torch.zeros(AI.shape + (4,), dtype=V.dtype)
    for g in range(T.ngp):
        ua = _recon(Ug, T, ox, IR, g, AI, AJ)
        ub = _recon(Ug, T, ox, IL, g, AI + 1, AJ)
        Fx = Fx + hllc(ua, ub, 0) * T.gw[g]
    # y-faces between ghosted rows b = NG-1 .. NG+n-1 and b+1
    b = torch.arange(NG - 1, NG + n)
    BI, BJ = torch.meshgrid(j, b, indexing="xy")
    oy = torch.minimum(M[BJ, BI], M[BJ + 1, BI])
    Fy = torch.zeros(BI.shape + (4,), dtype=V.dtype)
    for g in range(T.ngp):
        ua = _recon(Ug, T, oy, IT, g, BI, BJ)
        ub = _recon(Ug, T, oy, IB, g, BI, BJ + 1)
        Fy = Fy + hllc(ua, ub, 1) * T.gw[g]
    # Fx[j, k]: face between cells k-1 and k of row j; Fy[k, i]: between rows k-1 and k
    fxl, fxr = Fx[:, :-1], Fx[:, 1:]
    fyb, fyt = Fy[:-1, :], Fy[1:, :]
    W = V + dtdy * fyb
    W = W + dtdx * fxl
    W = W - dtdx * fxr
    W = W - dtdy * fyt
    return W


def ssprk3(U, masks, dt, dx, dy, T, bc):
    """One SSP-RK3 step with frozen per-stage orders masks[3][n][n]."""
    dtdx, dtdy = dt / dx, dt / dy
    U1 = stage(U, masks[0], dtdx, dtdy, T, bc)
    U2 = 0.75 * U + 0.25 * stage(U1, masks[1], dtdx, dtdy, T, bc)
    return (1.0 / 3) * U + (2.0 / 3) * stage(U2, masks[2], dtdx, dtdy, T, bc)


def run(U, masks, dts, dx, dy, T, bc):
    """K frozen steps; masks [K][3][n][n], dts [K]."""
    for k in range(len(dts)):
        U = ssprk3(U, masks[k], float(dts[k]), dx, dy, T, bc)
    return U


def kinetic(U, i0, j0, w):
    """J = sum over cells [j0, j0+w) x [i0, i0+w) of rho (u^2 + v^2)."""
    S = U[j0:j0 + w, i0:i0 + w]
    return ((S[..., 1] * S[..., 1] + S[..., 2] * S[..., 2]) / S[..., 0]).sum()
