"""A PyTorch restatement of the periodic MHD step with HLL fluxes and no
background field (the forward map the device adjoint differentiates), used
by the tests as a gradient checker: torch.autograd of this function gives
the exact vector-Jacobian product of one step, against which the
hand-written adjoint is compared component by component.

Test infrastructure only.  It follows the same formulas as the oracle's
step (oracle/pamhd_oracle.c: or_step_periodic with or_flux_hll), not its
operation order, so values agree to rounding, not bit for bit.
"""
import torch


def _roll(a, shift, d):
    """roll along grid dimension d (0 = x, the last tensor axis)"""
    return torch.roll(a, shifts=shift, dims=a.dim() - 1 - d)


def _smax(a, b):
    return torch.where(a < b, b, a)


def _smin(a, b):
    return torch.where(b < a, b, a)


def _hll(L, R, gamma, mu0):
    """source/mhd/hll_athena.hpp:57-275 without background field; L, R are
    lists of 8 tensors in the face frame"""
    half_inv = 0.5 / mu0
    inv = 1.0 / mu0
    sides = []
    for u in (L, R):
        rho, m, E, B = u[0], u[1:4], u[4], u[5:8]
        kin = 0.5 / rho * (m[0] * m[0] + m[1] * m[1] + m[2] * m[2])
        mag = half_inv * (B[0] * B[0] + B[1] * B[1] + B[2] * B[2])
        p = (E - kin - mag) * (gamma - 1)
        ir = 1.0 / rho
        v = [m[0] * ir, m[1] * ir, m[2] * ir]
        bmag = torch.sqrt(B[0] * B[0] + B[1] * B[1] + B[2] * B[2])
        sound = torch.sqrt(gamma * p / rho)
        alfven = bmag / torch.sqrt(mu0 * rho)
        s2, a2 = sound * sound, alfven * alfven
        ratio = B[0] / bmag
        to_sqrt = s2 * s2 + a2 * a2 + 2 * s2 * a2 * (1 - 2 * ratio * ratio)
        cf = torch.sqrt(0.5 * (s2 + a2 + torch.sqrt(to_sqrt)))
        pB = 0.5 * inv * (B[0] * B[0] + B[1] * B[1] + B[2] * B[2])
        F = [m[0],
             m[0] * v[0] - B[0] * B[0] * inv + p + pB,
             m[1] * v[0] - B[1] * B[0] * inv,
             m[2] * v[0] - B[2] * B[0] * inv,
             v[0] * (E + p + pB) - B[0] * (v[0] * B[0] + v[1] * B[1] + v[2] * B[2]) * inv,
             torch.zeros_like(rho),
             B[1] * v[0] - v[1] * B[0],
             B[2] * v[0] - v[2] * B[0]]
        sides.append((u, v, cf, F))
    (uL, vL, cfL, FL), (uR, vR, cfR, FR) = sides
    ms = _smax(cfL, cfR)
    lo = vL[0] <= vR[0]
    msn = torch.where(lo, vL[0] - ms, vR[0] - ms)
    msp = torch.where(lo, vR[0] + ms, vL[0] + ms)
    bm = _smin(msn, torch.zeros_like(msn))
    bp = _smax(msp, torch.zeros_like(msp))
    den, s = bp - bm, bp + bm
    out = []
    for q in range(8):
        fl = FL[q] - uL[q] * bm
        fr = FR[q] - uR[q] * bp
        if q == 5:
            fl = fr = torch.zeros_like(fl)
        out.append(0.5 * (fl + fr) + 0.5 * (fl - fr) * s / den)
    out[5] = torch.zeros_like(out[5])
    return out


def step(U, dt, length, gamma, mu0, refits=3):
    """One step of length dt; U is (8, nz, ny, nx): rho, m (3), E, face B
    (-x, -y, -z faces of each cell)."""
    rho, m, E, Fb = U[0], U[1:4], U[4], U[5:8]
    vb = torch.stack([0.5 * (Fb[d] + _roll(Fb[d], -1, d)) for d in range(3)])
    cell = torch.cat([rho[None], m, E[None], vb])
    P = []
    for d in range(3):
        rot = [d, (d + 1) % 3, (d + 2) % 3]
        Rs = [cell[0], cell[1 + rot[0]], cell[1 + rot[1]], cell[1 + rot[2]], cell[4],
              cell[5 + rot[0]], cell[5 + rot[1]], cell[5 + rot[2]]]
        Ls = [_roll(x, 1, d) for x in Rs]           # L = c - e_d, R = c
        F = _hll(Ls, Rs, gamma, mu0)
        g = [None] * 8
        g[0], g[4] = F[0], F[4]
        for q in range(3):
            g[1 + rot[q]] = F[1 + q]
            g[5 + rot[q]] = F[5 + q]
        P.append([dt * x for x in g])
    dx, dy, dz = length
    sh = lambda a, ox, oy, oz: _roll(_roll(_roll(a, -ox, 0), -oy, 1), -oz, 2)
    EZX = lambda o: sh(P[0][6], *o) * dz / 4
    EYX = lambda o: -(sh(P[0][7], *o) * dy / 4)
    EXY = lambda o: sh(P[1][7], *o) * dx / 4
    EZY = lambda o: -(sh(P[1][5], *o) * dz / 4)
    EXZ = lambda o: -(sh(P[2][6], *o) * dx / 4)
    EYZ = lambda o: sh(P[2][5], *o) * dy / 4
    dBx = (EZX((0, 1, 0)) - EZX((0, -1, 0)) - EYX((0, 0, 1)) + EYX((0, 0, -1))
           + EZY((0, 1, 0)) + EZY((-1, 1, 0)) - EZY((0, 0, 0)) - EZY((-1, 0, 0))
           - EYZ((0, 0, 1)) - EYZ((-1, 0, 1)) + EYZ((0, 0, 0)) + EYZ((-1, 0, 0)))
    dBy = (EXY((0, 0, 1)) - EXY((0, 0, -1)) - EZY((1, 0, 0)) + EZY((-1, 0, 0))
           + EXZ((0, 0, 1)) + EXZ((0, -1, 1)) - EXZ((0, 0, 0)) - EXZ((0, -1, 0))
           - EZX((1, 0, 0)) + EZX((0, 0, 0)) - EZX((1, -1, 0)) + EZX((0, -1, 0)))
    dBz = (EYZ((1, 0, 0)) - EYZ((-1, 0, 0)) - EXZ((0, 1, 0)) + EXZ((0, -1, 0))
           + EYX((1, 0, 0)) + EYX((1, 0, -1)) - EYX((0, 0, 0)) - EYX((0, 0, -1))
           - EXY((0, 1, 0)) - EXY((0, 1, -1)) + EXY((0, 0, 0)) + EXY((0, 0, -1)))
    Fn = torch.stack([Fb[0] + dBx / (dy * dz), Fb[1] + dBy / (dx * dz), Fb[2] + dBz / (dx * dy)])
    new = []
    for q in range(8):
        acc = cell[q]
        for d in range(3):
            acc = acc + (P[d][q] - _roll(P[d][q], -1, d)) / length[d]
        new.append(acc)
    rho1, m1, E1, vb_flux = new[0], torch.stack(new[1:4]), new[4], torch.stack(new[5:8])
    vb_new = torch.stack([0.5 * (Fn[d] + _roll(Fn[d], -1, d)) for d in range(3)])
    for r in range(refits):
        vold = vb_flux if r == 0 else vb_new
        kin = 0.5 / rho1 * (m1 * m1).sum(0)
        p = (E1 - kin - 0.5 / mu0 * (vold * vold).sum(0)) * (gamma - 1)
        v = m1 / rho1
        E1 = p / (gamma - 1) + 0.5 * rho1 * (v * v).sum(0) + 0.5 / mu0 * (vb_new * vb_new).sum(0)
    return torch.cat([rho1[None], m1, E1[None], Fn])
