"""The reference's uniform-grid regression goldens (tests/mhd/reference/**.dc)
and the initial states that produce them.

Test infrastructure.  The .dc files under tests/golden/dc are copies of the
reference's own result files (data).  This module restates
  * the .dc v4 reader of tests/common.py:35-207 (header + 'mhd     ' block of
    8 doubles per cell: rho, m[3], E, vol B[3]),
  * the comparison rule of tests/mhd/mhd_diff.py:189-241 (relative 1e-15
    after subtracting an absolute 1e-15 floor, on rho, |v|, E and |B|),
  * the initial conditions of the golden configs
    (tests/mhd/config_files/**.json) with the face-B initialisation of
    source/mhd/initialize.hpp:361-549 (7-point Gauss-Legendre face averages,
    emulated 1-D/2-D geometry of source/grid/amr.hpp:288-310) and the fluid
    initialisation of initialize.hpp:106-174.
"""
import math
import os
import struct

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DC_DIR = os.path.join(HERE, "golden", "dc")

# 7-point Gauss-Legendre nodes and weights (textbook values)
_GL_X = (0.0, 0.405845151377397166906606412076961,
         0.741531185599394439863864773280788, 0.949107912342758524526189684047851)
_GL_W = (0.417959183673469387755102040816327, 0.381830050505118944950369775488975,
         0.279705391489276667901467771423780, 0.129484966168869693270611432679082)


def gl7(f, a, b):
    """Integral of f over [a, b] with the 7-point rule: scale * sum."""
    avg = (a + b) * 0.5
    scale = (b - a) * 0.5
    r = f(avg + scale * 0.0) * _GL_W[0]
    for x, w in zip(_GL_X[1:], _GL_W[1:]):
        fp = f(avg + scale * x)
        fm = f(avg + scale * (-x))
        r += (fp + fm) * w
    return scale * r


def read_dc(path):
    """Header fields, cell ids and the (ncells, 8) 'mhd' block of a v4 file."""
    with open(path, "rb") as f:
        raw = f.read()
    off = 0

    def take(fmt):
        nonlocal off
        v = struct.unpack_from("<" + fmt, raw, off)
        off += struct.calcsize("<" + fmt)
        return v

    meta = {}
    meta["file_version"], meta["sim_step"] = take("QQ")
    (meta["sim_time"], meta["adiabatic_index"], meta["proton_mass"],
     meta["vacuum_permeability"], meta["particle_temp_nrj_ratio"]) = take("5d")
    nvar = take("B")[0]
    offsets = take("%dQ" % nvar)
    meta["endianness"] = take("Q")[0]
    meta["ref_lvl_0_cells"] = take("3Q")
    meta["max_ref_lvl"] = take("i")[0]
    meta["neighborhood_length"] = take("I")[0]
    meta["periodicity"] = take("3B")
    meta["geometry_id"] = take("i")[0]
    meta["grid_start"] = take("3d")
    meta["lvl_0_cell_length"] = take("3d")
    ncell = take("Q")[0]
    cells = np.frombuffer(raw, dtype="<u8", count=ncell, offset=off).copy()
    var = {}
    for o in offsets:
        name = raw[o:o + 8].decode("ascii")
        var[name] = o + 8
    mhd = np.frombuffer(raw, dtype="<f8", count=8 * ncell,
                        offset=var["mhd     "]).reshape(ncell, 8).copy()
    assert meta["file_version"] == 4 and meta["endianness"] == 0x1234567890abcdef
    return meta, cells, mhd


def dc_to_grid(meta, cells, mhd):
    """Reorder the per-cell 'mhd' rows into (8, n) arrays in i-fastest order
    (level-0 dccrg ids are 1 + i + nx*(j + ny*k))."""
    nx, ny, nz = meta["ref_lvl_0_cells"]
    n = nx * ny * nz
    out = np.zeros((8, n))
    out[:, cells.astype(np.int64) - 1] = mhd.T
    return out


def mhd_diff(a, b, tol=1e-15, floor=1e-15):
    """Largest violation ratio of the mhd_diff.py rule between two (8, n)
    arrays of {rho, m, E, vol B}: returns dict of max (|x-y| - floor)/max(|x|,|y|)
    per checked quantity (the rule passes where all are <= tol) and the
    plain maximum relative difference."""
    res = {}
    rho1, rho2 = a[0], b[0]
    v1 = np.sqrt(((a[1:4] / rho1) ** 2).sum(0))
    v2 = np.sqrt(((b[1:4] / rho2) ** 2).sum(0))
    B1 = np.sqrt((a[5:8] ** 2).sum(0))
    B2 = np.sqrt((b[5:8] ** 2).sum(0))
    for name, x, y in (("mass", rho1, rho2), ("velocity", v1, v2),
                       ("energy", a[4], b[4]), ("magnetic_field", B1, B2)):
        den = np.maximum(np.abs(x), np.abs(y))
        d = np.abs(x - y)
        with np.errstate(invalid="ignore", divide="ignore"):
            viol = np.where(den != 0, (d - floor) / den, -np.inf)
            rel = np.where(den != 0, d / den, 0.0)
        res[name] = (float(viol.max()), float(rel.max()))
    return res


# --- golden configurations (tests/mhd/config_files/**.json) ---------------

def _adv_density(x, y, z):
    return 1.01 if abs(x - 0.5) < 0.25 else 1.0


def _phase2d(y, z):
    return math.pi * (math.sqrt(3) * y + z)


GOLDENS = {
    "density_advection_1d_px": dict(
        cells=(200, 1, 1), volume=(1.0, 1.0, 1.0), gamma=1.6666666667, mu0=1.0,
        mp=1.0, cfl=0.5, t_end=1.0, steps=917,
        n=_adv_density, v=lambda x, y, z: (1.0, 0.0, 0.0), p=lambda x, y, z: 1.0,
        B=lambda x, y, z: (0.0, 0.0, 0.0), bg=None),
    "density_advection_1d_mx": dict(
        cells=(200, 1, 1), volume=(1.0, 1.0, 1.0), gamma=1.6666666667, mu0=1.0,
        mp=1.0, cfl=0.5, t_end=1.0, steps=917,
        n=_adv_density, v=lambda x, y, z: (-1.0, 0.0, 0.0), p=lambda x, y, z: 1.0,
        B=lambda x, y, z: (0.0, 0.0, 0.0), bg=None),
    "alfven_1d_px": dict(
        cells=(100, 1, 1), volume=(1.0, 1.0, 1.0), gamma=1.6666666666666667,
        mu0=1.0, mp=1.0, cfl=0.5, t_end=1.0, steps=202,
        n=lambda x, y, z: 1.0,
        v=lambda x, y, z: (0.0, 0.1 * math.sin(2 * math.pi * x),
                           0.1 * math.cos(2 * math.pi * x)),
        p=lambda x, y, z: 0.1,
        B=lambda x, y, z: (1.0, 0.1 * math.sin(2 * math.pi * x),
                           0.1 * math.cos(2 * math.pi * x)), bg=None),
    "alfven_1d_mx": dict(
        cells=(100, 1, 1), volume=(1.0, 1.0, 1.0), gamma=1.6666666666666667,
        mu0=1.0, mp=1.0, cfl=0.5, t_end=1.0, steps=202,
        n=lambda x, y, z: 1.0,
        v=lambda x, y, z: (0.0, -0.1 * math.sin(2 * math.pi * x),
                           -0.1 * math.cos(2 * math.pi * x)),
        p=lambda x, y, z: 0.1,
        B=lambda x, y, z: (1.0, 0.1 * math.sin(2 * math.pi * x),
                           0.1 * math.cos(2 * math.pi * x)), bg=None),
    "alfven_2d_x_1": dict(
        cells=(1, 10, 10), volume=(1.0, 2 / math.sqrt(3), 2.0),
        gamma=1.6666666666666667, mu0=1.0, mp=1.0, cfl=0.5, t_end=1.0, steps=18,
        n=lambda x, y, z: 1.0,
        v=lambda x, y, z: (math.sin(_phase2d(y, z)) / 1000,
                           -math.cos(_phase2d(y, z)) / 2000,
                           math.sqrt(3) * math.cos(_phase2d(y, z)) / 2000),
        p=lambda x, y, z: 0.1,
        B=lambda x, y, z: (math.sin(_phase2d(y, z)) / 1000,
                           math.sqrt(3) / 2 + math.cos(_phase2d(y, z)) / 2000,
                           0.5 - math.sqrt(3) * math.cos(_phase2d(y, z)) / 2000),
        bg=None),
    "alfven_2d_x_2": dict(
        cells=(1, 10, 10), volume=(1.0, 2 / math.sqrt(3), 2.0),
        gamma=1.6666666666666667, mu0=1.0, mp=1.0, cfl=0.5, t_end=1.0, steps=33,
        n=lambda x, y, z: 1.0,
        v=lambda x, y, z: (math.sin(_phase2d(y, z)) / 1000,
                           math.sqrt(3) / 2 - math.cos(_phase2d(y, z)) / 2000,
                           0.5 + math.sqrt(3) * math.cos(_phase2d(y, z)) / 2000),
        p=lambda x, y, z: 0.1,
        B=lambda x, y, z: (math.sin(_phase2d(y, z)) / 1000,
                           math.sqrt(3) / 2 + math.cos(_phase2d(y, z)) / 2000,
                           0.5 - math.sqrt(3) * math.cos(_phase2d(y, z)) / 2000),
        bg=None),
    "alfven_2d_x_bgB1": dict(
        cells=(1, 10, 10), volume=(1.0, 2 / math.sqrt(3), 2.0),
        gamma=1.6666666666666667, mu0=1.0, mp=1.0, cfl=0.5, t_end=1.0, steps=18,
        n=lambda x, y, z: 1.0,
        v=lambda x, y, z: (math.sin(_phase2d(y, z)) / 1000,
                           -math.cos(_phase2d(y, z)) / 2000,
                           math.sqrt(3) * math.cos(_phase2d(y, z)) / 2000),
        p=lambda x, y, z: 0.1,
        B=lambda x, y, z: (math.sin(_phase2d(y, z)) / 1000,
                           math.cos(_phase2d(y, z)) / 2000,
                           -math.sqrt(3) * math.cos(_phase2d(y, z)) / 2000),
        bg=(0.0, 0.86603, 0.5)),
    "alfven_2d_x_bgB2": dict(
        cells=(1, 10, 10), volume=(1.0, 2 / math.sqrt(3), 2.0),
        gamma=1.6666666666666667, mu0=1.0, mp=1.0, cfl=0.5, t_end=1.0, steps=33,
        n=lambda x, y, z: 1.0,
        v=lambda x, y, z: (math.sin(_phase2d(y, z)) / 1000,
                           math.sqrt(3) / 2 - math.cos(_phase2d(y, z)) / 2000,
                           0.5 + math.sqrt(3) * math.cos(_phase2d(y, z)) / 2000),
        p=lambda x, y, z: 0.1,
        B=lambda x, y, z: (math.sin(_phase2d(y, z)) / 1000,
                           math.cos(_phase2d(y, z)) / 2000,
                           -math.sqrt(3) * math.cos(_phase2d(y, z)) / 2000),
        bg=(0.0, 0.86603, 0.5)),
}


def _face_average(func, comp, dim, lo, hi):
    """integrate_face_B of source/mhd/initialize.hpp:364-497 for the face on
    the negative side of a cell spanning lo..hi (emulated dims have lo == hi)."""
    sx, sy, sz = lo
    ex, ey, ez = hi
    if dim == 2:
        z = sz
        dx, dy = ex - sx, ey - sy
        if dx == 0 and dy == 0:
            return func(sx, sy, z)[comp]
        if dx == 0:
            return gl7(lambda y: func(sx, y, z)[comp], sy, ey) / dy
        if dy == 0:
            return gl7(lambda x: func(x, sy, z)[comp], sx, ex) / dx
        return gl7(lambda y: gl7(lambda x: func(x, y, z)[comp], sx, ex), sy, ey) / dx / dy
    if dim == 1:
        y = sy
        dx, dz = ex - sx, ez - sz
        if dx == 0 and dz == 0:
            return func(sx, y, sz)[comp]
        if dx == 0:
            return gl7(lambda z: func(sx, y, z)[comp], sz, ez) / dz
        if dz == 0:
            return gl7(lambda x: func(x, y, sz)[comp], sx, ex) / dx
        return gl7(lambda z: gl7(lambda x: func(x, y, z)[comp], sx, ex), sz, ez) / dx / dz
    x = sx
    dy, dz = ey - sy, ez - sz
    if dy == 0 and dz == 0:
        return func(x, sy, sz)[comp]
    if dy == 0:
        return gl7(lambda z: func(x, sy, z)[comp], sz, ez) / dz
    if dz == 0:
        return gl7(lambda y: func(x, y, sz)[comp], sy, ey) / dy
    return gl7(lambda z: gl7(lambda y: func(x, y, z)[comp], sy, ey), sz, ez) / dy / dz


def initial_state(cfg):
    """(shape, lengths, state (8, n), bg (9, n) or None) after
    initialize_magnetic_field + update_vol_B + initialize_fluid."""
    nx, ny, nz = cfg["cells"]
    n = nx * ny * nz
    length = [cfg["volume"][d] / cfg["cells"][d] for d in range(3)]
    cells = (nx, ny, nz)
    mid = [cfg["volume"][d] / 2 for d in range(3)]
    state = np.zeros((8, n))
    bg = None if cfg["bg"] is None else np.zeros((9, n))
    gamma, mu0, mp = cfg["gamma"], cfg["mu0"], cfg["mp"]

    def geom(i, j, k):
        idx = (i, j, k)
        lo, hi, ctr = [], [], []
        for d in range(3):
            if cells[d] == 1:
                lo.append(mid[d]); hi.append(mid[d]); ctr.append(mid[d])
            else:
                mn = idx[d] * length[d]
                lo.append(mn); hi.append(mn + length[d]); ctr.append(mn + length[d] / 2)
        return lo, hi, ctr

    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                c = i + nx * (j + ny * k)
                lo, hi, ctr = geom(i, j, k)
                for d in range(3):
                    state[5 + d, c] = _face_average(cfg["B"], d, d, lo, hi)
                if bg is not None:
                    # Bg_B(cell)(+dir): face on the positive side
                    for d in range(3):
                        plo = list(lo)
                        plo[d] = hi[d]
                        for comp in range(3):
                            bg[3 * d + comp, c] = _face_average(
                                lambda x, y, z: cfg["bg"], comp, d, plo, hi)
    # update_vol_B(constant_p = false), then initialize_fluid
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                c = i + nx * (j + ny * k)
                cp = ((i + 1) % nx) + nx * (j + ny * k)
                cq = i + nx * (((j + 1) % ny) + ny * k)
                cr = i + nx * (j + ny * ((k + 1) % nz))
                vb = (0.5 * (state[5, c] + state[5, cp]),
                      0.5 * (state[6, c] + state[6, cq]),
                      0.5 * (state[7, c] + state[7, cr]))
                _, _, ctr = geom(i, j, k)
                rho = mp * cfg["n"](*ctr)
                v = cfg["v"](*ctr)
                p = cfg["p"](*ctr)
                state[0, c] = rho
                for q in range(3):
                    state[1 + q, c] = v[q] * rho
                kin = 0.5 * rho * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2])
                mag = 0.5 / mu0 * (vb[0] * vb[0] + vb[1] * vb[1] + vb[2] * vb[2])
                state[4, c] = p / (gamma - 1) + kin + mag
    return cells, length, state, bg


def volume_field(cells, state):
    """(3, n) cell-centred B = average of the two faces."""
    nx, ny, nz = cells
    f = state[5:8].reshape(3, nz, ny, nx)
    out = np.empty_like(f)
    out[0] = 0.5 * (f[0] + np.roll(f[0], -1, axis=2))
    out[1] = 0.5 * (f[1] + np.roll(f[1], -1, axis=1))
    out[2] = 0.5 * (f[2] + np.roll(f[2], -1, axis=0))
    return out.reshape(3, -1)
