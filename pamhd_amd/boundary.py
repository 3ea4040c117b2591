"""Cell and face classification of uniform grids with value and copy
boundaries, and the per-step values of time-dependent value boundaries.

Restates, for one refinement level:
* Boundaries::classify (source/boundaries/boundaries.hpp:100-226), per
  variable (number density, velocity, pressure, magnetic field): cells in a
  value boundary's geometry are value cells, cells in a copy boundary's
  geometry copy cells (unless already value cells); boundary cells without
  a normal neighbour, and copy cells without a normal neighbour to copy
  from, are not solved;
* set_solver_info (source/mhd/boundaries.hpp:67-245): cell type 1 normal,
  0 boundary (a value or copy cell of any variable), -1 not solved (the
  variables must agree on these);
* classify_faces (:1039-1217): face types 1 normal, 0 boundary, -1 not
  solved, for the faces -x, +x, -y, +y, -z, +z of every cell;
* Value_Boundary::get_data (source/boundaries/value_boundary.hpp:60-340):
  the value of the time stamp nearest to t (the earlier one on a tie).

Neighbours are those of a neighbourhood of length 1, offsets -1..1 with z
slowest and x fastest; periodic dimensions wrap (a one-cell periodic
dimension makes a cell its own neighbour), closed dimensions end.
"""
import bisect

import numpy as np

VARIABLES = ("number-density", "velocity", "pressure", "magnetic-field")


def offsets():
    """the 26 offsets in neighbourhood order with their face direction
    (+-1..3, 0 if none) and edge description (dimension, sign in the lower
    other dimension, sign in the higher one; dimension -1 if none)"""
    out = []
    for oz in (-1, 0, 1):
        for oy in (-1, 0, 1):
            for ox in (-1, 0, 1):
                o = (ox, oy, oz)
                nz = sum(1 for v in o if v)
                if nz == 0:
                    continue
                fn, en = 0, (-1, 0, 0)
                if nz == 1:
                    d = [a for a in range(3) if o[a]][0]
                    fn = o[d] * (d + 1)
                elif nz == 2:
                    e = [a for a in range(3) if not o[a]][0]
                    d1, d2 = [a for a in range(3) if a != e]
                    en = (e, o[d1], o[d2])
                out.append((o, fn, en))
    return out


def neighbour_index(cells, periodic):
    """(26, n) index of each cell's neighbour at each offset, -1 where a
    closed dimension ends"""
    nx, ny, nz = cells
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    idx = [i.reshape(-1), j.reshape(-1), k.reshape(-1)]
    N = (nx, ny, nz)
    out = []
    for o, _, _ in offsets():
        p = []
        ok = np.ones(idx[0].shape, dtype=bool)
        for d in range(3):
            v = idx[d] + o[d]
            if periodic[d]:
                v = v % N[d]
            else:
                ok &= (v >= 0) & (v < N[d])
                v = np.clip(v, 0, N[d] - 1)
            p.append(v)
        out.append(np.where(ok, p[0] + nx * (p[1] + ny * p[2]), -1))
    return np.stack(out)


def _fidx(direction):
    return 2 * (abs(direction) - 1) + (1 if direction > 0 else 0)


class ClassificationError(ValueError):
    pass


def classify(cells, periodic, in_value, in_copy):
    """(vtype (4, n), ctype (n,), finfo (6, n)) from the (4, n) masks of
    cells in value / copy boundary geometries of each variable.  vtype: 1
    normal, 2 not solved, 3 value cell, 4 copy cell"""
    nb = neighbour_index(cells, periodic)
    offs = offsets()
    n = nb.shape[1]
    me = np.arange(n)
    vtype = np.ones((4, n), dtype=np.int8)
    for v in range(4):
        t = vtype[v]
        t[np.asarray(in_copy[v], dtype=bool)] = 4
        t[np.asarray(in_value[v], dtype=bool)] = 3
        normal = t == 1
        have = np.zeros(n, dtype=bool)
        sources = np.zeros(n, dtype=bool)
        for q in range(len(offs)):
            b = nb[q]
            ok = b >= 0
            nrm = ok & normal[np.where(ok, b, 0)]
            have |= nrm & (b != me)
            sources |= nrm
        t[(t != 1) & ~have] = 2
        t[(t == 4) & ~sources] = 2
    ctype = np.ones(n, dtype=np.int8)
    ctype[((vtype == 3) | (vtype == 4)).any(0)] = 0
    ds = vtype == 2
    if not (ds == ds[0]).all():
        raise ClassificationError("the variables' not-solved cells differ")
    ctype[ds[0]] = -1
    # classify_faces
    F = np.full((6, n), -99, dtype=np.int16)
    F[:, ctype == 1] = 1
    non_normal = ctype != 1
    for q, (o, fn, en) in enumerate(offs):
        b = nb[q]
        ok = non_normal & (b >= 0)
        ok &= ctype[np.where(b >= 0, b, 0)] >= 1
        if fn:
            for direction in (-3, -2, -1, 1, 2, 3):
                if direction == -fn:
                    continue
                F[_fidx(direction), ok] = 1
        if en[0] >= 0:
            e = en[0]
            d1, d2 = [a for a in range(3) if a != e]
            F[_fidx(-(d1 + 1) if en[1] < 0 else d1 + 1), ok] = 1
            F[_fidx(-(d2 + 1) if en[2] < 0 else d2 + 1), ok] = 1
    has_normal = (F == 1).any(0)
    F[:, has_normal] = np.where(F[:, has_normal] < 0, 0, F[:, has_normal])
    G = F.copy()
    for q, (o, fn, en) in enumerate(offs):
        b = nb[q]
        ok = b >= 0
        bb = np.where(ok, b, 0)
        if fn:
            for direction in (-3, -2, -1, 1, 2, 3):
                f = _fidx(direction)
                nrm = ok & (F[f, bb] >= 1)
                if direction == -fn:
                    bad = nrm & (F[_fidx(fn), me] < 1) & (G[_fidx(fn), me] < 1)
                    if bad.any():
                        raise ClassificationError("inconsistent face types at cell %d"
                                                  % int(np.nonzero(bad)[0][0]))
                G[f] = np.where(nrm & (G[f] < 0), 0, G[f])
        if en[0] >= 0:
            e = en[0]
            d1, d2 = [a for a in range(3) if a != e]
            s1 = -1 if en[1] < 0 else 1
            s2 = -1 if en[2] < 0 else 1
            hit = ok & ((F[_fidx(-s1 * (d1 + 1)), bb] == 1) | (F[_fidx(-s2 * (d2 + 1)), bb] == 1))
            for f in (_fidx(s1 * (d1 + 1)), _fidx(s2 * (d2 + 1))):
                G[f] = np.where(hit & (G[f] < 0), 0, G[f])
    G[G < -1] = -1
    return vtype, ctype, G.astype(np.int8)


def nearest_stamp(stamps, t):
    """index of the time stamp nearest to t, the earlier one on a tie"""
    k = bisect.bisect_left(stamps, t)
    if k == 0:
        return 0
    if k == len(stamps):
        return len(stamps) - 1
    return k - 1 if abs(t - stamps[k - 1]) <= abs(t - stamps[k]) else k
