"""Pins the oracle's solar-wind box (tests/mhd/solar_wind_box.cpp restated for
uniform grids) before it is trusted as the checker of the device path.

No reference output exists for this program (it needs dccrg, Boost and
MPI), so the checks come from what the reference's code fixes:
  * prepare_grid's classification (source/grid/solar_wind_box.hpp:262-406):
    counts of boundary, inflow, planet and unsolved cells in closed form, and
    every copy rule pointing at a normal cell across the right walls;
  * a uniform flow with a field across it is preserved exactly by the whole
    step (fluxes, constrained transport, refits, outflow copies, inflow);
  * the inflow face rule (source/mhd/solar_wind_box.hpp:275-299);
  * the face divergence of normal cells away from the boundary stays at
    roundoff (constrained transport signs);
  * the background field: point-dipole closed form and its min-distance
    clamp (source/background_magnetic_field.hpp:466-491).
"""
import math

import numpy as np
import pytest

import oracle_lib as O
from pamhd_amd import configs


def _box(edge=12, **kw):
    sb = configs.solar_wind_box(edge, 5e6, 5e5, 3e-9, -4e-9)
    for k, v in kw.items():
        setattr(sb, k, v)
    return sb


@pytest.mark.parametrize("cells,direction", [((12, 12, 12), "+x"), ((10, 8, 6), "-y"),
                                             ((6, 7, 9), "+z")])
def test_classification_counts(cells, direction):
    sb = _box(cells=cells, direction=direction)
    h = sb.cell_length[0]
    sb.start = tuple(-c * h / 2 for c in cells)
    b = O.Box(sb)
    nx, ny, nz = cells
    n = nx * ny * nz
    interior = (nx - 2) * (ny - 2) * (nz - 2)
    dim = "xyz".index(direction[1])
    wall = n // cells[dim]
    planet = (b.rule == O.RULE_PLANET).sum()
    assert (b.rule == O.RULE_SW).sum() == wall
    assert (b.ctype == 0).sum() == n - interior + planet
    assert (b.ctype == 1).sum() == interior - planet
    # the inflow wall is the +/- side named by the direction
    idx = np.unravel_index(np.flatnonzero(b.rule == O.RULE_SW), (nz, ny, nx))[2 - dim]
    assert np.all(idx == (cells[dim] - 1 if direction[0] == "+" else 0))
    # copy rules: offsets point inwards, at a normal cell
    k, j, i = np.unravel_index(np.arange(n), (nz, ny, nx))
    for c in np.flatnonzero((b.rule >= 1) & (b.rule <= 27)):
        r = int(b.rule[c]) - 1
        o = (r % 3 - 1, (r // 3) % 3 - 1, r // 9 - 1)
        pos = (i[c], j[c], k[c])
        for d in range(3):
            want = 1 if pos[d] == 0 else (-1 if pos[d] == cells[d] - 1 else 0)
            assert o[d] == want
        src = (i[c] + o[0]) + nx * ((j[c] + o[1]) + ny * (k[c] + o[2]))
        assert b.ctype[src] == 1
    # 8 vertex cells (3 offsets), 12 edges, 6 faces, minus the inflow wall
    nnz = np.array([sum(1 for d in range(3) if ((int(r) - 1) // 3 ** d) % 3 != 1)
                    for r in b.rule[(b.rule >= 1) & (b.rule <= 27)]])
    assert (nnz == 3).sum() == 4


def test_planet_cells_and_unsolved_interior():
    """Inner radius 2.2 cells around the box centre: cells with a corner
    inside are boundary cells; those without a face or edge neighbour of
    type 1 become type -1."""
    sb = _box(edge=16)
    h = sb.cell_length[0]
    sb.inner_radius = 2.2 * h
    b = O.Box(sb)
    nx = 16
    k, j, i = np.unravel_index(np.arange(nx ** 3), (nx, nx, nx))
    lo = np.stack([i, j, k]) * h + sb.start[0]
    corners = []
    for q in range(8):
        c = lo + np.array([[(q >> 2) & 1], [(q >> 1) & 1], [q & 1]]) * h
        corners.append(np.sqrt((c ** 2).sum(0)))
    inside = (np.min(corners, axis=0) <= sb.inner_radius)
    planet_or_unsolved = (b.rule == O.RULE_PLANET) | (b.ctype == -1)
    np.testing.assert_array_equal(planet_or_unsolved, inside)
    assert (b.ctype == -1).sum() > 0
    for c in np.flatnonzero(b.ctype == -1):
        for dz in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    if (dx != 0) + (dy != 0) + (dz != 0) in (1, 2):
                        nb = (i[c] + dx) + nx * ((j[c] + dy) + nx * (k[c] + dz))
                        assert b.ctype[nb] != 1


def _uniform_state(b, rho, v, p, B, gamma, mu0):
    n = b.n
    s = np.zeros((14, n))
    s[0] = rho
    for d in range(3):
        s[1 + d] = rho * v[d]
        s[5 + d] = B[d]
        s[8 + 2 * d] = B[d]
        s[9 + 2 * d] = B[d]
    kin = 0.5 * rho * sum(x * x for x in v)
    mag = 0.5 / mu0 * sum(x * x for x in B)
    s[4] = p / (gamma - 1) + kin + mag
    return s


@pytest.mark.parametrize("sid", [O.RUSANOV, O.HLL, O.HLLD, O.ROE])
def test_uniform_flow_is_preserved(sid):
    """Box away from the origin (no planet cells), no background field, the
    whole grid in the solar-wind state: steps with boundaries keep it, up to
    the energy refits' rounding."""
    sb = _box(edge=10, start=(1e9, 1e9, 1e9), dipoles=[])
    b = O.Box(sb)
    assert (b.rule == O.RULE_PLANET).sum() == 0
    s = _uniform_state(b, sb.number_density * sb.proton_mass, sb.velocity, sb.pressure,
                       sb.magnetic_field, sb.adiabatic_index, sb.vacuum_permeability)
    ref = s.copy()
    dt = 0.3 * sb.cell_length[0] / 6e5
    for _ in range(5):
        st, sm, zc = b.step(sid, dt, s)
        assert st == 0
        assert b.boundaries(s) == 0
        b.sync(s)
    scale = np.abs(ref).max(1, keepdims=True) + 1e-300
    assert np.abs(s - ref).max() <= 1e-13 * scale.max()
    for q in range(14):
        assert np.abs(s[q] - ref[q]).max() <= 1e-13 * scale[q, 0], q


def test_inflow_face_rule():
    sb = _box(edge=10)
    b = O.Box(sb)
    s = b.init()
    rng = np.random.default_rng(4)
    s[8:] += rng.normal(0, 1e-9, size=(6, b.n))
    s[5:8] += rng.normal(0, 1e-9, size=(3, b.n))
    before = s.copy()
    assert b.boundaries(s) == 0
    nx = 10
    sw = np.flatnonzero(b.rule == O.RULE_SW)
    nb = sw - 1          # "+x" inflow: interior neighbour at -x
    # face toward the inflow cell set to the solar-wind B_x (0); Vol_B x refit
    assert np.all(s[9, nb] == sb.magnetic_field[0])
    np.testing.assert_array_equal(s[5, nb], s[9, nb] / 2 + s[8, nb] / 2)
    # inflow cells hold the solar-wind state on all their copies
    for d in range(3):
        assert np.all(s[8 + 2 * d, sw] == sb.magnetic_field[d])
        assert np.all(s[9 + 2 * d, sw] == sb.magnetic_field[d])
    # normal cells that are not inflow neighbours are untouched
    normal = np.flatnonzero(b.ctype == 1)
    other = np.setdiff1d(normal, nb)
    np.testing.assert_array_equal(s[:, other], before[:, other])
    _ = nx


def test_divergence_stays_at_roundoff_inside():
    sb = _box(edge=14)
    b = O.Box(sb)
    s = b.init()
    st, steps, t = b.run(O.HLL, 0.3, 0.0, 1e30, s, max_steps=15)
    assert st == 0 and steps == 15
    n = 14
    h = sb.cell_length
    div = ((s[9] - s[8]) / h[0] + (s[11] - s[10]) / h[1] + (s[13] - s[12]) / h[2]).reshape(n, n, n)
    ct = b.ctype.reshape(n, n, n)
    deep = np.zeros_like(ct, dtype=bool)
    deep[2:-2, 2:-2, 2:-2] = True
    for ax in range(3):
        for sh in (-1, 1):
            deep &= np.roll(ct, sh, ax) == 1
    deep &= ct == 1
    Bscale = np.abs(s[8:]).max() / min(h)
    assert np.abs(div[deep]).max() <= 1e-12 * Bscale
    # the flow has changed the field by much more than that
    assert np.abs(s[8:] - b.init()[8:]).max() > 1e-6 * np.abs(s[8:]).max()


def test_dipole_background_closed_form():
    """One dipole at the origin, no constant field: Bg_B on the +x face of a
    cell equals mu0/(4 pi) (3 (m.r) r / r^5 - m / r^3) at the face centre,
    and mu0 m / (2 pi d^3) inside min-distance d."""
    sb = _box(edge=8)
    m = np.array([1e16, -2e16, -5e16])
    sb.dipoles = [(tuple(m), (0.0, 0.0, 0.0))]
    sb.bg_value = (0.0, 0.0, 0.0)
    h = sb.cell_length[0]
    sb.bg_min_distance = 1.2 * h
    b = O.Box(sb)
    mu0 = sb.vacuum_permeability
    for (i, j, k) in [(0, 0, 0), (5, 2, 7), (3, 3, 3), (4, 3, 4)]:
        c = i + 8 * (j + 8 * k)
        r = np.array([sb.start[0] + (i + 1) * h, sb.start[1] + (j + 0.5) * h,
                      sb.start[2] + (k + 0.5) * h])
        rr = math.sqrt((r ** 2).sum())
        if rr < sb.bg_min_distance:
            want = mu0 * m / (2 * math.pi * sb.bg_min_distance ** 3)
        else:
            want = mu0 / (4 * math.pi) * (3 * (m @ r) * r / rr ** 5 - m / rr ** 3)
        np.testing.assert_allclose(b.bg[0:3, c], want, rtol=1e-13, atol=1e-30)


@pytest.mark.parametrize("sid", [O.RUSANOV, O.HLL])
def test_box_step_matches_literal_face_loop(sid):
    """rho and m of normal cells after one box step against a literal loop
    over faces between cells of type >= 0 (get_fluxes,
    source/mhd/solve.hpp:262-393): L = cell, R = its +d neighbour,
    bg = Bg_B(L)(+d), through the batched oracle solver; faces next to
    unsolved (-1) cells or outside the box carry no flux."""
    sb = _box(edge=12)
    h = sb.cell_length[0]
    sb.inner_radius = 2.2 * h
    sb.dipoles = [((3e18, -1e18, -9e18), (0.3 * h, -0.2 * h, 0.1 * h))]
    sb.bg_min_distance = 0.7 * h
    b = O.Box(sb)
    assert (b.ctype == -1).sum() > 0
    s = b.init()
    rng = np.random.default_rng(9)
    # density up (the pressure rises), momentum jitter well below the
    # thermal share of E
    s[0] *= 1 + 0.05 * rng.random(b.n)
    s[1:4] += 1e-4 * np.abs(s[1]).max() * rng.normal(size=(3, b.n))
    n, nx = b.n, 12
    ref = s.copy()
    dt = 0.2 * h / 8e5
    st, sm, zc = b.step(sid, dt, s)
    assert st == 0
    c3 = ref.reshape(14, nx, nx, nx)
    ct = b.ctype.reshape(nx, nx, nx)
    expect = ref[0:4].reshape(4, nx, nx, nx).copy()
    for d in range(3):
        rot = [d, (d + 1) % 3, (d + 2) % 3]
        ax = 2 - d
        U = np.concatenate([c3[0:1], c3[1:4][rot], c3[4:5], c3[5:8][rot]])
        sl_l = [slice(None)] * 3
        sl_r = [slice(None)] * 3
        sl_l[ax] = slice(0, nx - 1)
        sl_r[ax] = slice(1, nx)
        L = U[(slice(None),) + tuple(sl_l)].reshape(8, -1)
        R = U[(slice(None),) + tuple(sl_r)].reshape(8, -1)
        bgl = b.bg[3 * d:3 * d + 3][rot].reshape(3, nx, nx, nx)[(slice(None),) + tuple(sl_l)].reshape(3, -1)
        ok = ((ct[tuple(sl_l)] >= 0) & (ct[tuple(sl_r)] >= 0)).reshape(-1)
        F, sp, stt = O.flux_batch(sid, L, R, bgl, sb.adiabatic_index, sb.vacuum_permeability)
        F[:, ~ok] = 0
        Fg = np.zeros((4,) + L.shape[1:])
        Fg[0] = F[0]
        for q in range(3):
            Fg[1 + rot[q]] = F[1 + q]
        shape = [4, nx, nx, nx]
        shape[ax + 1] = nx - 1
        Fg = Fg.reshape(shape) * dt
        full = np.zeros((4, nx, nx, nx))
        full[(slice(None),) + tuple(sl_r)] = Fg           # face owned by its +side cell
        hi = np.zeros_like(full)
        hi[(slice(None),) + tuple(sl_l)] = Fg             # the cell's +d face
        expect += (full - hi) / sb.cell_length[d]
        assert np.isclose(sm[d], sp[ok].max(), rtol=0, atol=0)
    normal = b.ctype == 1
    got = s[0:4][:, normal]
    want = expect.reshape(4, -1)[:, normal]
    np.testing.assert_allclose(got[0], want[0], rtol=1e-13)
    mscale = np.abs(want[1:4]).max()
    assert np.abs(got[1:4] - want[1:4]).max() <= 1e-12 * mscale
    # cells that are not normal keep rho and m
    np.testing.assert_array_equal(s[0:4][:, ~normal], ref[0:4][:, ~normal])
