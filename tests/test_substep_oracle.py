"""The oracle's per-cell 2^N substepping (timestep() of
source/mhd/solve.hpp:1837-1924 with source/substepping.hpp) pinned by:

* equality with the unsubstepped periodic step when every period is 1
  (substep-max 0: the reference's global time step);
* the periods: a brute-force min-plus closure of the restriction (each cell
  at most one level above any face or edge neighbour), the per-cell time
  step of minimize_timestep and the substep length, recomputed in numpy;
* conservation of mass and momentum and preservation of the face
  divergence over substepped steps (catches a face flux applied to one
  cell with a different weight than to the other, or a face copy that
  misses edge terms);
* invariance under a periodic translation of the grid and the time
  accounting (step length = largest period x substep length).
"""
import numpy as np
import pytest

import oracle_lib as O
from pamhd_amd import configs


def _neighbourhood(shape):
    nx, ny, nz = shape
    offs = [(di, dj, dk) for dk in (-1, 0, 1) for dj in (-1, 0, 1) for di in (-1, 0, 1)
            if 0 < (di != 0) + (dj != 0) + (dk != 0) < 3]
    idx = np.arange(nx * ny * nz).reshape(nz, ny, nx)
    return [np.roll(idx, (-dk, -dj, -di), (0, 1, 2)).reshape(-1) for di, dj, dk in offs]


def _maxv_after_steps(g, state, solver, steps=2):
    """per-face speeds of a short substepped run (the lagged input of the periods)"""
    n = state.shape[1]
    maxv = -np.ones((3, n))
    st, dt, mp = O.step_substep(g, solver, 0.3, 0.0, 0, 3, state, maxv)
    assert st == 0
    for _ in range(steps):
        st, dt, mp = O.step_substep(g, solver, 0.3, 1e9, 0, 3, state, maxv)
        assert st == 0
        O.refit_periodic(g, state)
    return maxv


@pytest.mark.parametrize("sid", [O.RUSANOV, O.HLL, O.HLLD])
def test_one_substep_equals_the_periodic_step(sid):
    cells = (12, 10, 8)
    state, length, gamma, mu0 = configs.orszag_tang_vortex(cells, noise=0.02, seed=4)
    g = O.make_grid(cells, length, gamma, mu0)
    a = state.copy()
    st_a, steps_a, t_a = O.run_periodic(g, sid, 0.3, 0.0, 1e9, a, max_steps=6)
    b = state.copy()
    st_b, steps_b, t_b, mp = O.run_periodic_substep(g, sid, 0.3, 0, 0, 0.0, 1e9, b, max_steps=6)
    assert st_a == st_b == 0 and steps_a == steps_b == 6 and mp == 1
    assert t_a == t_b
    np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("smax", [1, 2, 3])
def test_periods_match_a_brute_force_restriction(smax):
    cells = (8, 8, 8)
    state, length, gamma, mu0 = configs.hot_blob(cells, contrast=30.0, radius=0.2, seed=2)
    g = O.make_grid(cells, length, gamma, mu0)
    maxv = _maxv_after_steps(g, state, O.HLL)
    cfl, max_dt = 0.3, 1e9
    st, period, minsub, sub_dt, mp = O.substep_periods(g, cfl, max_dt, 0, smax, maxv)
    assert st == 0
    n = state.shape[1]
    nx, ny, nz = cells
    # minimize_timestep: min over the six faces of cfl * len / speed
    f = maxv.reshape(3, nz, ny, nx)
    lens = length
    ts = np.full((nz, ny, nx), np.finfo(float).max)
    for d in range(3):
        for v in (f[d], np.roll(f[d], -1, 2 - d)):
            ts = np.minimum(ts, np.where(v > 0, cfl * lens[d] / np.where(v > 0, v, 1), 0.0))
    ts = ts.reshape(-1)
    assert sub_dt == ts.min()
    ratio = np.minimum(ts, max_dt) / sub_dt
    mant, ex = np.frexp(ratio)
    step = np.minimum(smax, np.maximum(0, ex - 1))
    # restriction: e(c) = min over cells m of step(m) + graph distance(c, m)
    nbrs = _neighbourhood(cells)
    dist = np.full((n, n), 10 ** 6, dtype=np.int64)
    for c in range(n):
        dist[c, c] = 0
        frontier = [c]
        level = 0
        while frontier:
            level += 1
            nxt = []
            for a in frontier:
                for nb in nbrs:
                    b = nb[a]
                    if dist[c, b] > level:
                        dist[c, b] = level
                        nxt.append(b)
            frontier = nxt
    e = (step[None, :] + dist).min(1)
    np.testing.assert_array_equal(period, 2 ** e)
    ms = period.copy()
    for nb in nbrs:
        ms = np.minimum(ms, period[nb])
    np.testing.assert_array_equal(minsub, ms)
    assert mp == period.max() and mp > 1


@pytest.mark.parametrize("sid", [O.HLL, O.HLLD])
def test_substepped_steps_conserve_and_keep_divergence(sid):
    cells = (16, 12, 10)
    state, length, gamma, mu0 = configs.hot_blob(cells, contrast=30.0, radius=0.25, seed=3)
    g = O.make_grid(cells, length, gamma, mu0)
    s = state.copy()
    st, steps, t, mp = O.run_periodic_substep(g, sid, 0.3, 0, 3, 0.0, 1e9, s, max_steps=5)
    assert st == 0 and steps == 5 and mp > 1
    vol = length[0] * length[1] * length[2]
    m0 = state[0].sum() * vol
    m1 = s[0].sum() * vol
    assert abs(m1 - m0) <= 1e-13 * m0
    for q in (1, 2, 3):
        assert abs(s[q].sum() - state[q].sum()) * vol <= 1e-13 * m0
    assert O.div_face2volume(g, s) < 1e-13
    assert np.all(s[0] > 0)
    assert not np.array_equal(s, state)


def test_translation_invariance_and_time_accounting():
    cells = (12, 8, 8)
    state, length, gamma, mu0 = configs.hot_blob(cells, contrast=30.0, radius=0.25, seed=5)
    g = O.make_grid(cells, length, gamma, mu0)
    shift = 5
    sh = np.ascontiguousarray(np.roll(state.reshape(8, cells[2], cells[1], cells[0]), shift, 3)
                              .reshape(8, -1))
    a = state.copy()
    b = sh.copy()
    maxv_a = -np.ones((3, a.shape[1]))
    maxv_b = -np.ones((3, a.shape[1]))
    ts = []
    for max_dt in (0.0, 1e9, 1e9, 1e9):
        st_a, dt_a, mp_a = O.step_substep(g, O.HLL, 0.3, max_dt, 0, 3, a, maxv_a)
        st_b, dt_b, mp_b = O.step_substep(g, O.HLL, 0.3, max_dt, 0, 3, b, maxv_b)
        assert st_a == st_b == 0 and mp_a == mp_b and dt_a == dt_b
        ts.append((dt_a, mp_a))
    st, period, minsub, sub_dt, mp = O.substep_periods(g, 0.3, 1e9, 0, 3, maxv_a)
    # the last step was as long as its largest period times the substep
    assert mp > 1
    back = np.roll(b.reshape(8, cells[2], cells[1], cells[0]), -shift, 3).reshape(8, -1)
    np.testing.assert_allclose(back, a, rtol=1e-13, atol=1e-15)


def _cell_radius(cells):
    nx, ny, nz = cells
    x = (np.arange(nx) + 0.5) / nx - 0.5
    y = (np.arange(ny) + 0.5) / ny - 0.5
    z = (np.arange(nz) + 0.5) / nz - 0.5
    zz, yy, xx = np.meshgrid(z, y, x, indexing="ij")
    return np.sqrt(xx * xx + yy * yy + zz * zz).reshape(-1), xx.reshape(-1)


def test_per_cell_options_match_a_brute_force_restriction():
    """Each cell's own substep-min / substep-max (the string-valued options):
    the minimum shifted so that its smallest is 0, the substep length as the
    smallest time step / 2^substep-min, each cell's level capped by its own
    substep-max."""
    cells = (8, 8, 8)
    state, length, gamma, mu0 = configs.hot_blob(cells, contrast=30.0, radius=0.2, seed=2)
    g = O.make_grid(cells, length, gamma, mu0)
    maxv = _maxv_after_steps(g, state, O.HLL)
    r, x = _cell_radius(cells)
    lo = np.where(r > 0.35, 3, 2).astype(np.int32)     # shifted to 0 / 1
    hi = np.where(x < 0, 4, 2).astype(np.int32)
    cfl, max_dt = 0.3, 1e9
    st, period, minsub, sub_dt, mp = O.substep_periods(g, cfl, max_dt, lo, hi, maxv)
    assert st == 0
    nx, ny, nz = cells
    f = maxv.reshape(3, nz, ny, nx)
    ts = np.full((nz, ny, nx), np.finfo(float).max)
    for d in range(3):
        for v in (f[d], np.roll(f[d], -1, 2 - d)):
            ts = np.minimum(ts, np.where(v > 0, cfl * length[d] / np.where(v > 0, v, 1), 0.0))
    ts = ts.reshape(-1)
    lmin = lo - lo.min()
    ret = (ts / 2.0 ** lmin).min()
    assert sub_dt == ret and ret < ts.min()
    _, ex = np.frexp(np.minimum(ts, max_dt) / ret)
    step = np.minimum(hi, np.maximum(0, ex - 1))
    assert np.all(lmin <= step)
    nbrs = _neighbourhood(cells)
    e = step.copy()
    for _ in range(int(step.max()) + 1):   # min-plus closure by relaxation
        for nb in nbrs:
            e = np.minimum(e, e[nb] + 1)
    np.testing.assert_array_equal(period, 2 ** e)
    assert mp == period.max() and mp >= 4


def test_substep_min_above_the_allowed_level_is_an_error():
    cells = (8, 8, 8)
    state, length, gamma, mu0 = configs.hot_blob(cells, contrast=30.0, radius=0.2, seed=2)
    g = O.make_grid(cells, length, gamma, mu0)
    maxv = _maxv_after_steps(g, state, O.HLL)
    r, _ = _cell_radius(cells)
    lo = np.where(r < 0.15, 2, 0).astype(np.int32)
    st, *_ = O.substep_periods(g, 0.3, 1e9, lo, 1, maxv)
    assert st == O.SUBSTEP_ERROR
    st, *_ = O.substep_periods(g, 0.3, 1e9, lo, 999, maxv)
    assert st == 0


def test_flux_calcs_counts_the_faces_solved():
    """timestep()'s third return value (source/mhd/solve.hpp:253-313,
    :1918): every face once per global step; with substepping a face of
    owner o (its negative-side cell) once per minsub(o) substeps, i.e.
    3 maxp / minsub(o) solutions over a step of maxp substeps."""
    cells = (12, 10, 8)
    n = 12 * 10 * 8
    state, length, gamma, mu0 = configs.hot_blob(cells, contrast=30.0, radius=0.25, seed=3)
    g = O.make_grid(cells, length, gamma, mu0)
    st, new, sm, zc = O.step(g, O.HLL, 1e-4, state, refits=3)
    assert st == 0 and O.last_flux_calcs() == 3 * n
    s = state.copy()
    maxv = _maxv_after_steps(g, s, O.HLL)
    st, period, minsub, sub_dt, maxp = O.substep_periods(g, 0.3, 1e9, 0, 3, maxv.copy())
    assert st == 0 and maxp > 1 and minsub.min() < maxp
    st, dt, mp = O.step_substep(g, O.HLL, 0.3, 1e9, 0, 3, s, maxv)
    assert st == 0 and mp == maxp
    assert O.last_flux_calcs() == int((3 * (maxp // minsub)).sum())
    assert O.last_flux_calcs() < 3 * n * maxp
