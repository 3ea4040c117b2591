"""Per-cell 2^N substepping on the GPU against the oracle's restatement
(oracle/pamhd_oracle.c: or_run_periodic_substep), through the C ABI, bit for
bit: a hot sphere in a cold periodic box gives the outer cells periods of
up to 2^substep_max substeps."""
import numpy as np
import pytest

import oracle_lib as O
from pamhd_amd import configs

pytestmark = pytest.mark.gpu

SOLVERS = [("rusanov", O.RUSANOV), ("hll-athena", O.HLL), ("hlld-athena", O.HLLD)]


@pytest.mark.parametrize("name,sid", SOLVERS)
@pytest.mark.parametrize("smax,cells", [(1, (16, 12, 10)), (3, (16, 12, 10)), (2, (20, 14, 1))])
def test_substepped_run_bit_exact(gpu, name, sid, smax, cells):
    from pamhd_amd.solver import Grid
    state, length, gamma, mu0 = configs.hot_blob(cells, contrast=30.0, radius=0.25, seed=3)
    g_o = O.make_grid(cells, length, gamma, mu0)
    ref = state.copy()
    st, steps_o, t_o, mp_o = O.run_periodic_substep(g_o, sid, 0.3, 0, smax, 0.0, 1e9, ref, max_steps=5)
    assert st == 0 and steps_o == 5 and mp_o > 1
    fc_o = O.last_flux_calcs()
    with Grid(cells, length, gamma, mu0) as g:
        g.upload(state)
        g.set_substepping(0, smax)
        steps_g, t_g = g.run(name, 0.3, 0.0, 1e9, max_steps=5)
        mp_g = g.last_max_substep()
        fc_g = g.last_flux_calcs()
        out = g.download()
    print(name, "smax", smax, "largest period", mp_g, "t", t_g, "flux_calcs", fc_g, flush=True)
    assert steps_g == 5 and mp_g == mp_o and t_g == t_o
    # timestep()'s third return value: the faces solved over the last step
    assert fc_g == fc_o
    np.testing.assert_array_equal(out, ref)


def test_state_changed_between_substepped_runs(gpu):
    """Consecutive substepped steps keep the working layout (the expansion of
    the state) and skip re-expanding it; any other call on the handle in
    between makes the next step expand the state again.  Here the state is
    changed through the ABI between two runs (uploaded with a scaled
    density, and read back in the middle of a run's worth of single steps),
    and the result must equal the oracle doing the same, bit for bit."""
    from pamhd_amd.solver import Grid
    cells = (16, 12, 10)
    state, length, gamma, mu0 = configs.hot_blob(cells, contrast=30.0, radius=0.25, seed=5)
    g_o = O.make_grid(cells, length, gamma, mu0)
    ref = state.copy()
    st, steps_o, t1_o, _ = O.run_periodic_substep(g_o, O.HLLD, 0.3, 0, 3, 0.0, 1e9, ref, max_steps=3)
    assert st == 0 and steps_o == 3
    ref[0] *= 1.01
    ref[4] *= 1.01
    st, steps_o, t2_o, _ = O.run_periodic_substep(g_o, O.HLLD, 0.3, 0, 3, 0.0, 1e9, ref, max_steps=3)
    assert st == 0 and steps_o == 3
    with Grid(cells, length, gamma, mu0) as g:
        g.upload(state)
        g.set_substepping(0, 3)
        steps_g, t1_g = g.run("hlld-athena", 0.3, 0.0, 1e9, max_steps=3)
        mid = g.download()
        mid[0] *= 1.01
        mid[4] *= 1.01
        g.upload(mid)
        steps_g2, t2_g = g.run("hlld-athena", 0.3, 0.0, 1e9, max_steps=3)
        out = g.download()
    assert steps_g == 3 and steps_g2 == 3 and t1_g == t1_o and t2_g == t2_o
    np.testing.assert_array_equal(out, ref)


def test_reads_between_substepped_steps(gpu):
    """Single substepped steps with a state read-back between some of them
    (each read makes the next step expand the state again) give the same
    state as the uninterrupted run, bit for bit."""
    from pamhd_amd.solver import Grid
    cells = (16, 12, 10)
    state, length, gamma, mu0 = configs.hot_blob(cells, contrast=30.0, radius=0.25, seed=7)
    outs = []
    for read_between in (False, True):
        with Grid(cells, length, gamma, mu0) as g:
            g.upload(state)
            g.set_substepping(0, 3)
            g.apply_boundaries(0.0)
            g.timestep("hlld-athena", 0.0, 0.3)
            for q in range(5):
                g.step("hlld-athena", 1e300, 0.3)
                if read_between and q % 2 == 0:
                    g.download()
            outs.append(g.download())
    np.testing.assert_array_equal(outs[0], outs[1])


@pytest.mark.parametrize("walk", [0, 7])
def test_substepped_run_many_blocks_per_reduction_slot(gpu, walk):
    """The step's reductions (smallest time step, largest level and period,
    faces solved) go to 64 slots by block index; a grid of 320 blocks puts
    several blocks into each slot (the reference's default options 0, 999).
    With 7 workgroups walking the substep passes' block lists, each walks
    many entries of a list, as on the 256^3 grid's 65 536 blocks."""
    from pamhd_amd.solver import Grid
    cells = (64, 32, 40)
    state, length, gamma, mu0 = configs.hot_blob(cells, contrast=30.0, radius=0.25, seed=6)
    g_o = O.make_grid(cells, length, gamma, mu0)
    ref = state.copy()
    st, steps_o, t_o, mp_o = O.run_periodic_substep(g_o, O.HLLD, 0.3, 0, 999, 0.0, 1e9, ref,
                                                    max_steps=4)
    assert st == 0 and steps_o == 4 and mp_o > 1
    fc_o = O.last_flux_calcs()
    with Grid(cells, length, gamma, mu0) as g:
        g.upload(state)
        g.set_substep_walk(walk)
        g.set_substepping(0, 999)
        steps_g, t_g = g.run("hlld-athena", 0.3, 0.0, 1e9, max_steps=4)
        mp_g = g.last_max_substep()
        fc_g = g.last_flux_calcs()
        periods = g.last_substep_periods()
        out = g.download()
    print("largest period", mp_g, "periods", np.unique(periods, return_counts=True), flush=True)
    assert steps_g == 4 and mp_g == mp_o and t_g == t_o and fc_g == fc_o
    assert periods.max() == mp_g and periods.min() >= 1
    np.testing.assert_array_equal(out, ref)


def test_substep_max_zero_is_the_global_step(gpu):
    from pamhd_amd.solver import Grid
    cells = (12, 10, 8)
    state, length, gamma, mu0 = configs.hot_blob(cells, contrast=30.0, radius=0.25, seed=4)
    outs = []
    for smax in (None, 0):
        with Grid(cells, length, gamma, mu0) as g:
            g.upload(state)
            if smax is not None:
                g.set_substepping(0, smax)
            g.run("hll-athena", 0.3, 0.0, 1e9, max_steps=4)
            outs.append(g.download())
    np.testing.assert_array_equal(outs[0], outs[1])


def test_substepping_needs_fp64(gpu):
    from pamhd_amd import _lib
    from pamhd_amd.solver import Grid
    cells = (8, 8, 8)
    state, length, gamma, mu0 = configs.hot_blob(cells)
    with Grid(cells, length, gamma, mu0, precision="mixed") as g:
        with pytest.raises(_lib.PamhdError) as e:
            g.set_substepping(0, 2)
        assert e.value.code == _lib.PAMHD_E_UNSUPPORTED


def _cell_coordinates(cells):
    nx, ny, nz = cells
    x = (np.arange(nx) + 0.5) / nx - 0.5
    y = (np.arange(ny) + 0.5) / ny - 0.5
    z = (np.arange(nz) + 0.5) / nz - 0.5
    zz, yy, xx = np.meshgrid(z, y, x, indexing="ij")
    return xx.reshape(-1), np.sqrt(xx * xx + yy * yy + zz * zz).reshape(-1)


@pytest.mark.parametrize("name,sid", [("hll-athena", O.HLL), ("hlld-athena", O.HLLD)])
def test_per_cell_time_dependent_options_bit_exact(gpu, name, sid):
    """string-valued substep-min / substep-max: each cell's options, set
    again before every step from the cell centre and the time, as the main
    loop of tests/mhd/test.cpp evaluates them"""
    from pamhd_amd.solver import Grid
    cells = (16, 12, 10)
    state, length, gamma, mu0 = configs.hot_blob(cells, contrast=30.0, radius=0.25, seed=3)
    x, r = _cell_coordinates(cells)

    def options(t, step):
        lo = np.where(r > 0.4, 3, 2).astype(np.int32)
        hi = np.where(x < -0.5 + 0.25 * step, 1, 3).astype(np.int32)
        return lo, hi

    g_o = O.make_grid(cells, length, gamma, mu0)
    ref = state.copy()
    maxv = -np.ones((3, ref.shape[1]))
    lo, hi = options(0.0, 0)
    st, dt0, _ = O.step_substep(g_o, sid, 0.3, 0.0, lo, hi, ref, maxv)
    assert st == 0
    with Grid(cells, length, gamma, mu0) as g:
        g.upload(state)
        g.set_substepping(lo, hi)
        assert g.timestep(name, 0.0, 0.3) == dt0 == 0.0
        t_o = t_g = 0.0
        periods = []
        for step in range(1, 6):
            lo, hi = options(t_o, step)
            st, dt_o, mp_o = O.step_substep(g_o, sid, 0.3, 1e9 - t_o, lo, hi, ref, maxv)
            assert st == 0
            O.refit_periodic(g_o, ref)
            g.set_substepping(lo, hi)
            dt_g = g.step(name, 1e9 - t_g, 0.3)
            assert dt_g == dt_o and g.last_max_substep() == mp_o
            t_o += dt_o
            t_g += dt_g
            periods.append(mp_o)
        out = g.download()
    print(name, "largest periods", periods, flush=True)
    assert max(periods) > 1
    np.testing.assert_array_equal(out, ref)


def test_reference_default_options_bit_exact(gpu):
    """substep-min 0, substep-max 999 (the options' defaults): levels are
    bounded by the time-step ratios alone"""
    from pamhd_amd.solver import Grid
    cells = (16, 12, 10)
    state, length, gamma, mu0 = configs.hot_blob(cells, contrast=30.0, radius=0.25, seed=3)
    g_o = O.make_grid(cells, length, gamma, mu0)
    ref = state.copy()
    st, steps_o, t_o, mp_o = O.run_periodic_substep(g_o, O.HLL, 0.3, 0, 999, 0.0, 1e9, ref, max_steps=4)
    assert st == 0 and mp_o > 1
    with Grid(cells, length, gamma, mu0) as g:
        g.upload(state)
        g.set_substepping(0, 999)
        steps_g, t_g = g.run("hll-athena", 0.3, 0.0, 1e9, max_steps=4)
        assert g.last_max_substep() == mp_o and t_g == t_o
        out = g.download()
    np.testing.assert_array_equal(out, ref)


def test_substep_min_above_the_allowed_level_fails_the_step(gpu):
    from pamhd_amd import _lib
    from pamhd_amd.solver import Grid
    cells = (8, 8, 8)
    state, length, gamma, mu0 = configs.hot_blob(cells, contrast=30.0, radius=0.2, seed=2)
    _, r = _cell_coordinates(cells)
    lo = np.where(r < 0.15, 2, 0).astype(np.int32)
    with Grid(cells, length, gamma, mu0) as g:
        g.upload(state)
        g.set_substepping(lo, 1)
        g.timestep("hll-athena", 0.0, 0.3)
        with pytest.raises(_lib.PamhdError) as e:
            g.step("hll-athena", 1e9, 0.3)
        assert e.value.code == _lib.PAMHD_E_SUBSTEP
