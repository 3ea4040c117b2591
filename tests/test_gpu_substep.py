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
    with Grid(cells, length, gamma, mu0) as g:
        g.upload(state)
        g.set_substepping(0, smax)
        steps_g, t_g = g.run(name, 0.3, 0.0, 1e9, max_steps=5)
        mp_g = g.last_max_substep()
        out = g.download()
    print(name, "smax", smax, "largest period", mp_g, "t", t_g, flush=True)
    assert steps_g == 5 and mp_g == mp_o and t_g == t_o
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
