"""The reference's JSON configurations run on the GPU through the loader
(pamhd_amd.config.run: test.cpp's program loop over the C ABI) against the
same configurations in the oracle (oracle_lib.run_config, or
run_bounded_config for grids with boundaries or closed dimensions), bit for
bit, and the golden ones against the reference's .dc results."""
import os

import numpy as np
import pytest

import golden_cases as G
import oracle_lib as O
from pamhd_amd import config, dc

pytestmark = pytest.mark.gpu

CFG_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "config")
CONFIGS = sorted(f[:-5] for f in os.listdir(CFG_DIR)
                 if not f.startswith(("two_pop", "bow_shock")))
# whole runs of the periodic configurations, but at most 60 steps of the
# 3-D Kelvin-Helmholtz case and 300 of those with boundaries
MAX_STEPS = {"kelvin_helmholtz_3d_roe_athena": 60}
BOUNDED_STEPS = 300


@pytest.mark.parametrize("name", CONFIGS)
def test_config_runs_bit_exact(gpu, name):
    cfg = config.load(os.path.join(CFG_DIR, name + ".json"))
    if cfg.bounded:
        limit = BOUNDED_STEPS
        st, steps_o, t_o, periods, state = O.run_bounded_config(cfg, max_steps=limit)
    else:
        limit = MAX_STEPS.get(name, 1 << 40)
        state, bg = cfg.initial_state()
        st, steps_o, t_o, periods = O.run_config(cfg, state, bg, max_steps=limit)
    assert st == 0
    res = config.run(cfg, max_steps=limit)
    print(name, cfg.cells, cfg.solver, "steps", res["steps"], "largest period", max(periods),
          flush=True)
    assert res["steps"] == steps_o and res["time"] == t_o
    assert res["max_substep"] == periods[-1]
    np.testing.assert_array_equal(res["state"], state)
    if name in G.GOLDENS:
        meta, ids, mhd = G.read_dc(os.path.join(G.DC_DIR, name + ".dc"))
        assert res["steps"] == meta["sim_step"]
        gold = G.dc_to_grid(meta, ids, mhd)
        out = res["state"]
        mine = np.concatenate([out[:5], G.volume_field(cfg.cells, out)])
        for k, (viol, rel) in G.mhd_diff(mine, gold).items():
            assert viol <= 1e-15, (name, k, viol, rel)


def test_time_dependent_substep_options(gpu):
    """substep-max as an expression of t, set again before every step"""
    import json
    obj = json.load(open(os.path.join(CFG_DIR, "blast_wave_3d_rusanov.json")))
    obj["substep-max"] = "t < 0.01 ? 0 : 999"
    obj["time-length"] = 0.1
    cfg = config.load(obj)
    assert cfg.substeps_depend_on_time()
    state, bg = cfg.initial_state()
    st, steps_o, t_o, periods = O.run_config(cfg, state, bg)
    # the global step while substep-max is 0, then periods up to 4
    assert st == 0 and periods[0] == 1 and max(periods) == 4
    res = config.run(cfg)
    assert res["steps"] == steps_o and res["time"] == t_o
    np.testing.assert_array_equal(res["state"], state)


def test_saves_follow_save_mhd_n(gpu, tmp_path):
    """save-mhd-n 0: the state after the final init (step 0) and at the end,
    as .dc files that read back to the run's state"""
    cfg = config.load(os.path.join(CFG_DIR, "alfven_2d_x_1.json"))
    res = config.run(cfg, output_directory=str(tmp_path))
    names = sorted(os.path.basename(p) for p in res["saves"])
    assert names == ["mhd_000000000.dc", "mhd_%09d.dc" % res["steps"]]
    meta, ids, rows = dc.read(res["saves"][-1])
    assert meta["sim_step"] == res["steps"] and meta["sim_time"] == res["time"]
    mhd_rows, _ = dc.cells_of_grid_state(res["state"], cfg.cells)
    np.testing.assert_array_equal(rows["mhd     "], mhd_rows)
    fast = config.run(cfg)
    np.testing.assert_array_equal(fast["state"], res["state"])


@pytest.mark.parametrize("name", ["shock_tube_1d_px_hll_athena", "rarefaction_1d_px_hll_athena",
                                  "reconnection_2d_y1_hll_athena"])
def test_bounded_grid_through_pamhd_run(gpu, name):
    """pamhd_run on a grid with constant boundaries: apply_boundaries at
    time-start, the final init and the loop, all inside the library, give
    the oracle's bits"""
    from pamhd_amd.solver import Grid
    cfg = config.load(os.path.join(CFG_DIR, name + ".json"))
    assert cfg.bounded and not cfg.boundaries_depend_on_time()
    assert not cfg.substeps_depend_on_time()
    st, steps_o, t_o, periods, want = O.run_bounded_config(cfg, max_steps=40)
    assert st == 0
    state, bg, types = cfg.bounded_initial_state()
    with Grid(cfg.cells, cfg.cell_length, cfg.adiabatic_index, cfg.vacuum_permeability,
              cfg.proton_mass, start=cfg.start, periodic=cfg.periodic, types=types) as g:
        assert g.kind == "bounded"
        g.upload(state)
        if bg is not None:
            g.set_background_faces(bg)
        g.set_substepping(*cfg.substep_options(cfg.time_start))
        g.set_boundary_values(*cfg.boundary_values(cfg.time_start))
        steps, t = g.run(cfg.solver, cfg.time_step_factor, cfg.time_start,
                         cfg.time_start + cfg.time_length, max_steps=40)
        assert steps == steps_o and t == t_o
        np.testing.assert_array_equal(g.download(), want)
