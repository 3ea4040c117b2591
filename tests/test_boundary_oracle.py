"""Uniform grids with value and copy boundaries: the oracle's restatement
(or_bnd_classify / or_bnd_apply / or_bnd_sync / or_bnd_step) and the
product's classification (pamhd_amd.boundary), pinned by
* equality of the cell-typed substepped step with the periodic one
  (or_step_substep) on a periodic grid without boundaries, bit for bit, all
  four solvers, with apply_boundaries reduced to its refit and the sync;
* the product's classification against the oracle's on every boundary
  configuration of the fixtures (two independent restatements of
  Boundaries::classify, set_solver_info and classify_faces);
* hand-worked cell and face types of the 1-D shock tube;
* a uniform state held by value boundaries of the same state staying
  uniform to the last bit (a wrong boundary value, a flux through a face to
  a missing neighbour or an update of a boundary cell breaks it);
* the 1-D shock tube along x, y and z giving the same bits up to the
  relabelling of the axes;
* time-dependent value boundaries taking the nearest time stamp's
  expression at the step's end time.
"""
import json
import math
import os

import numpy as np
import pytest

import oracle_lib as O
from pamhd_amd import boundary, config, configs

CFG_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "config")
BOUNDED = ["shock_tube_1d_px_hll_athena", "shock_tube_1d_py_hll_athena",
           "shock_tube_1d_pz_hll_athena", "shock_tube_1d_px_roe_athena",
           "shock_tube_1d_px_hlld_athena", "rarefaction_1d_px_hll_athena",
           "solar_wind_1d_density_10mHz", "solar_wind_1d_constant",
           "reconnection_2d_y1_hll_athena", "orszag_tang_2d_hll_athena"]


def load(name):
    return config.load(os.path.join(CFG_DIR, name + ".json"))


def _expand(state, cells):
    nx, ny, nz = cells
    n = nx * ny * nz
    S = np.zeros((14, n))
    S[:5] = state[:5]
    f = state[5:8].reshape(3, nz, ny, nx)
    plus = [np.roll(f[0], -1, 2).reshape(-1), np.roll(f[1], -1, 1).reshape(-1),
            np.roll(f[2], -1, 0).reshape(-1)]
    for d in range(3):
        S[8 + 2 * d] = state[5 + d]
        S[9 + 2 * d] = plus[d]
        S[5 + d] = 0.5 * (state[5 + d] + plus[d])
    return S


@pytest.mark.parametrize("sid", [O.RUSANOV, O.HLL, O.HLLD, O.ROE])
def test_cell_typed_step_equals_the_periodic_step(sid):
    cells = (12, 10, 8)
    state, length, gamma, mu0 = configs.hot_blob(cells, contrast=30.0, radius=0.25, seed=3)
    n = state.shape[1]
    g = O.make_grid(cells, length, gamma, mu0)
    none = np.zeros((4, n))
    G = O.BoundedGrid(cells, length, gamma, mu0, 1.0, (1, 1, 1), none, none)
    assert G.status == 0 and (G.ctype == 1).all() and (G.finfo == 1).all()
    a = state.copy()
    S = _expand(state, cells)
    ma = -np.ones((3, n))
    mb = -np.ones((3, n))
    zeros = (np.zeros((4, n)), np.zeros(n), np.zeros((3, n)), np.zeros(n), np.zeros((6, n)))
    periods = []
    for max_dt in (0.0, 1e9, 1e9, 1e9, 1e9):
        r1 = O.step_substep(g, sid, 0.3, max_dt, 0, 3, a, ma)
        r2 = G.step(sid, 0.3, max_dt, 0, 3, S, mb)
        assert r1 == r2 and r1[0] == 0
        periods.append(r1[2])
        if max_dt > 0:
            O.refit_periodic(g, a)
            assert G.apply(S, *zeros) == 0
            G.sync(S)
        np.testing.assert_array_equal(S[:5], a[:5])
        np.testing.assert_array_equal(S[8:14:2], a[5:8])
    assert max(periods) > 1


@pytest.mark.parametrize("name", BOUNDED)
def test_classification_matches_the_oracle(name):
    cfg = load(name)
    assert cfg.bounded
    vtype, ctype, finfo = cfg.classify()
    iv, ic = cfg.boundary_masks()
    G = O.BoundedGrid(cfg.cells, cfg.cell_length, cfg.adiabatic_index, cfg.vacuum_permeability,
                      cfg.proton_mass, cfg.periodic, iv, ic)
    assert G.status == 0
    np.testing.assert_array_equal(G.vtype, vtype)
    np.testing.assert_array_equal(G.ctype, ctype)
    np.testing.assert_array_equal(G.finfo, finfo)


def test_shock_tube_types_by_hand():
    """1002 x 1 x 1, closed: value cells at both ends (all variables), the
    rest normal; an end cell's face toward the interior and its y / z faces
    are normal (pass one of classify_faces: every face but the one opposite
    the normal neighbour), its outer face a boundary face"""
    cfg = load("shock_tube_1d_px_hll_athena")
    vtype, ctype, finfo = cfg.classify()
    n = cfg.n
    want = np.ones(n, dtype=np.int8)
    want[[0, n - 1]] = 0
    np.testing.assert_array_equal(ctype, want)
    assert (vtype[:, [0, n - 1]] == 3).all() and (vtype[:, 1:-1] == 1).all()
    np.testing.assert_array_equal(finfo[:, 0], [0, 1, 1, 1, 1, 1])
    np.testing.assert_array_equal(finfo[:, n - 1], [1, 0, 1, 1, 1, 1])
    assert (finfo[:, 1:-1] == 1).all()


@pytest.mark.parametrize("solver", ["rusanov", "hll-athena", "hlld-athena", "roe-athena"])
def test_uniform_state_held_by_value_boundaries(solver):
    obj = json.load(open(os.path.join(CFG_DIR, "shock_tube_1d_px_hll_athena.json")))
    obj["solver-mhd"] = solver
    for var, val in (("number-density", 3e6), ("pressure", 3e-12)):
        obj[var]["default"] = val
        for b in obj[var]["value-boundaries"]:
            b["values"] = [val]
    obj["velocity"]["default"] = [2e4, -1e3, 5e2]
    for b in obj["velocity"]["value-boundaries"]:
        b["values"] = [[2e4, -1e3, 5e2]]
    obj["magnetic-field"]["default"] = [1.5e-9, 1e-9, -2e-10]
    for b in obj["magnetic-field"]["value-boundaries"]:
        b["values"] = [[1.5e-9, 1e-9, -2e-10]]
    cfg = config.load(obj)
    S0, _, _ = cfg.bounded_initial_state()
    st, steps, t, periods, S = O.run_bounded_config(cfg, max_steps=30)
    assert st == 0 and steps == 30
    for q in range(14):
        assert np.all(S[q] == S[q][0]), q
    np.testing.assert_array_equal(S[[0, 1, 2, 3, 8, 9, 10, 11, 12, 13]],
                                  S0[[0, 1, 2, 3, 8, 9, 10, 11, 12, 13]])


def test_shock_tube_is_invariant_under_axis_exchange():
    outs = []
    for name in ("shock_tube_1d_px_hll_athena", "shock_tube_1d_py_hll_athena",
                 "shock_tube_1d_pz_hll_athena"):
        st, steps, t, periods, S = O.run_bounded_config(load(name), max_steps=40)
        assert st == 0
        outs.append((steps, t, S))
    assert outs[0][:2] == outs[1][:2] == outs[2][:2]
    # field (1.5, +-1, 0) nT along +x becomes (+-1, 1.5, 0) along +y (x and y
    # exchanged) and (+-1, 0, 1.5) along +z (x -> z, y -> x, z -> y)
    sx, sy, sz = outs[0][2], outs[1][2], outs[2][2]
    perm_y = [0, 2, 1, 3, 4, 6, 5, 7, 10, 11, 8, 9, 12, 13]
    perm_z = [0, 3, 1, 2, 4, 7, 5, 6, 12, 13, 8, 9, 10, 11]
    np.testing.assert_array_equal(sy[perm_y], sx)
    np.testing.assert_array_equal(sz[perm_z], sx)


def test_time_dependent_inflow():
    """solar wind, 20 % density oscillation at 10 mHz on the +x end: after
    each step the inflow cell holds proton mass x 5e6 (1 + 0.2 sin(0.01 pi t))
    of the step's end time"""
    cfg = load("solar_wind_1d_density_10mHz")
    mask, nv, vv, pv, bv = cfg.boundary_values(123.0)
    cells = np.nonzero(mask[0])[0]
    assert len(cells) == 1
    want = 5e6 * (1 + 0.2 * math.sin(0.01 * math.pi * 123.0))
    assert nv[cells[0]] == want
    st, steps, t, periods, S = O.run_bounded_config(cfg, max_steps=12)
    assert st == 0 and steps == 12
    want = cfg.proton_mass * (5e6 * (1 + 0.2 * math.sin(0.01 * math.pi * t)))
    assert S[0][cells[0]] == want


def test_nearest_time_stamp():
    stamps = [0.0, 10.0, 20.0]
    assert [boundary.nearest_stamp(stamps, t) for t in (-5, 0, 4.9, 5.0, 5.1, 15, 25)] == \
        [0, 0, 0, 0, 1, 1, 2]


def test_copy_boundary_cells_of_the_rarefaction():
    """copy cells at both ends: density, velocity and pressure of the one
    normal face neighbour (the reference's integer 1/count leaves the
    velocity factor at 1 with one source)"""
    cfg = load("rarefaction_1d_px_hll_athena")
    st, steps, t, periods, S = O.run_bounded_config(cfg, max_steps=25)
    assert st == 0
    n = cfg.n
    g = cfg.adiabatic_index
    for c, src in ((0, 1), (n - 1, n - 2)):
        assert S[0][c] == S[0][src]
        v_src = S[1:4, src] * (1 / S[0][src])
        np.testing.assert_array_equal(S[1:4, c], v_src * S[0][c])
        kin = 0.5 / S[0] * (S[1:4] ** 2).sum(0)
        mag = 0.5 / cfg.vacuum_permeability * (S[5:8] ** 2).sum(0)
        p = (S[4] - kin - mag) * (g - 1)
        assert abs(p[c] - p[src]) <= 1e-13 * abs(p[src])


def test_rejected_corner_copy_cells():
    with pytest.raises(config.ConfigError, match="no normal face"):
        load("bow_shock_2d_hll_athena").classify()
