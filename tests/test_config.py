"""The JSON configuration loader and its expression language.

Pinned by
* the reference's own regression results: the eight uniform-grid golden
  configurations (copied under tests/golden/config) loaded, initialised and
  run in the oracle reproduce the reference's .dc files under
  tests/mhd/mhd_diff.py's rule, with the reference's step counts;
* closed forms: operator precedence and associativity worked out by hand,
  Gauss-Legendre face averages against exact integrals, the geometry
  overlap tests against brute-force distances, the grid options against
  their arithmetic;
* the reference's quirks stated where they matter (fluid regions read the
  value of region number geometry-id).
The expression language is muparserx's (an absent, unpinned dependency):
what the configurations never exercise stays unpinned.
"""
import json
import math
import os

import numpy as np
import pytest

import golden_cases as G
import oracle_lib as O
from pamhd_amd import config
from pamhd_amd.expr import Expression, ExpressionError, evaluate

CFG_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "config")


def cfg_path(name):
    return os.path.join(CFG_DIR, name + ".json")


# --- expressions ----------------------------------------------------------

@pytest.mark.parametrize("text,value", [
    ("1 + 2 * 3", 7.0),
    ("(1 + 2) * 3", 9.0),
    ("-2^2", -4.0),                 # a sign binds looser than ^
    ("2^3^2", 512.0),               # ^ is right associative
    ("2 - 3 - 4", -5.0),            # - is left associative
    ("8 / 4 / 2", 1.0),
    ("1/2 + 1/4*8", 2.5),           # / divides in floating point
    ("3.0/5.0 * 9e-4", 3.0 / 5.0 * 9e-4),
    ("1 < 2 ? 3 : 4", 3.0),
    ("1 > 2 ? 3 : 1 > 0 ? 5 : 6", 5.0),   # ?: nests to the right
    ("1 > 2 or 3 > 2 and 1 > 0 ? 1 : 0", 1.0),
    ("1 > 2 or 3 > 2 and 1 > 2 ? 1 : 0", 0.0),   # and binds tighter than or
    ("{4, 5, 6}[2]", 6.0),
    ("{1, 2, 3}[0] + {1, 2, 3}[1] * 2", 5.0),
    ("x < 0.5 ? -1 : +1", -1.0),
    ("max(1, 3, 2) + min(4, -1)", 2.0),
    ("fmod(7, 3)", 1.0),
    ("abs(-2.5) + sqrt(16)", 6.5),
    ("pi", math.pi),
    ("true ? 1 : 2", 1.0),
])
def test_expression_values(text, value):
    assert float(evaluate(text, x=0.25)) == value


def test_functions_are_the_c_librarys():
    x = np.linspace(-3, 3, 101)
    for name in ("sin", "cos", "tanh", "cosh", "exp", "atan"):
        got = evaluate("%s(x)" % name, x=x)
        want = np.array([getattr(math, name)(v) for v in x])
        np.testing.assert_array_equal(got, want)
    got = evaluate("x^3 + 2^x", x=x)
    np.testing.assert_array_equal(got, np.array([math.pow(v, 3) + math.pow(2, v) for v in x]))


def test_vectors_and_cells():
    v = evaluate("{sin(2*pi*x), x < 0.5 ? 1e-9 : -1e-9, 0}", x=np.array([0.25, 0.75]))
    assert len(v) == 3
    np.testing.assert_array_equal(v[0], [math.sin(2 * math.pi * 0.25), math.sin(2 * math.pi * 0.75)])
    np.testing.assert_array_equal(v[1], [1e-9, -1e-9])
    # grid-options start of the magnetosphere configurations
    s = evaluate("{-(1/2 + 1/4*cells[0]) * volume[0] / cells[0], -volume[1] / 2, -volume[2] / 2}",
                 cells=(66.0, 50.0, 1.0), volume=(2.0, 3.0, 4.0))
    assert s == (-(0.5 + 0.25 * 66) * 2.0 / 66, -1.5, -2.0)
    assert Expression("{radius < 0.1 or x > 0 ? 0 : 1, y, z}").variables() == {"radius", "x", "y", "z"}


@pytest.mark.parametrize("text", ["1 +", "sin(1", "foo(1)", "{1, 2", "1 ? 2", "x $ 2", "{1,2}[5]"])
def test_expression_errors(text):
    with pytest.raises(ExpressionError):
        evaluate(text, x=1.0)


# --- configurations --------------------------------------------------------

GOLDEN_CONFIGS = sorted(G.GOLDENS)


def test_golden_configs_match_their_restated_initial_states():
    """the loader against the hand-restated initial states of the goldens
    (the same up to the face synchronisation, which moves the copies of the
    2-D cases by parts in 1e19)"""
    for case in GOLDEN_CONFIGS:
        cfg = config.load(cfg_path(case))
        state, bg = cfg.initial_state()
        cells, length, st0, bg0 = G.initial_state(G.GOLDENS[case])
        assert cfg.cells == cells
        np.testing.assert_allclose(cfg.cell_length, length, rtol=0, atol=0)
        np.testing.assert_array_equal(state[:5], st0[:5])
        np.testing.assert_allclose(state[5:], st0[5:], rtol=0, atol=1e-18)
        assert (bg is None) == (bg0 is None)
        if bg is not None:
            np.testing.assert_array_equal(bg, bg0)


@pytest.mark.parametrize("case", GOLDEN_CONFIGS)
def test_golden_configs_reproduce_the_reference_results(case):
    """config file -> initial state -> the test.cpp loop with the default
    substep options (0, 999) -> the reference's .dc result"""
    cfg = config.load(cfg_path(case))
    assert (cfg.substep_min, cfg.substep_max) == (0, 999)
    state, bg = cfg.initial_state()
    st, steps, t, periods = O.run_config(cfg, state, bg)
    assert st == 0 and max(periods) == 1
    meta, ids, mhd = G.read_dc(os.path.join(G.DC_DIR, case + ".dc"))
    assert steps == meta["sim_step"]
    assert meta["sim_time"] == t
    gold = G.dc_to_grid(meta, ids, mhd)
    mine = np.concatenate([state[:5], G.volume_field(cfg.cells, state)])
    for k, (viol, rel) in G.mhd_diff(mine, gold).items():
        assert viol <= 1e-15, (case, k, viol, rel)


def test_grid_options_arithmetic():
    obj = json.load(open(cfg_path("alfven_2d_x_1")))
    obj["grid-options"] = {
        "periodic": "{true, true, true}",
        "cells": "{2*15 + 2, 3*15 + 2, 1}",
        "volume": "{2 * (1 + 2 / (cells[0] - 2)), 3 * (1 + 2 / (cells[1] - 2)), 2}",
        "start": "{-(1/2 + 1/4*cells[0]) * volume[0] / cells[0], -volume[1] / 2, -volume[2] / 2}",
    }
    cfg = config.load(obj)
    assert cfg.cells == (32, 47, 1)
    vol = (2 * (1 + 2 / 30), 3 * (1 + 2 / 45), 2.0)
    assert cfg.volume == vol
    assert cfg.start == (-(0.5 + 0.25 * 32) * vol[0] / 32, -vol[1] / 2, -1.0)
    assert cfg.cell_length == (vol[0] / 32, vol[1] / 47, 2.0)


def test_face_averages_are_exact_integrals():
    """B = (cos 2 pi y, cos 2 pi z, sin 2 pi x) averaged over the faces of a
    3-D grid: each face average has the closed form
    (sin(2 pi b) - sin(2 pi a)) / (2 pi h) over its one varying direction"""
    obj = json.load(open(cfg_path("alfven_2d_x_1")))
    obj["grid-options"] = {"periodic": "{true, true, true}", "cells": "{6, 5, 4}",
                           "volume": "{1, 1, 1}", "start": "{0, 0, 0}"}
    obj["magnetic-field"] = {"default": "{cos(2*pi*y), cos(2*pi*z), sin(2*pi*x)}"}
    cfg = config.load(obj)
    state, _ = cfg.initial_state()
    lo, hi, _ = cfg.cell_bounds()

    def avg_cos(a, b):
        return (np.sin(2 * np.pi * b) - np.sin(2 * np.pi * a)) / (2 * np.pi * (b - a))

    def avg_sin(a, b):
        return -(np.cos(2 * np.pi * b) - np.cos(2 * np.pi * a)) / (2 * np.pi * (b - a))
    np.testing.assert_allclose(state[5], avg_cos(lo[1], hi[1]), rtol=0, atol=1e-14)
    np.testing.assert_allclose(state[6], avg_cos(lo[2], hi[2]), rtol=0, atol=1e-14)
    np.testing.assert_allclose(state[7], avg_sin(lo[0], hi[0]), rtol=0, atol=1e-14)
    # the field is divergence-free: so is its face average, to roundoff
    g = O.make_grid(cfg.cells, cfg.cell_length, cfg.adiabatic_index, cfg.vacuum_permeability)
    assert O.div_face2volume(g, state) < 1e-13


def test_blast_wave_region_by_overlap():
    """pressure 10 in the cells overlapping the sphere of radius 0.1001 at
    the origin (the distance from the centre to the cell box), 0.1 elsewhere"""
    cfg = config.load(cfg_path("blast_wave_2d_hll_athena"))
    S, _, _ = cfg.bounded_initial_state()      # z is closed
    state = np.concatenate([S[:5], S[8:14:2]])
    lo, hi, _ = cfg.cell_bounds()
    d2 = np.zeros(cfg.n)
    for d in range(3):
        gap = np.maximum(np.maximum(lo[d] - 0.0, 0.0), np.maximum(0.0 - hi[d], 0.0))
        d2 += gap * gap
    inside = d2 < 0.1001 ** 2
    assert 0 < inside.sum() < cfg.n
    g = cfg.adiabatic_index
    rho = state[0]
    mag = 0.5 / cfg.vacuum_permeability * (S[5:8] ** 2).sum(0)
    kin = 0.5 / rho * (state[1:4] ** 2).sum(0)
    p = (state[4] - kin - mag) * (g - 1)
    np.testing.assert_allclose(p[inside], 10.0, rtol=1e-14)
    np.testing.assert_allclose(p[~inside], 0.1, rtol=1e-14)


def test_fluid_regions_read_the_region_numbered_by_geometry_id():
    """initialize_fluid (source/mhd/initialize.hpp:183-195) applies region i
    to the cells of its geometry with the value of region number
    geometry-id; the magnetic field regions use their own value"""
    obj = json.load(open(cfg_path("blast_wave_2d_hll_athena")))
    obj["grid-options"]["periodic"] = "{true, true, true}"
    obj["geometries"] = [{"box": {"start": [-9, -9, -9], "end": [0, 9, 9]}},
                         {"box": {"start": [0, -9, -9], "end": [9, 9, 9]}}]
    obj["number-density"] = {"default": 1, "initial-conditions": [
        {"geometry-id": 1, "value": 2}, {"geometry-id": 0, "value": 3}]}
    cfg = config.load(obj)
    state, _ = cfg.initial_state()
    _, _, ctr = cfg.cell_bounds()
    left = ctr[0] < 0
    # region 0 (geometry 1, x > 0) takes region 1's value 3; region 1
    # (geometry 0, x < 0) takes region 0's value 2
    np.testing.assert_array_equal(state[0][~left], 3.0)
    np.testing.assert_array_equal(state[0][left], 2.0)


def test_substep_options_per_cell():
    cfg = config.load(cfg_path("substep_x1"))
    lo, hi = cfg.substep_options(0.0)
    _, _, ctr = cfg.cell_bounds()
    want = np.where(ctr[0] < 0.5, 0, 1)
    np.testing.assert_array_equal(lo, want)
    np.testing.assert_array_equal(hi, want)
    assert not cfg.substeps_depend_on_time()
    obj = json.load(open(cfg_path("substep_x1")))
    obj["substep-max"] = "t > 0.5 ? 2 : 1"
    cfg = config.load(obj)
    assert cfg.substeps_depend_on_time()
    assert cfg.substep_options(0.0)[1].max() == 1 and cfg.substep_options(0.7)[1].min() == 2


@pytest.mark.parametrize("name", ["substep_x1", "substep_x2", "substep_y1", "substep_z2"])
def test_substep_configs_run_in_the_oracle(name):
    """the reference's substepping configurations: cells with substep
    options 1 run at half the pace of their neighbours; mass is conserved"""
    cfg = config.load(cfg_path(name))
    state, bg = cfg.initial_state()
    m0 = state[0].sum()
    st, steps, t, periods = O.run_config(cfg, state, bg, max_steps=40)
    assert st == 0 and steps > 0 and max(periods) == 2
    assert abs(state[0].sum() - m0) <= 1e-13 * m0


@pytest.mark.parametrize("edit,match", [
    (lambda o: o["grid-options"].update({"max-ref-lvl": 1}), "refinement"),
    (lambda o: o.update({"solver-mhd": "rusanov-staggered"}), "unsupported solver"),
    (lambda o: o["pressure"].update({"default": {"x": [0, 1], "y": [0], "z": [0], "data": [1, 2]}}),
     "point-cloud"),
    (lambda o: o["pressure"].update({"value-boundaries": [{"geometry-id": 0, "value": 1}]}),
     "time-stamps"),
    (lambda o: o.update({"substep-max": -1}), "negative"),
    (lambda o: o.update({"background-magnetic-field": {"dipoles": [{"moment": [1, 2]}]}}),
     "dipoles"),
])
def test_unsupported_configurations_are_rejected(edit, match):
    obj = json.load(open(cfg_path("alfven_1d_px")))
    edit(obj)
    with pytest.raises(config.ConfigError, match=match):
        config.load(obj)


def test_two_population_configurations_are_rejected():
    with pytest.raises(config.ConfigError, match="two-population"):
        config.load(cfg_path("two_pop1"))


def test_all_fixture_configs_initialise():
    for f in sorted(os.listdir(CFG_DIR)):
        if f.startswith("two_pop") or f.startswith("bow_shock"):
            continue
        cfg = config.load(os.path.join(CFG_DIR, f))
        if cfg.bounded:
            state, bg, types = cfg.bounded_initial_state()
            assert state.shape == (14, cfg.n)
        else:
            state, bg = cfg.initial_state()
            assert state.shape == (8, cfg.n)
        assert np.all(np.isfinite(state)) and np.all(state[0] > 0) and np.all(state[4] > 0), f


def _bg_config(bg):
    obj = json.load(open(cfg_path("alfven_2d_x_1")))
    obj["background-magnetic-field"] = bg
    return config.load(obj)


def test_point_dipole_background():
    """B = mu0 / (4 pi r^3) (3 (m . r^) r^ - m) outside min-distance,
    mu0 m / (2 pi d^3) inside it, plus the constant"""
    m = (1e15, -2e15, 3e15)
    cfg = _bg_config({"value": [1e-9, 0, 0], "min-distance": 0.5,
                      "dipoles": [{"moment": list(m), "position": [0.1, 0.2, -0.3]}]})
    rng = np.random.default_rng(3)
    p = rng.uniform(-3, 3, (3, 200))
    B = cfg.background_field(*p)
    r = p - np.array([[0.1], [0.2], [-0.3]])
    r1 = np.sqrt((r * r).sum(0))
    u = r / r1
    mv = np.array(m)[:, None]
    far = cfg.vacuum_permeability / (4 * np.pi * r1 ** 3) * (3 * (mv * u).sum(0) * u - mv)
    near = cfg.vacuum_permeability * mv / (2 * np.pi * 0.5 ** 3) * np.ones_like(r1)
    want = np.where(r1 < 0.5, near, far) + np.array([[1e-9], [0], [0]])
    np.testing.assert_allclose(B, want, rtol=1e-12, atol=1e-30)


@pytest.mark.parametrize("dim,fdir", [(1, 2), (1, -3), (2, 1), (2, 3), (3, -1), (3, 2),
                                      (1, 1), (2, -2), (3, 3)])
def test_line_dipole_background(dim, fdir):
    """the line dipole of background_magnetic_field.hpp:338-458: with
    k = mu0 m / (2 pi r^4) and (a, b) the coordinates across the line
    (r = position - dipole position), a field through the line in a
    direction across it has the components k a b and k (b^2 - a^2) (in the
    reference's assignment of the two to the two directions), one along the
    line -k r / 2; it does not depend on the coordinate along the line,
    scales as 1 / r^2 (1 / r^3 along the line), changes sign with the
    direction (outside min-distance) and is the clamped value inside it.  (It is the
    reference's expression, not a potential field.)"""
    mu0 = 1.2566370614359173e-06
    obj = json.load(open(cfg_path("alfven_2d_x_1")))
    obj["vacuum-permeability"] = mu0
    pos = np.array([0.3, -0.2, 0.1])

    def cfg_for(f):
        obj["background-magnetic-field"] = {"min-distance": 0.1, "line-dipoles": [
            {"line-dimension": dim, "field-direction-through-line": f, "magnitude": 1e12,
             "position": list(pos)}]}
        return config.load(obj)
    cfg = cfg_for(fdir)
    rng = np.random.default_rng(dim * 10 + fdir + 5)
    p = rng.uniform(1.0, 2.0, (3, 50)) * rng.choice([-1, 1], (3, 50))
    B = cfg.background_field(*p)
    others = [a for a in range(3) if a != dim - 1]
    r = p - pos[:, None]
    r[dim - 1] = 0
    r1 = np.sqrt((r * r).sum(0))
    k = mu0 * 1e12 / (2 * np.pi * r1 ** 4)
    want = np.zeros((3, 50))
    c1, c2 = others
    if abs(fdir) == dim:
        want[dim - 1] = -k * r1 / 2
    else:
        v1, v2 = (c2, c1) if abs(fdir) - 1 == c1 else (c1, c2)
        want[v1] = k * r[c1] * r[c2]
        want[v2] = k * (r[c2] ** 2 - r[c1] ** 2)
    want *= np.sign(fdir)
    np.testing.assert_allclose(B, want, rtol=1e-12, atol=1e-40)
    q = p.copy()
    q[dim - 1] += 7.5
    np.testing.assert_array_equal(cfg.background_field(*q), B)
    for a in others:
        q[a] = pos[a] + 2 * (p[a] - pos[a])
    np.testing.assert_allclose(cfg.background_field(*q), B / (8 if abs(fdir) == dim else 4),
                               rtol=1e-12, atol=1e-40)
    np.testing.assert_allclose(cfg_for(-fdir).background_field(*p), -B, rtol=0, atol=0)
    inside = cfg.background_field(pos[0] + 0.01, pos[1] + 0.02, pos[2] - 0.03)
    clamp = np.zeros(3)
    # along the line the clamped value keeps its sign (the reference adds it
    # whatever the direction's sign)
    sign = 1 if abs(fdir) == dim else np.sign(fdir)
    clamp[abs(fdir) - 1] = sign * mu0 * 1e12 / (2 * np.pi * 0.1 ** 3)
    np.testing.assert_allclose(inside, clamp, rtol=1e-14)
