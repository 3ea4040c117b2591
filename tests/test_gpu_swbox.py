"""Solar-wind boxes and ensembles on the GPU against the FP64 oracle's
restatement of tests/mhd/solar_wind_box.cpp, through the C ABI.

Every kernel evaluates each value in the oracle's operation order (no FMA
contraction, IEEE division and square root), so results are compared bit for
bit: classification, background field, the initial state after the
reference's init sequence, and states after runs of main-loop steps.
"""
import numpy as np
import pytest

import oracle_lib as O
from pamhd_amd import configs

pytestmark = pytest.mark.gpu

SOLVERS = [("rusanov", O.RUSANOV), ("hll-athena", O.HLL), ("hlld-athena", O.HLLD),
           ("roe-athena", O.ROE)]


def _test_box(cells=(12, 10, 14), direction="+x", seed=1):
    """Small box with unsolved cells inside a 2.2-cell planet, an off-centre
    dipole clamped within 0.7 cells and a tilted field."""
    rng = np.random.default_rng(seed)
    sb = configs.solar_wind_box(cells[0], float(rng.uniform(2e6, 9e6)), float(rng.uniform(3e5, 7e5)),
                                float(rng.uniform(-8e-9, 8e-9)), float(rng.uniform(-8e-9, 8e-9)))
    h = sb.cell_length[0]
    sb.cells = tuple(cells)
    sb.start = tuple(-c * h / 2 for c in cells)
    sb.direction = direction
    # planet cells must not touch the boundary layer's copy sources
    sb.inner_radius = min(2.2, min(cells) / 2 - 2.5) * h
    sb.dipoles = [((3e18, -1e18, -9e18), (0.3 * h, -0.2 * h, 0.1 * h))]
    sb.bg_min_distance = 0.7 * h
    sb.bg_value = (1e-9, 0.0, -2e-9)
    if direction[1] != "x":
        dim = "xyz".index(direction[1])
        v = [0.0, 0.0, 0.0]
        v[dim] = -sb.velocity[0] if direction[0] == "-" else sb.velocity[0]
        sb.velocity = tuple(v)
        B = [1e-9, -2e-9, 3e-9]
        B[dim] = 0.0
        sb.magnetic_field = tuple(B)
    return sb


def test_box_setup_matches_oracle(gpu):
    from pamhd_amd.swbox import BoxGrid
    sb = _test_box()
    ob = O.Box(sb)
    with BoxGrid(sb) as g:
        assert g.kind == "solar wind box"
        ct, ru = g.cell_types()
        bg = g.background()
        s = g.download()
    np.testing.assert_array_equal(ct, ob.ctype)
    np.testing.assert_array_equal(ru, ob.rule)
    assert (ct == -1).sum() > 0
    np.testing.assert_array_equal(bg, ob.bg)
    np.testing.assert_array_equal(s, ob.init())


@pytest.mark.parametrize("name,sid", SOLVERS[:3])
def test_box_run_bit_exact(gpu, name, sid):
    """Final-init step and 8 main-loop steps with boundaries."""
    from pamhd_amd.swbox import BoxGrid
    sb = _test_box(seed=2)
    ob = O.Box(sb)
    ref = ob.init()
    st, steps_o, t_o = ob.run(sid, 0.3, 0.0, 1e30, ref, max_steps=8)
    assert st == 0 and steps_o == 8
    with BoxGrid(sb) as g:
        steps_g, t_g = g.run(name, 0.3, 0.0, 1e30, max_steps=8)
        out = g.download()
    print(name, "t", t_g, "max |diff| per component", np.abs(out - ref).max(1), flush=True)
    assert steps_g == 8 and t_g == t_o
    np.testing.assert_array_equal(out, ref)


def test_box_roe_run_bit_exact(gpu):
    """Roe with the wind's field in the background only: 8 steps, bit for bit."""
    from pamhd_amd.swbox import BoxGrid
    sb = _test_box(seed=2)
    ob = O.Box(sb)
    ref = ob.init()
    st, steps_o, t_o = ob.run(O.ROE, 0.3, 0.0, 1e30, ref, max_steps=8)
    assert st == 0 and steps_o == 8
    with BoxGrid(sb) as g:
        steps_g, t_g = g.run("roe-athena", 0.3, 0.0, 1e30, max_steps=8)
        out = g.download()
    assert steps_g == 8 and t_g == t_o
    np.testing.assert_array_equal(out, ref)


def test_box_roe_reports_the_oracles_error(gpu):
    """With a field of a few nT in the wind's own state, Roe's intermediate
    state of this cold, fast wind is unphysical at the first solved step (the
    reference throws); the device path stops with the same error class as
    the oracle."""
    from pamhd_amd import _lib
    from pamhd_amd.swbox import BoxGrid
    sb = _test_box(seed=2)
    sb.magnetic_field = (0.0, -5e-9, 4e-9)
    ob = O.Box(sb)
    ref = ob.init()
    st, steps_o, t_o = ob.run(O.ROE, 0.3, 0.0, 1e30, ref, max_steps=8)
    assert st != 0
    with BoxGrid(sb) as g:
        with pytest.raises(_lib.PamhdError) as e:
            g.run("roe-athena", 0.3, 0.0, 1e30, max_steps=8)
    assert e.value.code == st


@pytest.mark.parametrize("direction", ["-x", "+y", "-z"])
def test_box_inflow_directions(gpu, direction):
    from pamhd_amd.swbox import BoxGrid
    sb = _test_box(cells=(10, 12, 11), direction=direction, seed=3)
    ob = O.Box(sb)
    ref = ob.init()
    st, steps_o, t_o = ob.run(O.HLL, 0.3, 0.0, 1e30, ref, max_steps=6)
    assert st == 0
    with BoxGrid(sb) as g:
        g.run("hll-athena", 0.3, 0.0, 1e30, max_steps=6)
        out = g.download()
    np.testing.assert_array_equal(out, ref)


def _ensemble_boxes():
    sizes = [(12, 10, 14), (8, 8, 8), (20, 6, 9), (16, 16, 16), (6, 11, 7)]
    return [_test_box(cells=c, seed=10 + q) for q, c in enumerate(sizes)]


@pytest.mark.parametrize("use_graph,tile,nsteps", [
    (False, 0, 7), (True, 0, 7), (True, 1, 7), (True, 2, 7),
    ("persistent", 0, 7), ("persistent", 1, 7), ("persistent", 2, 7),
    (False, 1, 34), (True, 1, 34), ("persistent", 1, 34)])
def test_ensemble_matches_single_boxes(gpu, use_graph, tile, nsteps):
    """Boxes of different sizes advanced together (each with its own time
    step) give every box's oracle run bit for bit, with and without the
    captured graph, in one persistent launch, and for every tile shape.
    Consecutive steps of one call leave their last face sync to the next
    step's update: 7 steps run as a chain of 1 before a graph of 2 steps
    replayed three times (or as one chain of 7), 34 steps as a chain of 2
    before a graph of 16 steps replayed twice (or one chain of 34)."""
    from pamhd_amd.swbox import Ensemble
    boxes = _ensemble_boxes()
    refs = []
    for sb in boxes:
        ob = O.Box(sb)
        r = ob.init()
        st, steps, t = ob.run(O.HLL, 0.3, 0.0, 1e30, r, max_steps=nsteps)
        assert st == 0 and steps == nsteps
        refs.append(r)
    with Ensemble(boxes) as e:
        e.set_tile(tile)
        e.timestep("hll-athena", 0.0, 0.3)
        if use_graph == "persistent":
            e.run("hll-athena", 0.3, nsteps, persistent=True)
        else:
            e.run("hll-athena", 0.3, nsteps, use_graph=use_graph)
        dts = e.check()
        outs = [e.download(q) for q in range(len(boxes))]
    assert np.all(dts > 0)
    for q, (o, r) in enumerate(zip(outs, refs)):
        np.testing.assert_array_equal(o, r, err_msg="box %d" % q)


def test_persistent_run_gives_up_on_a_barrier_timeout(gpu):
    """A persistent run whose barrier waits may poll only once: the first
    workgroup to arrive gives up and sets the error flag, every workgroup
    leaves, and the run returns an error instead of waiting; the set (and
    the GPU) then runs again normally."""
    from pamhd_amd import _lib
    from pamhd_amd.swbox import Ensemble
    boxes = _ensemble_boxes()
    with Ensemble(boxes) as e:
        e.timestep("hll-athena", 0.0, 0.3)
        e.set_barrier_polls(1)
        with pytest.raises(_lib.PamhdError, match="barrier timed out"):
            e.run("hll-athena", 0.3, 3, persistent=True)
    refs = []
    for sb in boxes:
        ob = O.Box(sb)
        r = ob.init()
        st, steps, t = ob.run(O.HLL, 0.3, 0.0, 1e30, r, max_steps=3)
        assert st == 0
        refs.append(r)
    with Ensemble(boxes) as e:
        e.timestep("hll-athena", 0.0, 0.3)
        e.run("hll-athena", 0.3, 3, persistent=True)
        outs = [e.download(q) for q in range(len(boxes))]
    for q, (o, r) in enumerate(zip(outs, refs)):
        np.testing.assert_array_equal(o, r, err_msg="box %d" % q)


def test_ensemble_autotune_keeps_state(gpu):
    from pamhd_amd.swbox import Ensemble
    boxes = _ensemble_boxes()[:3]
    refs = []
    for sb in boxes:
        ob = O.Box(sb)
        r = ob.init()
        ob.run(O.RUSANOV, 0.3, 0.0, 1e30, r, max_steps=4)
        refs.append(r)
    with Ensemble(boxes) as e:
        e.timestep("rusanov", 0.0, 0.3)
        e.run("rusanov", 0.3, 2, use_graph=False)
        report = e.autotune("rusanov", 0.3)
        print(report, flush=True)
        assert "chosen" in report
        e.run("rusanov", 0.3, 2, use_graph=False)
        outs = [e.download(q) for q in range(len(boxes))]
    for o, r in zip(outs, refs):
        np.testing.assert_array_equal(o, r)


def test_config_box_runs_and_keeps_inflow(gpu):
    """A 32^3 box of the benchmark ensemble: 40 HLL steps stay physical, the
    inflow wall keeps the solar-wind state and the interior divergence stays
    at roundoff."""
    from pamhd_amd.swbox import BoxGrid
    sb = configs.solar_wind_ensemble(nbox=1, seed=7, edges=(32,), imf=False)[0]
    with BoxGrid(sb) as g:
        steps, t = g.run("hll-athena", 0.3, 0.0, 1e30, max_steps=40)
        s = g.download()
        ct, ru = g.cell_types()
    assert steps == 40 and t > 0
    assert np.all(s[0][ct >= 0] > 0)
    sw = ru == O.RULE_SW
    np.testing.assert_array_equal(s[0][sw], sb.number_density * sb.proton_mass)
