"""GPU parity of the HIP path against the FP64 CPU oracle, through the C ABI.

FP64 results must be bit-identical to the oracle (same operation order, no
FMA contraction, IEEE division and square root).  Mixed precision (FP32
fluxes, FP64 state and pressure recovery) is checked against the FP64 run
with the tolerances stated in each test.
"""
import numpy as np
import pytest

import golden_cases as G
import oracle_lib as O
import states

pytestmark = pytest.mark.gpu

SOLVERS = [("rusanov", O.RUSANOV), ("hll-athena", O.HLL),
           ("hlld-athena", O.HLLD), ("roe-athena", O.ROE)]
KINDS = ["generic", "zero_field", "zero_bx", "supersonic", "equal",
         "strong_field", "weak_bx"]


def _to_dev(a, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("name,sid", SOLVERS)
def test_face_flux_fp64_bit_exact(gpu, name, sid, kind):
    from pamhd_amd import solver as S
    seed = 1000 + 17 * sid + KINDS.index(kind)
    L, R, bg = states.random_faces(4096, seed, kind=kind)
    Fo, so, sto = O.flux_batch(sid, L, R, bg)
    Fg, sg, stg = S.face_flux_batch(name, _to_dev(L, gpu), _to_dev(R, gpu),
                                    bg=_to_dev(bg, gpu))
    Fg, sg, stg = Fg.cpu().numpy(), sg.cpu().numpy(), stg.cpu().numpy()
    print("solver", name, "kind", kind, "errors", int((sto != 0).sum()), flush=True)
    np.testing.assert_array_equal(stg, sto)
    ok = sto == 0
    assert ok.sum() > 0.5 * len(ok) or sid == O.ROE
    np.testing.assert_array_equal(Fg[:, ok], Fo[:, ok])
    np.testing.assert_array_equal(sg[ok], so[ok])


@pytest.mark.parametrize("name,sid", SOLVERS)
def test_face_flux_mixed_precision(gpu, name, sid):
    """FP32 fluxes: per face, |F32 - F64| <= 2e-5 * max(|F64|) + 1e-7 * max|U|
    (FP32 rounding of the conserved inputs bounds the attainable accuracy)."""
    from pamhd_amd import solver as S
    L, R, bg = states.random_faces(8192, 77 + sid, kind="generic")
    Fo, so, sto = O.flux_batch(sid, L, R, bg)
    Fg, sg, stg = S.face_flux_batch(name, _to_dev(L, gpu), _to_dev(R, gpu),
                                    bg=_to_dev(bg, gpu), precision="mixed")
    Fg, sg, stg = Fg.cpu().numpy(), sg.cpu().numpy(), stg.cpu().numpy()
    ok = (sto == 0) & (stg == 0)
    assert ok.sum() > 0.5 * len(ok)
    scale = np.abs(Fo[:, ok]).max(0) + np.maximum(np.abs(L[:, ok]).max(0), np.abs(R[:, ok]).max(0))
    err = np.abs(Fg[:, ok] - Fo[:, ok]).max(0) / scale
    print("solver", name, "max rel face error", err.max(), "median", np.median(err), flush=True)
    assert err.max() < 2e-5
    serr = np.abs(sg[ok] - so[ok]) / so[ok].clip(1e-30)
    assert serr.max() < 2e-5


def _vortex(cells, seed=3):
    from pamhd_amd import configs
    return configs.orszag_tang_vortex(cells, noise=0.01, seed=seed)


@pytest.mark.parametrize("name,sid", SOLVERS)
def test_grid_steps_fp64_bit_exact(gpu, name, sid):
    """Final init + 6 main-loop steps of a 20x18x16 vortex (uneven sizes)."""
    from pamhd_amd import solver as S
    cells = (20, 18, 16)
    state, length, gamma, mu0 = _vortex(cells)
    og = O.make_grid(cells, length, gamma, mu0)
    ref = state.copy()
    st, steps_o, t_o = O.run_periodic(og, sid, 0.5, 0.0, 1e9, ref, max_steps=6)
    assert st == 0
    with S.Grid(cells, length, gamma, mu0) as g:
        assert g.kind == "periodic"
        g.upload(state)
        steps_g, t_g = g.run(name, 0.5, 0.0, 1e9, max_steps=6)
        out = g.download()
    print("solver", name, "steps", steps_g, "t", t_g, flush=True)
    assert steps_g == steps_o == 6
    assert t_g == t_o
    diff = np.abs(out - ref).max(1)
    print("max abs diff per variable", diff, flush=True)
    np.testing.assert_array_equal(out, ref)


@pytest.mark.parametrize("name,sid", SOLVERS[:2])
def test_grid_steps_face_varying_background(gpu, name, sid):
    """A random background field that differs on every face (Rusanov and HLL
    use it): each face reads Bg_B of its negative-side cell's +d face."""
    from pamhd_amd import solver as S
    cells = (12, 10, 8)
    state, length, gamma, mu0 = _vortex(cells, seed=21)
    n = cells[0] * cells[1] * cells[2]
    bg = np.random.default_rng(5).uniform(-0.2, 0.2, size=(9, n))
    og = O.make_grid(cells, length, gamma, mu0, bg)
    ref = state.copy()
    st, steps_o, t_o = O.run_periodic(og, sid, 0.3, 0.0, 1e9, ref, max_steps=4)
    assert st == 0
    with S.Grid(cells, length, gamma, mu0) as g:
        g.upload(state)
        g.set_background_faces(bg)
        steps_g, t_g = g.run(name, 0.3, 0.0, 1e9, max_steps=4)
        out = g.download()
    assert steps_g == steps_o and t_g == t_o
    np.testing.assert_array_equal(out, ref)


@pytest.mark.parametrize("variant", [(0, 0, 0), (1, 0, 1), (0, 4, 0), (1, 8, 1), (0, 32, 1),
                                     (2, 0, 2), (3, 0, 3), (4, 0, 0), (7, 0, 2),
                                     (0, 8, 4), (1, 4, 9), (0, 16, 8), (0, 8, 12), (1, 16, 17)])
def test_launch_variants_bit_exact(gpu, variant):
    """Every launch variant the autotuner may pick (block order, occupancy
    build of the flux kernels, per-cell or z-marching update with its planes
    per block) gives the oracle's bits;
    uneven sizes exercise partial tiles and chunk ends."""
    from pamhd_amd import solver as S
    cells = (70, 9, 13)
    state, length, gamma, mu0 = _vortex(cells, seed=11)
    og = O.make_grid(cells, length, gamma, mu0)
    ref = state.copy()
    st, steps_o, t_o = O.run_periodic(og, O.HLL, 0.3, 0.0, 1e9, ref, max_steps=4)
    assert st == 0
    with S.Grid(cells, length, gamma, mu0) as g:
        g.upload(state)
        g.set_launch(*variant)
        g.run("hll-athena", 0.3, 0.0, 1e9, max_steps=4)
        out = g.download()
    np.testing.assert_array_equal(out, ref)


@pytest.mark.parametrize("variant", [(8, 0, 0), (9, 0, 1), (12, 16, 1), (13, 8, 0),
                                     (16, 0, 0), (49, 16, 1), (52, 0, 0), (21, 8, 0),
                                     (64, 0, 0), (65, 8, 1), (96, 16, 0), (68, 0, 1)])
def test_hlld_shared_sides_bit_exact(gpu, variant):
    """HLLD with the per-cell part (pressure, 1/rho, fast speed) computed once
    per cell, shared through LDS or carried along marches of 16 or 4 cells
    (chunk ends inside the 9- and 13-cell lines), the x-faces too with each
    lane's part moved to the next lane (the 64-lane tile's left edge and the
    partial second tile): the oracle's bits in FP64, the per-face
    kernel's bits in mixed precision and through the z-slab pipeline; the
    70x9x13 grid leaves partial tiles in x, y and z."""
    from pamhd_amd import solver as S
    cells = (70, 9, 13)
    state, length, gamma, mu0 = _vortex(cells, seed=17)
    og = O.make_grid(cells, length, gamma, mu0)
    ref = state.copy()
    st, steps_o, t_o = O.run_periodic(og, O.HLLD, 0.3, 0.0, 1e9, ref, max_steps=4)
    assert st == 0
    outs = {}
    for prec in ("fp64", "mixed"):
        for v, slabs in (((0, 0, 0), 1), (variant, 1), (variant, 3)):
            with S.Grid(cells, length, gamma, mu0, precision=prec) as g:
                g.upload(state)
                g.set_launch(*v)
                g.set_slabs(slabs)
                g.run("hlld-athena", 0.3, 0.0, 1e9, max_steps=4)
                outs[prec, v, slabs] = g.download()
    for slabs in (1, 3):
        np.testing.assert_array_equal(outs["fp64", variant, slabs], ref)
        np.testing.assert_array_equal(outs["mixed", variant, slabs], outs["mixed", (0, 0, 0), 1])


@pytest.mark.parametrize("solver,sid", [("rusanov", O.RUSANOV), ("hll-athena", O.HLL)])
@pytest.mark.parametrize("variant", [(0, 0, 0), (16, 0, 0), (64, 0, 0), (80, 8, 1), (96, 16, 0)])
def test_hll_rusanov_cell_parts_bit_exact(gpu, solver, sid, variant):
    """HLL and Rusanov with the per-cell part (checks, pressure, fast
    speed, physical flux) computed once per cell along marches, the x-faces
    too with each lane's part moved to the next lane: the oracle's bits in
    FP64 (70x9x13: partial tiles, chunk ends), and the per-face kernel's
    bits in mixed precision"""
    from pamhd_amd import solver as S
    cells = (70, 9, 13)
    state, length, gamma, mu0 = _vortex(cells, seed=21)
    og = O.make_grid(cells, length, gamma, mu0)
    ref = state.copy()
    st, steps_o, t_o = O.run_periodic(og, sid, 0.3, 0.0, 1e9, ref, max_steps=4)
    assert st == 0
    outs = {}
    for prec in ("fp64", "mixed"):
        for v in ((0, 0, 0), variant):
            with S.Grid(cells, length, gamma, mu0, precision=prec) as g:
                g.upload(state)
                g.set_launch(*v)
                g.run(solver, 0.3, 0.0, 1e9, max_steps=4)
                outs[prec, v] = g.download()
    np.testing.assert_array_equal(outs["fp64", variant], ref)
    np.testing.assert_array_equal(outs["mixed", variant], outs["mixed", (0, 0, 0)])


@pytest.mark.parametrize("solver", ["rusanov", "hll-athena"])
@pytest.mark.parametrize("i", [5, 11])
def test_hll_rusanov_cell_parts_report_errors(gpu, solver, i):
    """a cell with negative thermal pressure stops the step with the same
    error per face or with the per-cell parts (at i = 11 it is the lane-0
    edge of the x-march)"""
    from pamhd_amd import _lib
    from pamhd_amd import solver as S
    cells = (12, 10, 9)
    state, length, gamma, mu0 = _vortex(cells)
    c = i + 12 * (4 + 10 * 3)
    state[4, c] = 0.01 * state[4, c]
    codes = []
    for flux_bits in (0, 16, 64):
        with S.Grid(cells, length, gamma, mu0) as g:
            g.upload(state)
            g.set_launch(flux_bits, 0, 0)
            with pytest.raises(_lib.PamhdError) as info:
                g.timestep(solver, 0.0, 0.3)
            codes.append(info.value.code)
    assert codes[0] == codes[1] == codes[2] != 0


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("i", [5, 11])
def test_hlld_shared_sides_report_errors(gpu, prec, i):
    """A cell with negative thermal pressure stops the HLLD step with the
    pressure error whether the per-cell part is shared or not (at i = 11 it
    is the left side of the periodic x-face at i = 0, the lane-0 edge of the
    x-march)."""
    from pamhd_amd import _lib
    from pamhd_amd import solver as S
    cells = (12, 10, 9)
    state, length, gamma, mu0 = _vortex(cells)
    c = i + 12 * (4 + 10 * 3)
    state[4, c] = 0.01 * state[4, c]
    codes = []
    for flux_bits in (0, 8, 16, 64):
        with S.Grid(cells, length, gamma, mu0, precision=prec) as g:
            g.upload(state)
            g.set_launch(flux_bits, 0, 0)
            with pytest.raises(_lib.PamhdError) as info:
                g.timestep("hlld-athena", 0.0, 0.3)
            codes.append(info.value.code)
    print("codes", codes, flush=True)
    assert codes == [2, 2, 2, 2]


@pytest.mark.parametrize("slabs", [2, 3, 4, 8])
def test_slab_pipeline_bit_exact(gpu, slabs):
    """Flux pass and update pipelined over z-slabs on two streams give the
    oracle's bits (FP64) and the unpipelined mixed-precision bits."""
    from pamhd_amd import solver as S
    cells = (20, 18, 16)
    state, length, gamma, mu0 = _vortex(cells, seed=13)
    og = O.make_grid(cells, length, gamma, mu0)
    ref = state.copy()
    st, steps_o, t_o = O.run_periodic(og, O.HLLD, 0.3, 0.0, 1e9, ref, max_steps=5)
    assert st == 0
    outs = {}
    for prec in ("fp64", "mixed"):
        for sl in (1, slabs):
            with S.Grid(cells, length, gamma, mu0, precision=prec) as g:
                g.upload(state)
                # the per-cell update: the pipeline's kernel
                g.set_launch(0, 0, 0)
                g.set_slabs(sl)
                g.run("hlld-athena", 0.3, 0.0, 1e9, max_steps=5)
                outs[prec, sl] = g.download()
    np.testing.assert_array_equal(outs["fp64", slabs], ref)
    np.testing.assert_array_equal(outs["fp64", 1], ref)
    np.testing.assert_array_equal(outs["mixed", slabs], outs["mixed", 1])


def test_autotune_keeps_results(gpu):
    from pamhd_amd import solver as S
    cells = (32, 24, 16)
    state, length, gamma, mu0 = _vortex(cells, seed=12)
    og = O.make_grid(cells, length, gamma, mu0)
    ref = state.copy()
    st, steps_o, t_o = O.run_periodic(og, O.HLLD, 0.3, 0.0, 1e9, ref, max_steps=5)
    with S.Grid(cells, length, gamma, mu0) as g:
        g.upload(state)
        report = g.autotune("hlld-athena")
        print(report, flush=True)
        assert "chosen" in report
        g.run("hlld-athena", 0.3, 0.0, 1e9, max_steps=5)
        out = g.download()
    np.testing.assert_array_equal(out, ref)


@pytest.mark.parametrize("case", sorted(G.GOLDENS))
def test_goldens_on_gpu(gpu, case):
    """The reference's uniform-grid regression runs, on the GPU: bit-identical
    to the oracle and within mhd_diff's default rule of the reference's .dc."""
    from pamhd_amd import solver as S
    cfg = G.GOLDENS[case]
    cells, length, state, bg = G.initial_state(cfg)
    og = O.make_grid(cells, length, cfg["gamma"], cfg["mu0"], bg)
    ref = state.copy()
    st, steps_o, t_o = O.run_periodic(og, O.RUSANOV, cfg["cfl"], 0.0, cfg["t_end"], ref)
    assert st == 0
    with S.Grid(cells, length, cfg["gamma"], cfg["mu0"]) as g:
        g.upload(state)
        if bg is not None:
            g.set_background_faces(bg)
        steps_g, t_g = g.run("rusanov", cfg["cfl"], 0.0, cfg["t_end"], max_steps=20000)
        out = g.download()
    assert steps_g == cfg["steps"] == steps_o
    np.testing.assert_array_equal(out, ref)
    meta, ids, mhd = G.read_dc(G.DC_DIR + "/" + case + ".dc")
    gold = G.dc_to_grid(meta, ids, mhd)
    mine = np.concatenate([out[:5], G.volume_field(cells, out)])
    res = G.mhd_diff(mine, gold)
    print(case, res, flush=True)
    for k, (viol, rel) in res.items():
        assert viol <= 1e-15, (k, viol, rel)


def test_mixed_precision_solution_error(gpu):
    """HLLD, 32^3 vortex, 60 steps at time-step factor 0.3: relative L2 error
    of each conserved variable (rho, momentum vector, E, face-field vector) of
    the mixed run against the FP64 run is at most 1e-6.  Vector variables are
    normalised by the vector's norm: m_z and B_z of this quasi-2-D state are
    roundoff-sized in the FP64 run."""
    from pamhd_amd import solver as S
    cells = (32, 32, 32)
    state, length, gamma, mu0 = _vortex(cells, seed=5)
    outs = {}
    for prec in ("fp64", "mixed"):
        with S.Grid(cells, length, gamma, mu0, precision=prec) as g:
            g.upload(state)
            g.run("hlld-athena", 0.3, 0.0, 1e9, max_steps=60)
            outs[prec] = g.download()
    a, b = outs["mixed"], outs["fp64"]
    rel = {}
    for name, rows in (("rho", [0]), ("m", [1, 2, 3]), ("E", [4]), ("B", [5, 6, 7])):
        rel[name] = float(np.sqrt(((a[rows] - b[rows]) ** 2).sum()) / np.sqrt((b[rows] ** 2).sum()))
    print("relative L2 error per variable", rel, flush=True)
    assert max(rel.values()) <= 1e-6


def test_conservation_and_divergence(gpu):
    """Periodic 64^3 HLLD run: total mass and momentum conserved to roundoff
    (the constant-pressure energy refit of the reference's update_vol_B does
    not conserve total energy), face divergence stays at roundoff level."""
    from pamhd_amd import solver as S
    cells = (64, 64, 64)
    state, length, gamma, mu0 = _vortex(cells, seed=9)
    with S.Grid(cells, length, gamma, mu0) as g:
        g.upload(state)
        div0 = g.div_face2volume()
        g.run("hlld-athena", 0.3, 0.0, 1e9, max_steps=20)
        out = g.download()
        div1 = g.div_face2volume()
    tot0 = state[:4].sum(1)
    tot1 = out[:4].sum(1)
    scale = np.abs(state[:4]).sum(1).max()
    print("conservation drift", (tot1 - tot0) / scale, "div", div0, div1, flush=True)
    assert np.all(np.abs(tot1 - tot0) <= 1e-13 * scale)
    assert div1 < 1e-12


def test_physics_error_is_reported(gpu):
    """A cell with negative density makes the step fail with the density
    error and a message naming the face (the reference aborts there)."""
    from pamhd_amd import _lib
    from pamhd_amd import solver as S
    cells = (8, 8, 8)
    state, length, gamma, mu0 = _vortex(cells)
    state[0, 3 + 8 * (4 + 8 * 5)] = -1.0
    with S.Grid(cells, length, gamma, mu0) as g:
        g.upload(state)
        with pytest.raises(_lib.PamhdError) as info:
            g.timestep("hll-athena", 0.0, 0.5)
        assert info.value.code == 1
        assert "density" in str(info.value)
