"""HLLD's step structures against the FP64 oracle and against each other,
through the C ABI: the per-direction flux passes and update pass
("passes"), the fused single pass ("fused": every face flux and the
conserved update in one pass) and the accumulating marches ("accumulate":
x, y and z marches each adding their faces' flux differences to the new
state), the latter two followed by the face-field update and energy refits.

The fused pass tiles x by 64 cells, y by 6 rows with a row on each side and
z by chunks of planes; the accumulating marches tile x by 64 lanes and walk
km cells along y or z; every index wraps periodically, so the grid shapes
below cover tiles cut short in each dimension, grids smaller than one tile
(lanes and waves wrapping several times), thin and 2-D grids, and chunks of
1 to 32 planes.  FP64 results must be bit-identical to the oracle; the
mixed path must give the same bits in every structure (the same FP32 face
arithmetic and FP64 accumulation in the same order).
"""
import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu


def _state(cells, seed):
    from pamhd_amd import configs
    return configs.orszag_tang_vortex(cells, noise=0.02, seed=seed)


def _run(cells, steps, fused, kc=0, precision="fp64", seed=5, state=None):
    from pamhd_amd import solver as S
    st, length, gamma, mu0 = _state(cells, seed) if state is None else state
    with S.Grid(cells, length, gamma, mu0, precision=precision) as g:
        g.set_fused_step(fused, kc)
        g.upload(st)
        n, t = g.run("hlld-athena", 0.3, 0.0, 1e9, max_steps=steps)
        out = g.download()
        speeds = g.max_signal_speeds()
        fc = g.last_flux_calcs()
    return out, n, t, speeds, fc


SHAPES = [(64, 12, 10), (70, 13, 33), (130, 7, 5), (16, 12, 10), (5, 3, 4),
          (64, 6, 1), (40, 40, 1), (1, 1, 8), (128, 18, 40)]


@pytest.mark.parametrize("cells", SHAPES)
@pytest.mark.parametrize("mode", ["fused", "accumulate"])
def test_fused_fp64_bit_exact_vs_oracle(gpu, cells, mode):
    print("cells", cells, mode, flush=True)
    state, length, gamma, mu0 = _state(cells, 5)
    ref = state.copy()
    og = O.make_grid(cells, length, gamma, mu0)
    st, steps_o, t_o = O.run_periodic(og, O.HLLD, 0.3, 0.0, 1e9, ref, max_steps=4)
    assert st == 0
    out, n, t, _, fc = _run(cells, 4, mode)
    assert n == steps_o == 4 and t == t_o
    np.testing.assert_array_equal(out, ref)
    # every face once per step (timestep()'s flux_calcs)
    assert fc == O.last_flux_calcs() == 3 * cells[0] * cells[1] * cells[2]


@pytest.mark.parametrize("mode,kc", [("fused", 1), ("fused", 3), ("fused", 7), ("fused", 32),
                                     ("accumulate", 1), ("accumulate", 3), ("accumulate", 7),
                                     ("accumulate", 32)])
def test_fused_chunks_match_two_pass(gpu, mode, kc):
    cells = (66, 14, 37)
    print("cells", cells, mode, "planes per chunk", kc, flush=True)
    a, na, ta, sa, _ = _run(cells, 5, mode, kc)
    b, nb, tb, sb, _ = _run(cells, 5, "passes")
    assert na == nb and ta == tb
    np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(sa, sb)


@pytest.mark.parametrize("cells", [(64, 12, 10), (70, 13, 33)])
def test_fused_mixed_matches_two_pass_mixed(gpu, cells):
    print("cells", cells, flush=True)
    b, nb, tb, _, _ = _run(cells, 6, "passes", precision="mixed")
    for mode in ("fused", "accumulate"):
        a, na, ta, _, _ = _run(cells, 6, mode, precision="mixed")
        assert na == nb and ta == tb
        np.testing.assert_array_equal(a, b)


def test_fused_reports_the_failing_face(gpu):
    """A cell of negative thermal pressure: every step structure fails with
    the pressure error on a face of that cell (which of the cell's faces is
    reported first depends on the order the faces are computed in)."""
    from pamhd_amd import _lib
    from pamhd_amd import solver as S
    cells = (64, 12, 10)
    state, length, gamma, mu0 = _state(cells, 9)
    i, j, k = 5, 7, 4
    c = i + 64 * (j + 12 * k)
    state[4, c] = 0.0                 # E = 0: negative thermal pressure
    for mode in ("fused", "accumulate", "passes"):
        with S.Grid(cells, length, gamma, mu0) as g:
            g.set_fused_step(mode)
            g.upload(state)
            with pytest.raises(_lib.PamhdError) as ei:
                g.run("hlld-athena", 0.3, 0.0, 1e9, max_steps=2)
        msg = str(ei.value)
        print(mode, msg, flush=True)
        assert ei.value.code == 2
        named = [(i, j, k), (i + 1, j, k), (i, j + 1, k), (i, j, k + 1)]
        assert any("(%d,%d,%d)" % p in msg for p in named)


@pytest.mark.parametrize("jump", [1e-6, 1e-3, 1e-1])
def test_mixed_hlld_error_scales_with_the_jump(gpu, jump):
    """The mixed HLLD (FP64 physical fluxes, FP32 wave speeds and star-state
    differences in delta form) against the FP64 HLLD of the oracle on faces
    whose two states differ by a relative `jump`: the error is FP32
    rounding of the fan corrections, so it scales with the jump, not with
    |F| (the plain FP32 form's error is ~1e-7 of the flux whatever the
    jump).  With scale = max|F| + max|U| max(c_f) of the face, 99.9 % of the
    faces are within 1e-6 * jump * scale (+ 1e-14 * scale) and all within
    3e-5 * jump * scale: the few faces above the first bound are near HLLD's
    degenerate case, where rho (S - u)(S - S_M) - B_n^2 / mu0 (the
    denominator of the star states' transverse parts) cancels to 1e-3..1e-5
    of its terms (|B_t| << |B_n| with c_A,n ~ c_f) and the FP32 wave
    speeds' rounding is amplified by that ratio."""
    from pamhd_amd import solver as S
    import states
    rng = np.random.default_rng(int(1e6 * jump) % 1000 + 3)
    n = 8192
    L, _, _ = states.random_faces(n, 91, kind="generic")
    rho = L[0] * (1 + jump * rng.normal(size=n))
    v = L[1:4] / L[0] + jump * rng.normal(size=(3, n))
    B = L[5:8] * (1 + jump * rng.normal(size=(3, n)))
    p = (5 / 3 - 1) * (L[4] - 0.5 * (L[1:4] ** 2).sum(0) / L[0] - 0.5 * (L[5:8] ** 2).sum(0))
    p = p * (1 + jump * rng.normal(size=n))
    R = states.conserved(rho, v, p, B, 5 / 3, 1.0)
    Fo, so, sto = O.flux_batch(O.HLLD, L, R)
    tod = lambda a: __import__("torch").from_numpy(np.ascontiguousarray(a)).to(gpu)
    Fg, sg, stg = S.face_flux_batch("hlld-athena", tod(L), tod(R), precision="mixed")
    Fg, stg = Fg.cpu().numpy(), stg.cpu().numpy()
    ok = (sto == 0) & (stg == 0)
    assert ok.sum() > 0.9 * n
    scale = np.abs(Fo[:, ok]).max(0) + np.maximum(np.abs(L[:, ok]).max(0), np.abs(R[:, ok]).max(0)) * so[ok]
    err = (np.abs(Fg[:, ok] - Fo[:, ok]).max(0) / scale)
    print("jump", jump, "max relative error", err.max(), "median", np.median(err),
          "quantiles 0.99/0.999/0.9999", np.quantile(err, [0.99, 0.999, 0.9999]),
          "faces above 1e-6 jump", int((err > 1e-6 * jump + 1e-14).sum()), "of", int(ok.sum()),
          flush=True)
    assert np.quantile(err, 0.999) <= 1e-6 * jump + 1e-14
    assert err.max() <= 3e-5 * jump + 1e-14
