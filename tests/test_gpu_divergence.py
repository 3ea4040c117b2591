"""GPU divergence cleaning (multigrid Poisson solve, face projection, energy
refit) against the FP64 oracle, through the C ABI.

Every kernel of the device solver evaluates each value in the oracle's
operation order, so for the same number of V-cycles the cleaned state is
bit-identical; the cycle count itself comes from residual norms that are
reduced in a different order on each side, so the tests also check that
both sides stopped after the same number of cycles.
"""
import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu


def _perturbed(cells, sigma=1e-3, seed=8):
    from pamhd_amd import configs
    return configs.orszag_tang_vortex(cells, noise=0.01, seed=seed, divergence_sigma=sigma)


def _oracle_clean(cells, length, gamma, mu0, state, tol, max_cycles=50):
    og = O.make_grid(cells, length, gamma, mu0)
    ref = state.copy()
    st, rel, cyc = O.clean_divergence(og, ref, tol=tol, max_cycles=max_cycles)
    return st, rel, cyc, ref, og


# (cells, smoother): all-even grids run the LDS-tiled fused red-black sweep
# on their large levels, grids with odd sizes damped Jacobi; 2-D grids skip z
@pytest.mark.parametrize("cells,fused", [
    ((16, 16, 16), True), ((16, 16, 16), False), ((64, 32, 48), True),
    ((20, 18, 16), True), ((24, 24, 1), True), ((40, 12, 10), True), ((9, 7, 5), True)])
def test_clean_divergence_bit_exact(gpu, cells, fused):
    from pamhd_amd import solver as S
    state, length, gamma, mu0 = _perturbed(cells)
    st_o, rel_o, cyc_o, ref, og = _oracle_clean(cells, length, gamma, mu0, state, 1e-10)
    with S.Grid(cells, length, gamma, mu0) as g:
        g.upload(state)
        g.set_poisson_smoother(fused)
        before, rel_g, cyc_g, st_g = g.clean_divergence(tol=1e-10, max_cycles=50, check=False)
        out = g.download()
        after = g.div_face2volume()
    print(cells, "fused", fused, "oracle", st_o, rel_o, cyc_o, "gpu", st_g, rel_g, cyc_g,
          "div", before, "->", after, flush=True)
    assert cyc_g == cyc_o
    assert (st_g == 0) == (st_o == 0)
    assert rel_g == pytest.approx(rel_o, rel=1e-6, abs=1e-300)
    np.testing.assert_array_equal(out, ref)
    assert before == pytest.approx(O.div_face2volume(og, state), rel=1e-12)
    assert after < 1e-6 * before


def test_clean_then_steps_keep_divergence(gpu):
    """Clean a divergence-carrying 32^3 state, then run 10 HLLD steps: the
    constrained-transport update keeps the face divergence at the cleaned
    level, and the whole sequence matches the oracle bit for bit."""
    from pamhd_amd import solver as S
    cells = (32, 32, 32)
    state, length, gamma, mu0 = _perturbed(cells, sigma=1e-3, seed=4)
    st_o, rel_o, cyc_o, ref, og = _oracle_clean(cells, length, gamma, mu0, state, 1e-10)
    assert st_o == 0
    st, steps_o, t_o = O.run_periodic(og, O.HLLD, 0.3, 0.0, 1e9, ref, max_steps=10)
    assert st == 0
    with S.Grid(cells, length, gamma, mu0) as g:
        g.upload(state)
        g.clean_divergence(tol=1e-10)
        d0 = g.div_face2volume()
        g.run("hlld-athena", 0.3, 0.0, 1e9, max_steps=10)
        d1 = g.div_face2volume()
        out = g.download()
    np.testing.assert_array_equal(out, ref)
    assert d1 < 10 * d0 + 1e-13


def test_clean_divergence_free_field_is_a_no_op(gpu):
    """The unperturbed vortex's face field is divergence-free to roundoff:
    the right-hand side is at the roundoff floor, no V-cycle runs and the
    field is unchanged."""
    from pamhd_amd import solver as S
    cells = (32, 32, 8)
    state, length, gamma, mu0 = _perturbed(cells, sigma=0.0)
    with S.Grid(cells, length, gamma, mu0) as g:
        g.upload(state)
        before, rel, cyc, st = g.clean_divergence(tol=1e-10)
        out = g.download()
    assert st == 0 and cyc == 0
    np.testing.assert_array_equal(out[5:], state[5:])
    assert np.abs(out[4] - state[4]).max() <= 1e-13 * np.abs(state[4]).max()


def test_clean_divergence_reports_unconverged(gpu):
    """With one V-cycle allowed and a 1e-14 tolerance the solve cannot
    converge: PAMHD_E_POISSON, the message names the residual, and the field
    is still projected (its divergence drops)."""
    from pamhd_amd import solver as S
    from pamhd_amd import _lib
    cells = (32, 32, 32)
    state, length, gamma, mu0 = _perturbed(cells)
    with S.Grid(cells, length, gamma, mu0) as g:
        g.upload(state)
        d0 = g.div_face2volume()
        before, rel, cyc, st = g.clean_divergence(tol=1e-14, max_cycles=1, check=False)
        assert st == _lib.PAMHD_E_POISSON
        assert "residual" in g.last_error()
        assert cyc == 1 and rel > 1e-14
        assert g.div_face2volume() < d0
        with pytest.raises(_lib.PamhdError):
            g.clean_divergence(tol=1e-14, max_cycles=1)
