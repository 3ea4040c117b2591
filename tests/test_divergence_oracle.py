"""Pins the oracle's divergence cleaning (periodic Poisson projection of the
face field).  The reference's divergence::remove delegates the solve to
dccrg's Poisson_Solve, which is not in the reference checkout, and its only
caller is stale, so nothing in the reference pins this function; it is
pinned here by the mathematics:
  * the discrete Laplacian's Fourier eigenmodes are solved to the tolerance;
  * B = curl A + grad psi (face field from random edge potentials A and cell
    potentials psi) is projected back to curl A;
  * a divergence-free field is left unchanged;
  * the new face divergence is the solver's residual.
"""
import math

import numpy as np
import pytest

import oracle_lib as O


def _lap_eig(n, h, m):
    return (2 * math.cos(2 * math.pi * m / n) - 2) / (h * h) if n > 1 else 0.0


@pytest.mark.parametrize("dims", [(16, 16, 16), (32, 8, 4), (12, 10, 6), (24, 1, 1), (1, 20, 20)])
def test_poisson_fourier_mode(dims):
    nx, ny, nz = dims
    h = (1.0 / nx, 0.7 / ny, 1.3 / nz)
    mx, my, mz = (1 if nx > 1 else 0), (2 if ny > 2 else 0), (1 if nz > 1 else 0)
    i = np.arange(nx)[None, None, :]
    j = np.arange(ny)[None, :, None]
    k = np.arange(nz)[:, None, None]
    phi = (np.cos(2 * math.pi * mx * i / nx + 0.3) * np.cos(2 * math.pi * my * j / ny + 0.1)
           * np.cos(2 * math.pi * mz * k / nz + 0.7))
    lam = _lap_eig(nx, h[0], mx) + _lap_eig(ny, h[1], my) + _lap_eig(nz, h[2], mz)
    rhs = (lam * phi).reshape(-1)
    rhs -= rhs.mean()
    st, sol, res, cyc = O.poisson(dims, h, rhs, tol=1e-11)
    assert st == 0 and res <= 1e-11, (res, cyc)
    sol = sol - sol.mean()
    ref = phi.reshape(-1) - phi.mean()
    print(dims, "cycles", cyc, "error", np.abs(sol - ref).max())
    assert np.abs(sol - ref).max() <= 1e-9 * np.abs(ref).max()
    assert cyc <= 30


def _curl_plus_grad(cells, length, seed):
    """Face field curl(A) + grad(psi) with A on edges and psi on cells."""
    nx, ny, nz = cells
    dx, dy, dz = length
    rng = np.random.default_rng(seed)
    Ax, Ay, Az = (rng.normal(size=(nz, ny, nx)) for _ in range(3))   # edge-centred
    psi = rng.normal(size=(nz, ny, nx))
    r = lambda a, ax: np.roll(a, -1, axis=ax)   # value at index + 1
    # Bx(i-1/2) = dAz/dy - dAy/dz with Az on (i-1/2, j-1/2) edges etc.
    Bx = (r(Az, 1) - Az) / dy - (r(Ay, 0) - Ay) / dz
    By = (r(Ax, 0) - Ax) / dz - (r(Az, 2) - Az) / dx
    Bz = (r(Ay, 2) - Ay) / dx - (r(Ax, 1) - Ax) / dy
    l = lambda a, ax: np.roll(a, 1, axis=ax)     # value at index - 1
    Gx = (psi - l(psi, 2)) / dx
    Gy = (psi - l(psi, 1)) / dy
    Gz = (psi - l(psi, 0)) / dz
    return np.stack([Bx, By, Bz]).reshape(3, -1), np.stack([Gx, Gy, Gz]).reshape(3, -1)


@pytest.mark.parametrize("cells", [(16, 16, 16), (20, 12, 8), (32, 16, 8)])
def test_projection_recovers_the_curl(cells):
    length = (1.0 / cells[0], 1.0 / cells[1], 2.0 / cells[2])
    curl, grad = _curl_plus_grad(cells, length, 5)
    n = curl.shape[1]
    state = np.zeros((8, n))
    state[0] = 1.0
    state[4] = 10.0
    state[5:8] = curl + 0.1 * grad
    g = O.make_grid(cells, length, 5 / 3, 1.0)
    div0 = O.div_face2volume(g, state)
    assert O.div_face2volume(g, np.concatenate([state[:5], curl])) < 1e-11 * np.abs(curl).max() / min(length)
    st, res, cyc = O.clean_divergence(g, state, tol=1e-10)
    assert st == 0
    err = np.abs(state[5:8] - curl).max() / np.abs(curl).max()
    div1 = O.div_face2volume(g, state)
    print(cells, "cycles", cyc, "residual", res, "div", div0, "->", div1, "error", err)
    assert err < 1e-8
    assert div1 < 1e-9 * div0
    assert cyc <= 30


def test_divergence_free_field_unchanged():
    cells = (16, 8, 8)
    length = (1 / 16, 1 / 8, 1 / 8)
    curl, _ = _curl_plus_grad(cells, length, 9)
    state = np.zeros((8, curl.shape[1]))
    state[0] = 1.0
    state[4] = 5.0
    state[5:8] = curl
    g = O.make_grid(cells, length, 5 / 3, 1.0)
    before = state.copy()
    st, res, cyc = O.clean_divergence(g, state, tol=1e-10)
    assert st == 0
    assert np.abs(state[5:8] - before[5:8]).max() <= 1e-10 * np.abs(curl).max()
