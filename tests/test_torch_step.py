"""Pins the PyTorch restatement of the HLL step (tests/torch_step.py, the
gradient checker of the adjoint) against the FP64 oracle's step."""
import numpy as np
import pytest
import torch

import oracle_lib as O
import torch_step
from pamhd_amd import configs


@pytest.mark.parametrize("cells", [(8, 8, 1), (6, 5, 4), (5, 7, 6)])
def test_restated_step_matches_oracle(cells):
    state, length, gamma, mu0 = configs.orszag_tang_vortex(cells, noise=0.05, seed=3)
    rng = np.random.default_rng(1)
    state[1:4] *= 1 + 0.1 * rng.normal(size=state[1:4].shape)
    state[3] += 0.01 * rng.normal(size=state[3].shape)
    state[7] += 0.01 * rng.normal(size=state[7].shape)
    state[5:8] += 0.005 * rng.normal(size=state[5:8].shape)
    g = O.make_grid(cells, length, gamma, mu0)
    dt = 0.2 * min(length) / 3
    st, new, sm, zc = O.step(g, O.HLL, dt, state, refits=3)
    assert st == 0
    nx, ny, nz = cells
    y = torch_step.step(torch.from_numpy(state.reshape(8, nz, ny, nx).copy()), dt, length,
                        gamma, mu0).numpy().reshape(8, -1)
    scale = np.abs(new).max(1, keepdims=True)
    assert np.all(np.abs(y - new) <= 1e-14 * scale)
