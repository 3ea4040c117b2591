"""The benchmark workload at its full size and horizon: the 256^3 periodic
vortex (+-1 % density noise, seed 2024, gamma 5/3), the final-init step and
100 main-loop steps (tests/mhd/test.cpp's loop, source/mhd/solve.hpp:1837-1924
per step) at time-step factor 0.3, with each of the four Riemann solvers,
bit for bit against the FP64 oracle: the oracle's final states are recorded
as SHA-256 digests (with the end time and per-variable sums) by
tests/golden/make_horizon_digests.py, which runs the OpenMP oracle for
~4-10 minutes per solver.  HLLD runs in each of its step structures."""
import hashlib
import json
import os

import numpy as np
import pytest

from pamhd_amd import configs
from pamhd_amd.solver import Grid

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "horizon_256.json")))
CELLS = tuple(GOLD["cells"])


def _digest(state):
    return hashlib.sha256(np.ascontiguousarray(state, dtype="<f8").tobytes()).hexdigest()


@pytest.fixture(scope="module")
def vortex():
    state, length, gamma, mu0 = configs.orszag_tang_vortex(CELLS, noise=GOLD["noise"],
                                                           seed=GOLD["seed"])
    # the digests apply only to the initial state they were computed from
    assert _digest(state) == GOLD["initial_sha256"], "initial state built differently on this host"
    return state, length, gamma, mu0


@pytest.mark.parametrize("solver,mode", [("rusanov", None), ("hll-athena", None),
                                         ("hlld-athena", "accumulate"), ("hlld-athena", "fused"),
                                         ("hlld-athena", "passes"), ("roe-athena", None)])
def test_horizon_bit_exact(gpu, vortex, solver, mode):
    state, length, gamma, mu0 = vortex
    want = GOLD["solvers"][solver]
    print(solver, mode or "", "100 steps at", CELLS, flush=True)
    with Grid(CELLS, length, gamma, mu0) as g:
        if mode:
            g.set_fused_step(mode)
        g.upload(state)
        steps, t = g.run(solver, GOLD["cfl"], 0.0, 1e9, max_steps=GOLD["steps"])
        out = g.download()
    assert steps == GOLD["steps"]
    assert t.hex() == want["t_end"]
    if _digest(out) != want["sha256"]:
        # name the variables that differ (the sums' rounding may depend on
        # the host's summation kernels, so they are a diagnostic only)
        sums = out.sum(axis=1)
        print("sums", sums.tolist(), "oracle", want["sums"], flush=True)
        pytest.fail("final state differs from the oracle's")
