"""Seeded random face states for the Riemann-solver tests (test infrastructure)."""
import numpy as np


def conserved(rho, v, p, B, gamma, mu0):
    v = np.asarray(v, dtype=np.float64)
    B = np.asarray(B, dtype=np.float64)
    E = p / (gamma - 1) + 0.5 * rho * (v * v).sum(0) + 0.5 / mu0 * (B * B).sum(0)
    return np.concatenate([np.atleast_2d(rho), rho * v, np.atleast_2d(E), B])


def random_faces(n, seed, gamma=5.0 / 3.0, mu0=1.0, kind="generic"):
    """(L, R, bg) with shapes (8, n), (8, n), (3, n).

    kind: generic | zero_field | zero_bx | supersonic | equal | strong_field |
          weak_bx
    """
    rng = np.random.default_rng(seed)

    def side():
        rho = rng.uniform(0.1, 2.0, n)
        p = rng.uniform(0.05, 2.0, n)
        v = rng.normal(0.0, 1.0, (3, n))
        B = rng.normal(0.0, 1.0, (3, n))
        return rho, v, p, B

    rl, vl, pl, Bl = side()
    rr, vr, pr, Br = side()
    if kind == "zero_field":
        Bl[:] = 0
        Br[:] = 0
    elif kind == "zero_bx":
        Bl[0] = 0
        Br[0] = 0
    elif kind == "weak_bx":
        Bl[0] = 1e-9 * rng.normal(size=n)
        Br[0] = Bl[0]
    elif kind == "supersonic":
        sgn = np.where(rng.random(n) < 0.5, -1.0, 1.0)
        vl[0] = sgn * rng.uniform(8, 20, n)
        vr[0] = vl[0]
    elif kind == "strong_field":
        Bl *= 10
        Br *= 10
        pl *= 0.01
        pr *= 0.01
    if kind == "equal":
        rr, vr, pr, Br = rl, vl, pl, Bl
    # a shared normal field component, as div B = 0 assumes between states
    if kind in ("generic", "strong_field", "supersonic"):
        half = n // 2
        Br[0, :half] = Bl[0, :half]
    L = conserved(rl, vl, pl, Bl, gamma, mu0)
    R = conserved(rr, vr, pr, Br, gamma, mu0)
    bg = rng.normal(0.0, 0.5, (3, n))
    return L, R, bg
