"""Initial states of the benchmark workloads, built on the host.

orszag_tang_vortex: the 3-D periodic smooth-vortex case on [0,1]^3,
    rho = 25/(36 pi) (1 + a (2u - 1)),  u ~ U[0,1) seeded per cell,
    p = 5/(12 pi),  v = (-sin 2 pi y, sin 2 pi x, 0),
    B = (-sin 2 pi y, sin 4 pi x, 0) / sqrt(4 pi),  gamma = 5/3, mu0 = 1.
Face B is the exact average of B over each face (closed form), so the face
field is discretely divergence-free; the cell state follows
initialize_fluid (source/mhd/initialize.hpp:106-174): m = rho v at the cell
centre and E from p, v and the face-average cell field.
"""
import math

import numpy as np


def orszag_tang_vortex(cells, noise=0.01, seed=12345, divergence_sigma=0.0):
    """Returns (state (8, n) float64, cell_length, gamma, mu0).

    divergence_sigma > 0 adds a seeded Gaussian perturbation of that standard
    deviation to every face B (a divergence-carrying field).
    """
    nx, ny, nz = cells
    dx, dy, dz = 1.0 / nx, 1.0 / ny, 1.0 / nz
    gamma, mu0 = 5.0 / 3.0, 1.0
    rng = np.random.default_rng(seed)
    x = (np.arange(nx) + 0.5) * dx
    y = (np.arange(ny) + 0.5) * dy
    xf = np.arange(nx + 1) * dx
    yf = np.arange(ny + 1) * dy
    s4 = 1.0 / math.sqrt(4 * math.pi)
    n = nx * ny * nz
    state = np.zeros((8, nz, ny, nx))
    rho = 25.0 / (36.0 * math.pi) * (1.0 + noise * (2.0 * rng.random((nz, ny, nx)) - 1.0))
    p = 5.0 / (12.0 * math.pi)
    vx = np.broadcast_to(-np.sin(2 * math.pi * y)[None, :, None], (nz, ny, nx))
    vy = np.broadcast_to(np.sin(2 * math.pi * x)[None, None, :], (nz, ny, nx))
    # exact face averages: B_x on x-faces averaged over y, B_y on y-faces over x
    bx_j = s4 * (np.cos(2 * math.pi * yf[1:]) - np.cos(2 * math.pi * yf[:-1])) / (2 * math.pi * dy)
    by_i = s4 * (np.cos(4 * math.pi * xf[:-1]) - np.cos(4 * math.pi * xf[1:])) / (4 * math.pi * dx)
    Bx = np.broadcast_to(bx_j[None, :, None], (nz, ny, nx)).copy()
    By = np.broadcast_to(by_i[None, None, :], (nz, ny, nx)).copy()
    Bz = np.zeros((nz, ny, nx))
    if divergence_sigma > 0:
        Bx += rng.normal(0.0, divergence_sigma, Bx.shape)
        By += rng.normal(0.0, divergence_sigma, By.shape)
        Bz += rng.normal(0.0, divergence_sigma, Bz.shape)
    vbx = 0.5 * (Bx + np.roll(Bx, -1, axis=2))
    vby = 0.5 * (By + np.roll(By, -1, axis=1))
    vbz = 0.5 * (Bz + np.roll(Bz, -1, axis=0))
    state[0] = rho
    state[1] = vx * rho
    state[2] = vy * rho
    state[3] = 0.0
    kin = 0.5 * rho * (vx * vx + vy * vy)
    mag = 0.5 / mu0 * (vbx * vbx + vby * vby + vbz * vbz)
    state[4] = p / (gamma - 1) + kin + mag
    state[5] = Bx
    state[6] = By
    state[7] = Bz
    return state.reshape(8, n), (dx, dy, dz), gamma, mu0


def hot_blob(cells, contrast=16.0, radius=0.2, seed=1, noise=0.01):
    """Periodic unit box at rest with a uniform field, a hot sphere of
    `contrast` times the background pressure (sound speed sqrt(contrast)
    times larger) at the centre and seeded +-`noise` density noise: the
    per-cell time steps differ by ~sqrt(contrast), so per-cell substepping
    gives the outer cells longer periods.  Returns (state (8, n), cell
    length, gamma, mu0)."""
    nx, ny, nz = cells
    dx, dy, dz = 1.0 / nx, 1.0 / ny, 1.0 / nz
    gamma, mu0 = 5.0 / 3.0, 1.0
    rng = np.random.default_rng(seed)
    x = (np.arange(nx) + 0.5) * dx - 0.5
    y = (np.arange(ny) + 0.5) * dy - 0.5
    z = (np.arange(nz) + 0.5) * dz - 0.5
    r = np.sqrt(z[:, None, None] ** 2 + y[None, :, None] ** 2 + x[None, None, :] ** 2)
    rho = 1.0 + noise * (2.0 * rng.random((nz, ny, nx)) - 1.0)
    p = np.where(r < radius, contrast, 1.0) * 0.1
    B = (0.05, 0.03, 0.02)
    n = nx * ny * nz
    state = np.zeros((8, nz, ny, nx))
    state[0] = rho
    state[4] = p / (gamma - 1) + 0.5 / mu0 * (B[0] ** 2 + B[1] ** 2 + B[2] ** 2)
    for d in range(3):
        state[5 + d] = B[d]
    return state.reshape(8, n), (dx, dy, dz), gamma, mu0


# ---------------------------------------------------------------- solar wind

K_BOLTZMANN = 1.380649e-23          # source/simulation_options.hpp:63
PROTON_MASS = 1.67262192595e-27
MU0_SI = 1.2566370614359173e-06     # "vacuum-permeability" of the SI configs
MERCURY_RADIUS = 244e4


def solar_wind_box(edge, number_density, speed, by, bz, temperature=1e5,
                   box_length=32 * 4.88e6, dipole_moment=(0.0, 0.0, -225e16)):
    """A cubic solar-wind box of `edge` cells (boundary layer included)
    centred on a Mercury-sized planet, with the wind entering through the +x
    wall moving along -x ("direction": "+x", velocity (-speed, 0, 0)),
    thermal pressure n k_B T, an inner boundary ten times denser and hotter
    than the wind at rest, and a dipole at the centre whose field is clamped
    within one cell (min-distance).  The IMF (0, By, Bz) is the constant of
    the background field and the wind's own field is 0, as in the
    reference's Mercury configuration
    (tests/mhd/config_files/magnetohydrodynamic/magnetosphere/3d/-z/mercury1.json):
    planet cells zero their face field, which a state-borne IMF of a few nT
    at these densities (plasma beta ~ 0.05) turns into negative pressures
    within a few steps."""
    from .swbox import SolarWindBox
    h = box_length / edge
    p = number_density * K_BOLTZMANN * temperature
    return SolarWindBox(
        cells=(edge, edge, edge), start=(-box_length / 2,) * 3, cell_length=(h, h, h),
        direction="+x", number_density=number_density, velocity=(-speed, 0.0, 0.0),
        pressure=p, magnetic_field=(0.0, 0.0, 0.0),
        inner_radius=MERCURY_RADIUS, inner_number_density=10 * number_density,
        inner_velocity=(0.0, 0.0, 0.0), inner_pressure=10 * p,
        bg_value=(0.0, by, bz), bg_min_distance=h,
        dipoles=[(tuple(dipole_moment), (0.0, 0.0, 0.0))],
        adiabatic_index=5.0 / 3.0, vacuum_permeability=MU0_SI, proton_mass=PROTON_MASS)


def solar_wind_ensemble(nbox=64, seed=2024, edges=(32, 48, 64, 96, 128), imf=True):
    """Seeded ensemble of solar-wind boxes: edge drawn from `edges`, number
    density log-uniform in 1-20 cm^-3, speed uniform in 300-800 km/s,
    By and Bz uniform in [-10, 10] nT (multiplied by 0 when imf is False;
    the draws are made either way, so the other parameters do not change),
    T = 1e5 K.

    With the IMF these boxes stop with a non-positive pressure within
    50-200 steps: the face between a normal and a boundary cell receives
    half of its constrained-transport change (only the normal cell's copy is
    updated, update_mhd_state, source/mhd/solve.hpp:1503-1593, and the copies
    are then averaged, common_functions.hpp:184-269), so the field next to
    the outflow walls is not divergence-free and grows where a field is
    advected along them.  The reference's own Mercury configuration stops
    the same way after 132 steps of the same restatement."""
    rng = np.random.default_rng(seed)
    boxes = []
    for _ in range(nbox):
        edge = int(rng.choice(edges))
        n = 1e6 * math.exp(rng.uniform(math.log(1.0), math.log(20.0)))
        v = rng.uniform(3e5, 8e5)
        by, bz = rng.uniform(-1e-8, 1e-8, size=2)
        f = 1.0 if imf else 0.0
        boxes.append(solar_wind_box(edge, n, v, f * float(by), f * float(bz)))
    return boxes


# ------------------------------------------------------- 2-D adjoint workload

def _fourier_field(n, kmax, rng):
    """Random periodic field on the n x n cell centres (and its values at the
    cell corners), sum over 0 < |k|_inf <= kmax (half plane) of
    Gaussian-weighted cos and sin modes, normalised to unit RMS at the
    centres.  Evaluated with an inverse FFT: the mode sum at x = (i + s)/n is
    Re sum (a - i b) e^{2 pi i k.(i + s)/n}."""
    ks = [(kx, ky) for kx in range(-kmax, kmax + 1) for ky in range(0, kmax + 1)
          if (ky > 0 or kx > 0)]
    a = rng.normal(size=len(ks))
    b = rng.normal(size=len(ks))
    out = []
    for shift in (0.5, 0.0):
        C = np.zeros((n, n), dtype=np.complex128)
        for (kx, ky), ca, cb in zip(ks, a, b):
            C[ky % n, kx % n] += (ca - 1j * cb) * np.exp(2j * math.pi * (kx + ky) * shift / n)
        out.append(np.real(np.fft.ifft2(C)) * (n * n))
    s = np.sqrt((out[0] ** 2).mean())
    return out[0] / s, out[1] / s


def adjoint_vortex_2d(n=512, seed=5, kmax=8):
    """Fixed parts of the 2-D periodic adjoint workload on [0,1]^2 (n x n x 1
    cells, gamma 5/3, mu0 1): the smooth vortex (rho, p, v, and the vector
    potential A_z of B = (-sin 2 pi y, sin 4 pi x, 0) / sqrt(4 pi) at cell
    corners) and seeded unit-RMS random-Fourier perturbations (|k| <= kmax)
    of rho, v_x, v_y, p and A_z.  Returns a dict of float64 numpy arrays
    indexed [j, i]."""
    rng = np.random.default_rng(seed)
    xc = (np.arange(n) + 0.5) / n
    xk = np.arange(n) / n
    Xc, Yc = np.meshgrid(xc, xc, indexing="xy")
    Xk, Yk = np.meshgrid(xk, xk, indexing="xy")
    s4 = 1.0 / math.sqrt(4 * math.pi)
    base = {
        "rho": np.full((n, n), 25.0 / (36.0 * math.pi)),
        "p": np.full((n, n), 5.0 / (12.0 * math.pi)),
        "vx": -np.sin(2 * math.pi * Yc),
        "vy": np.sin(2 * math.pi * Xc),
        "A": s4 * (np.cos(2 * math.pi * Yk) / (2 * math.pi) + np.cos(4 * math.pi * Xk) / (4 * math.pi)),
    }
    pert = {}
    for name in ("rho", "vx", "vy", "p"):
        pert[name] = _fourier_field(n, kmax, rng)[0]
    pert["A"] = _fourier_field(n, kmax, rng)[1]
    return {"n": n, "base": base, "pert": pert, "length": (1.0 / n, 1.0 / n, 1.0),
            "gamma": 5.0 / 3.0, "mu0": 1.0}


def adjoint_initial_state(w, amplitudes):
    """(8, n*n) float64 torch state of the workload `w` for the 4 amplitudes
    (rho, v, p, A_z; a torch tensor, differentiable): rho = rho0 (1 + a0 d_rho),
    v = v0 + a1 d_v, p = p0 (1 + a2 d_p), A_z = A0 + a3 0.05 d_A; face B from
    corner differences of A_z (discretely divergence-free), m = rho v and
    E = p/(gamma-1) + rho v^2/2 + |face-average B|^2/(2 mu0)."""
    import torch
    dev, dt = amplitudes.device, torch.float64
    T = lambda a: torch.as_tensor(a, dtype=dt, device=dev)
    n = w["n"]
    b, d = w["base"], w["pert"]
    h = 1.0 / n
    rho = T(b["rho"]) * (1 + amplitudes[0] * T(d["rho"]))
    vx = T(b["vx"]) + amplitudes[1] * T(d["vx"])
    vy = T(b["vy"]) + amplitudes[1] * T(d["vy"])
    p = T(b["p"]) * (1 + amplitudes[2] * T(d["p"]))
    A = T(b["A"]) + amplitudes[3] * 0.05 * T(d["A"])        # corners (x_i - h/2, y_j - h/2)
    Bx = (torch.roll(A, -1, 0) - A) / h                     # face (i - 1/2, j)
    By = -(torch.roll(A, -1, 1) - A) / h                    # face (i, j - 1/2)
    Bz = torch.zeros_like(Bx)
    vbx = 0.5 * (Bx + torch.roll(Bx, -1, 1))
    vby = 0.5 * (By + torch.roll(By, -1, 0))
    E = p / (w["gamma"] - 1) + 0.5 * rho * (vx * vx + vy * vy) + 0.5 / w["mu0"] * (vbx * vbx + vby * vby)
    U = torch.stack([rho, rho * vx, rho * vy, torch.zeros_like(rho), E, Bx, By, Bz])
    return U.reshape(8, n * n)


def window_magnetic_energy(U, n, lo=224, size=64, mu0=1.0):
    """sum over a size x size window of cells of |B_cell|^2 / (2 mu0) * dA,
    B_cell the face-average field of the final state (torch, differentiable)."""
    s = U.reshape(8, n, n)
    bx = 0.5 * (s[5] + s[5].roll(-1, 1))
    by = 0.5 * (s[6] + s[6].roll(-1, 0))
    bz = s[7]
    e = 0.5 / mu0 * (bx * bx + by * by + bz * bz) / (n * n)
    return e[lo:lo + size, lo:lo + size].sum()


def smooth_direction(n, scale, rng, kmax=8):
    """Random perturbation direction of an (8, n*n) 2-D state: each
    component a unit-RMS random-Fourier field (|k| <= kmax) times
    scale[component].  Smooth directions keep a central difference with
    h = 1e-6 clear of the flux's max / min switches, which rough per-cell
    noise of that size can cross."""
    rows = [_fourier_field(n, kmax, rng)[0].reshape(-1) * float(scale[q]) for q in range(8)]
    return np.stack(rows)
