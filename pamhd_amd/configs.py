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
    thermal pressure n k_B T, IMF (0, By, Bz), an inner boundary ten times
    denser and hotter than the wind at rest, and a dipole at the centre whose
    field is clamped within one cell (min-distance)."""
    from .swbox import SolarWindBox
    h = box_length / edge
    p = number_density * K_BOLTZMANN * temperature
    return SolarWindBox(
        cells=(edge, edge, edge), start=(-box_length / 2,) * 3, cell_length=(h, h, h),
        direction="+x", number_density=number_density, velocity=(-speed, 0.0, 0.0),
        pressure=p, magnetic_field=(0.0, by, bz),
        inner_radius=MERCURY_RADIUS, inner_number_density=10 * number_density,
        inner_velocity=(0.0, 0.0, 0.0), inner_pressure=10 * p,
        bg_value=(0.0, 0.0, 0.0), bg_min_distance=h,
        dipoles=[(tuple(dipole_moment), (0.0, 0.0, 0.0))],
        adiabatic_index=5.0 / 3.0, vacuum_permeability=MU0_SI, proton_mass=PROTON_MASS)


def solar_wind_ensemble(nbox=64, seed=2024, edges=(32, 48, 64, 96, 128)):
    """Seeded ensemble of solar-wind boxes: edge drawn from `edges`, number
    density log-uniform in 1-20 cm^-3, speed uniform in 300-800 km/s,
    By and Bz uniform in [-10, 10] nT, T = 1e5 K."""
    rng = np.random.default_rng(seed)
    boxes = []
    for _ in range(nbox):
        edge = int(rng.choice(edges))
        n = 1e6 * math.exp(rng.uniform(math.log(1.0), math.log(20.0)))
        v = rng.uniform(3e5, 8e5)
        by, bz = rng.uniform(-1e-8, 1e-8, size=2)
        boxes.append(solar_wind_box(edge, n, v, float(by), float(bz)))
    return boxes
