"""Solar-wind boxes and ensembles of them on the GPU.

Mirrors the options and program of the reference's solar-wind box
(tests/mhd/solar_wind_box.cpp, source/solar_wind_box_options.hpp,
source/mhd/solar_wind_box.hpp): a uniform non-periodic grid whose outer layer
of cells is the boundary, an inflow wall, outflow copies on the other walls,
planet cells inside the inner-boundary radius and a background field of a
constant plus point dipoles.  Physical units are those of the options (SI in
the reference's configuration files).
"""
import ctypes
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .solver import solver_id

# "direction" strings of source/solar_wind_box_options.hpp:100-120
DIRECTIONS = {"-x": -1, "+x": +1, "-y": -2, "+y": +2, "-z": -3, "+z": +3}


@dataclass
class SolarWindBox:
    cells: tuple                      # (nx, ny, nz), boundary layer included
    start: tuple                      # grid-options "start"
    cell_length: tuple                # volume / cells
    direction: str = "+x"
    number_density: float = 1e7       # 1/m^3
    velocity: tuple = (-4e5, 0.0, 0.0)
    pressure: float = 1e-11
    magnetic_field: tuple = (0.0, 0.0, 0.0)
    inner_radius: float = 0.0
    inner_number_density: float = 1e8
    inner_velocity: tuple = (0.0, 0.0, 0.0)
    inner_pressure: float = 1e-10
    bg_value: tuple = (0.0, 0.0, 0.0)
    bg_min_distance: float = 0.0
    dipoles: list = field(default_factory=list)   # [(moment (3), position (3))]
    adiabatic_index: float = 5.0 / 3.0
    vacuum_permeability: float = 4e-7 * np.pi
    proton_mass: float = 1.67262192595e-27

    @property
    def n(self):
        return self.cells[0] * self.cells[1] * self.cells[2]

    def grid_desc(self):
        return _lib.GridDesc((ctypes.c_int32 * 3)(*self.cells),
                             (ctypes.c_double * 3)(*self.start),
                             (ctypes.c_double * 3)(*self.cell_length),
                             (ctypes.c_int32 * 3)(0, 0, 0))

    def physics(self):
        return _lib.Physics(self.adiabatic_index, self.vacuum_permeability, self.proton_mass)

    def sw_desc(self):
        d = _lib.SolarWindDesc()
        d.direction = DIRECTIONS[self.direction]
        d.number_density = self.number_density
        d.velocity[:] = self.velocity
        d.pressure = self.pressure
        d.magnetic_field[:] = self.magnetic_field
        d.inner_radius = self.inner_radius
        d.inner_number_density = self.inner_number_density
        d.inner_velocity[:] = self.inner_velocity
        d.inner_pressure = self.inner_pressure
        d.bg_value[:] = self.bg_value
        d.bg_min_distance = self.bg_min_distance
        if len(self.dipoles) > _lib.MAX_DIPOLES:
            raise ValueError("at most %d dipoles" % _lib.MAX_DIPOLES)
        d.n_dipoles = len(self.dipoles)
        for q, (m, p) in enumerate(self.dipoles):
            d.dipole_moment[q][:] = m
            d.dipole_position[q][:] = p
        return d


class BoxGrid:
    """One solar-wind box resident on the GPU (pamhd_create_solar_wind_box).

    State layout (host side): float64 array (14, n): rho, m (3), E, Vol_B (3)
    and the cell's own copies of its -x, +x, -y, +y, -z, +z face B.
    """

    def __init__(self, box: SolarWindBox):
        L = _lib.load()
        self.box = box
        self.n = box.n
        h = ctypes.c_void_p()
        _lib.check(L.pamhd_create_solar_wind_box(ctypes.byref(box.grid_desc()),
                                                 ctypes.byref(box.physics()),
                                                 ctypes.byref(box.sw_desc()), ctypes.byref(h)))
        self._h = h
        self._L = L

    def close(self):
        if getattr(self, "_h", None):
            self._L.pamhd_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, code):
        return _lib.check(code, self._h)

    @property
    def kind(self):
        """"solar wind box" (pamhd_grid_kind)"""
        k = ctypes.c_int(-1)
        self._check(self._L.pamhd_grid_kind(self._h, ctypes.byref(k)))
        return _lib.GRID_KINDS[k.value]

    def download(self):
        out = np.empty((14, self.n))
        self._check(self._L.pamhd_download_state(self._h, out.ctypes.data))
        return out

    def upload(self, state):
        state = np.ascontiguousarray(state, dtype=np.float64)
        if state.shape != (14, self.n):
            raise ValueError("state must have shape (14, %d)" % self.n)
        self._check(self._L.pamhd_upload_state(self._h, state.ctypes.data))

    def cell_types(self):
        ct = np.empty(self.n, dtype=np.int8)
        ru = np.empty(self.n, dtype=np.uint8)
        self._check(self._L.pamhd_box_cell_types(self._h, ct.ctypes.data, ru.ctypes.data))
        return ct, ru

    def background(self):
        bg = np.empty((9, self.n))
        self._check(self._L.pamhd_box_background(self._h, bg.ctypes.data))
        return bg

    def timestep(self, solver, max_time_step, time_step_factor, stream=None):
        dt = ctypes.c_double(0)
        self._check(self._L.pamhd_timestep(self._h, solver_id(solver), max_time_step,
                                           time_step_factor, ctypes.byref(dt), stream))
        return dt.value

    def step(self, solver, max_time_step, time_step_factor, sync=True, stream=None):
        dt = ctypes.c_double(0)
        self._check(self._L.pamhd_step(self._h, solver_id(solver), max_time_step,
                                       time_step_factor, ctypes.byref(dt) if sync else None,
                                       stream))
        return dt.value if sync else None

    def run(self, solver, time_step_factor, t_start, t_end, max_steps=1 << 62, stream=None):
        steps, tf = ctypes.c_long(0), ctypes.c_double(0)
        self._check(self._L.pamhd_run(self._h, solver_id(solver), time_step_factor, t_start,
                                      t_end, max_steps, ctypes.byref(steps), ctypes.byref(tf),
                                      stream))
        return steps.value, tf.value

    def div_face2volume(self):
        v = ctypes.c_double(0)
        self._check(self._L.pamhd_div_face2volume(self._h, ctypes.byref(v)))
        return v.value

    def check(self):
        self._check(self._L.pamhd_check(self._h))


class Ensemble:
    """Boxes of different sizes advanced together (pamhd_ensemble_*)."""

    def __init__(self, boxes):
        L = _lib.load()
        self.boxes = list(boxes)
        nb = len(self.boxes)
        phys = self.boxes[0].physics()
        for b in self.boxes[1:]:
            if (b.adiabatic_index, b.vacuum_permeability, b.proton_mass) != (
                    self.boxes[0].adiabatic_index, self.boxes[0].vacuum_permeability,
                    self.boxes[0].proton_mass):
                raise ValueError("all boxes of an ensemble share the physics constants")
        grids = (_lib.GridDesc * nb)(*[b.grid_desc() for b in self.boxes])
        sws = (_lib.SolarWindDesc * nb)(*[b.sw_desc() for b in self.boxes])
        e = ctypes.c_void_p()
        _lib.check_ensemble(L.pamhd_ensemble_create(nb, grids, ctypes.byref(phys), sws,
                                                    ctypes.byref(e)))
        self._e = e
        self._L = L
        self.cells = sum(b.n for b in self.boxes)

    def close(self):
        if getattr(self, "_e", None):
            self._L.pamhd_ensemble_destroy(self._e)
            self._e = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, code):
        return _lib.check_ensemble(code, self._e)

    def timestep(self, solver, max_time_step, time_step_factor, stream=None):
        self._check(self._L.pamhd_ensemble_timestep(self._e, solver_id(solver), max_time_step,
                                                    time_step_factor, stream))

    def run(self, solver, time_step_factor, steps, use_graph=True, stream=None, persistent=False):
        """`steps` main-loop steps of every box: one launch per stage, the
        steps replayed from a HIP graph (use_graph), or all of them in one
        persistent launch (persistent)."""
        mode = 2 if persistent else (1 if use_graph else 0)
        self._check(self._L.pamhd_ensemble_run(self._e, solver_id(solver), time_step_factor,
                                               steps, mode, stream))

    def set_barrier_polls(self, polls=0):
        """Polls of a persistent run's barrier wait before it gives up (0:
        the default, about a second)."""
        self._check(self._L.pamhd_ensemble_set_barrier_polls(self._e, int(polls)))

    def check(self, stream=None):
        dts = np.empty(len(self.boxes))
        self._check(self._L.pamhd_ensemble_check(self._e, dts.ctypes.data, stream))
        return dts

    def download(self, box):
        out = np.empty((14, self.boxes[box].n))
        self._check(self._L.pamhd_ensemble_download(self._e, box, out.ctypes.data))
        return out

    def set_tile(self, tile):
        self._check(self._L.pamhd_ensemble_set_tile(self._e, tile))

    def autotune(self, solver, time_step_factor, stream=None):
        buf = ctypes.create_string_buffer(1024)
        self._check(self._L.pamhd_ensemble_autotune(self._e, solver_id(solver), time_step_factor,
                                                    buf, 1024, stream))
        return buf.value.decode()
