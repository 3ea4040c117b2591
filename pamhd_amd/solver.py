"""Host-side mirror of the reference's MHD solver interface over the C ABI.

The reference drives its step through
    pamhd::mhd::timestep(solver, grid, options, t, max_dt, cfl, gamma, mu0, ...)
        -> (dt, max_substep, flux_calcs)          source/mhd/solve.hpp:1837-1924
    pamhd::mhd::apply_boundaries(...)            source/mhd/boundaries.hpp:1249
    pamhd::math::get_divergence_face2volume(...) source/math/nabla.hpp:58-113
selecting the Riemann solver by the "solver-mhd" string of
source/mhd/options.hpp:97-117 / tests/mhd/test.cpp:294-316.  Grid here
wraps one device-resident periodic grid (include/pamhd_mi355x.h).
"""
import ctypes

import numpy as np

from . import _lib

# "solver-mhd" strings accepted by tests/mhd/test.cpp:294-316
SOLVER_NAMES = {
    "rusanov": _lib.RUSANOV,
    "hll-athena": _lib.HLL_ATHENA,
    "hlld-athena": _lib.HLLD_ATHENA,
    "roe-athena": _lib.ROE_ATHENA,
}


def solver_id(solver):
    if isinstance(solver, str):
        try:
            return SOLVER_NAMES[solver]
        except KeyError:
            raise ValueError("Invalid solver: %s, use rusanov, hll-athena, "
                             "hlld-athena or roe-athena" % solver) from None
    return int(solver)


class Grid:
    """A periodic nx*ny*nz grid of MHD cells resident on the GPU.

    State layout (host side): float64 array of shape (8, n), n = nx*ny*nz,
    rows rho, m_x, m_y, m_z, E, B_x(i-1/2), B_y(j-1/2), B_z(k-1/2), cell
    index i + nx*(j + ny*k).
    """

    def __init__(self, cells, cell_length, adiabatic_index=5.0 / 3.0,
                 vacuum_permeability=1.0, proton_mass=1.0, precision="fp64",
                 start=(0.0, 0.0, 0.0), periodic=(True, True, True), types=None):
        """types = (var_type (4, n), cell_type (n,), face_type (6, n)) makes a
        grid with boundaries (pamhd_create_bounded): FP64, 14 doubles per
        cell, closed where periodic is False"""
        L = _lib.load()
        self.cells = tuple(int(c) for c in cells)
        self.n = self.cells[0] * self.cells[1] * self.cells[2]
        self.cell_length = tuple(float(x) for x in cell_length)
        self.gamma = float(adiabatic_index)
        self.mu0 = float(vacuum_permeability)
        self.precision = {"fp64": _lib.FP64, "mixed": _lib.MIXED}[precision]
        self.components = 8 if types is None else 14
        desc = _lib.GridDesc((ctypes.c_int32 * 3)(*self.cells),
                             (ctypes.c_double * 3)(*start),
                             (ctypes.c_double * 3)(*self.cell_length),
                             (ctypes.c_int32 * 3)(*[1 if p else 0 for p in periodic]))
        phys = _lib.Physics(self.gamma, self.mu0, float(proton_mass))
        h = ctypes.c_void_p()
        if types is None:
            _lib.check(L.pamhd_create(ctypes.byref(desc), ctypes.byref(phys),
                                      self.precision, ctypes.byref(h)))
        else:
            if self.precision != _lib.FP64:
                raise ValueError("grids with boundaries run in FP64")
            bp = ctypes.POINTER(ctypes.c_byte)
            arrs = [np.ascontiguousarray(a, dtype=np.int8).reshape(-1) for a in types]
            for a, k in zip(arrs, (4, 1, 6)):
                if a.size != k * self.n:
                    raise ValueError("type arrays must hold 4 n, n and 6 n entries")
            _lib.check(L.pamhd_create_bounded(ctypes.byref(desc), ctypes.byref(phys),
                                              *[a.ctypes.data_as(bp) for a in arrs],
                                              ctypes.byref(h)))
        self._h = h
        self._L = L
        self.time = 0.0

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

    def last_error(self):
        return _lib.last_error(self._h)

    @property
    def kind(self):
        """"periodic" or "bounded" (pamhd_grid_kind)"""
        k = ctypes.c_int(-1)
        self._check(self._L.pamhd_grid_kind(self._h, ctypes.byref(k)))
        return _lib.GRID_KINDS[k.value]

    # --- state ----------------------------------------------------------
    def upload(self, state):
        state = np.ascontiguousarray(state, dtype=np.float64)
        if state.shape != (self.components, self.n):
            raise ValueError("state must have shape (%d, %d)" % (self.components, self.n))
        self._check(self._L.pamhd_upload_state(self._h, state.ctypes.data))

    def download(self):
        out = np.empty((self.components, self.n))
        self._check(self._L.pamhd_download_state(self._h, out.ctypes.data))
        return out

    def set_boundary_values(self, mask, number_density, velocity, pressure, face_field):
        """value boundaries resolved per cell (pamhd_set_boundary_values)"""
        m = np.ascontiguousarray(mask, dtype=np.uint8)
        arrs = [np.ascontiguousarray(a, dtype=np.float64)
                for a in (number_density, velocity, pressure, face_field)]
        dp = ctypes.POINTER(ctypes.c_double)
        self._check(self._L.pamhd_set_boundary_values(
            self._h, m.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte)),
            *[a.ctypes.data_as(dp) for a in arrs]))

    def device_ptr(self):
        p = ctypes.c_void_p()
        self._check(self._L.pamhd_state_device_ptr(self._h, ctypes.byref(p)))
        return p.value

    def set_background_faces(self, bg):
        if bg is None:
            self._check(self._L.pamhd_set_background_faces(self._h, None))
            return
        bg = np.ascontiguousarray(bg, dtype=np.float64)
        if bg.shape != (9, self.n):
            raise ValueError("background field must have shape (9, %d)" % self.n)
        self._check(self._L.pamhd_set_background_faces(self._h, bg.ctypes.data))

    # --- the step ---------------------------------------------------------
    def timestep(self, solver, max_time_step, time_step_factor, sync=True, stream=None):
        """pamhd::mhd::timestep: returns the step length (None when sync=False)."""
        dt = ctypes.c_double(0)
        self._check(self._L.pamhd_timestep(self._h, solver_id(solver), max_time_step,
                                           time_step_factor,
                                           ctypes.byref(dt) if sync else None, stream))
        return dt.value if sync else None

    def apply_boundaries(self, simulation_time=0.0, stream=None):
        self._check(self._L.pamhd_apply_boundaries(self._h, simulation_time, stream))

    def step(self, solver, max_time_step, time_step_factor, sync=True, stream=None):
        """timestep + apply_boundaries of one main-loop iteration."""
        dt = ctypes.c_double(0)
        self._check(self._L.pamhd_step(self._h, solver_id(solver), max_time_step,
                                       time_step_factor,
                                       ctypes.byref(dt) if sync else None, stream))
        return dt.value if sync else None

    def run(self, solver, time_step_factor, t_start, t_end, max_steps=1 << 62, stream=None):
        """The program loop of tests/mhd/test.cpp; returns (steps, t_final)."""
        steps = ctypes.c_long(0)
        tf = ctypes.c_double(0)
        self._check(self._L.pamhd_run(self._h, solver_id(solver), time_step_factor,
                                      t_start, t_end, max_steps, ctypes.byref(steps),
                                      ctypes.byref(tf), stream))
        self.time = tf.value
        return steps.value, tf.value

    def check(self):
        self._check(self._L.pamhd_check(self._h))

    def autotune(self, solver, stream=None):
        """Pick the fastest bit-identical launch variants; returns the report."""
        buf = ctypes.create_string_buffer(4096)
        self._check(self._L.pamhd_autotune(self._h, solver_id(solver), buf, 4096, stream))
        return buf.value.decode()

    def set_substep_walk(self, workgroups=0):
        """Workgroups of the substep passes that walk their block lists (0:
        the default); every value gives the same results."""
        self._check(self._L.pamhd_set_substep_walk(self._h, int(workgroups)))

    def set_substepping(self, substep_min=0, substep_max=0):
        """Per-cell 2^N substepping ("substep-min", "substep-max"): ints for
        every cell, or arrays of each cell's values (x fastest, set again
        before each step when they depend on time); substep_max 0 = global
        time step.  Set before the final-init timestep."""
        if np.ndim(substep_min) == 0 and np.ndim(substep_max) == 0:
            self._check(self._L.pamhd_set_substepping(self._h, int(substep_min), int(substep_max)))
            return
        n = self.n
        lo = np.ascontiguousarray(np.broadcast_to(substep_min, (n,)), dtype=np.int32)
        hi = np.ascontiguousarray(np.broadcast_to(substep_max, (n,)), dtype=np.int32)
        ip = ctypes.POINTER(ctypes.c_int)
        self._check(self._L.pamhd_set_substepping_cells(self._h, lo.ctypes.data_as(ip),
                                                        hi.ctypes.data_as(ip)))

    def last_max_substep(self):
        """Largest period (in substeps) of the last step."""
        v = ctypes.c_int(0)
        self._check(self._L.pamhd_last_max_substep(self._h, ctypes.byref(v)))
        return v.value

    def last_substep_periods(self):
        """Each cell's substep period in the last substepped step (int32 (n,))."""
        n = int(np.prod(self.cells))
        out = np.empty(n, dtype=np.int32)
        self._check(self._L.pamhd_last_substep_periods(self._h, out.ctypes.data))
        return out

    def last_flux_calcs(self):
        """Faces solved by the last step: timestep()'s third return value."""
        v = ctypes.c_ulonglong(0)
        self._check(self._L.pamhd_last_flux_calcs(self._h, ctypes.byref(v)))
        return v.value

    def set_launch(self, flux_remap=0, update_kz=0, update_remap=0):
        self._check(self._L.pamhd_set_launch(self._h, flux_remap, update_kz, update_remap))

    def enable_timing(self, enable=True):
        self._check(self._L.pamhd_enable_timing(self._h, 1 if enable else 0))

    def kernel_time(self, kernel):
        """(total milliseconds, launches) of one kernel since timing was enabled."""
        ms = ctypes.c_double(0)
        n = ctypes.c_long(0)
        self._check(self._L.pamhd_kernel_time(self._h, kernel.encode(), ctypes.byref(ms),
                                              ctypes.byref(n)))
        return ms.value, n.value

    def div_face2volume(self):
        v = ctypes.c_double(0)
        self._check(self._L.pamhd_div_face2volume(self._h, ctypes.byref(v)))
        return v.value

    def clean_divergence(self, tol=1e-10, max_cycles=50, stream=None, check=True):
        """Multigrid divergence cleaning of the face field (the role of
        pamhd::divergence::remove, source/divergence/remove.hpp:609-769).
        Returns (mean |div B| before, relative residual, V-cycles, status);
        with check=True a tolerance not reached raises PamhdError."""
        before, rel, cyc = ctypes.c_double(0), ctypes.c_double(0), ctypes.c_int(0)
        st = self._L.pamhd_clean_divergence(self._h, tol, max_cycles, ctypes.byref(before),
                                            ctypes.byref(rel), ctypes.byref(cyc),
                                            stream)
        if check:
            self._check(st)
        return before.value, rel.value, cyc.value, st

    # --- frozen time-step schedule and its adjoint ------------------------
    def step_fixed(self, solver, dt, stream=None):
        """One main-loop step of exactly dt (no CFL limit)."""
        self._check(self._L.pamhd_step_fixed(self._h, solver_id(solver), dt, stream))

    def step_fixed_from(self, solver, dt, in_ptr, out_ptr, fluxes_ptr=None, stream=None):
        """step_fixed from the device state at in_ptr into out_ptr (8 * n
        doubles each, not overlapping), leaving the grid's own state as it
        was; with fluxes_ptr also the step's magnetic fluxes (6 * n, as
        get_magnetic_fluxes)."""
        self._check(self._L.pamhd_step_fixed_from(self._h, solver_id(solver), dt, in_ptr, out_ptr,
                                                  fluxes_ptr, stream))

    def set_state_device(self, ptr, stream=None):
        """Copies 8 * n doubles from a device pointer into the grid."""
        self._check(self._L.pamhd_set_state_device(self._h, ptr, stream))

    def get_state_device(self, ptr, stream=None):
        """Copies the grid's 8 * n doubles to a device pointer."""
        self._check(self._L.pamhd_get_state_device(self._h, ptr, stream))

    def adjoint_step(self, solver, dt, state_ptr, lambda_ptr, stream=None):
        """lambda (device, 8 * n) <- dL/dU^n from dL/dU^{n+1}, for the step of
        length dt that started from the state at state_ptr."""
        self._check(self._L.pamhd_adjoint_step(self._h, solver_id(solver), dt, state_ptr,
                                               lambda_ptr, stream))

    def get_magnetic_fluxes(self, ptr, stream=None):
        """Copies the last step's six transverse magnetic fluxes (dt * F,
        6 * n doubles) to a device pointer, for adjoint_step_fluxes."""
        self._check(self._L.pamhd_get_magnetic_fluxes(self._h, ptr, stream))

    def adjoint_step_fluxes(self, solver, dt, state_ptr, fluxes_ptr, lambda_ptr, stream=None):
        """adjoint_step with the forward step's magnetic fluxes given
        (get_magnetic_fluxes after that step): the same lambda."""
        self._check(self._L.pamhd_adjoint_step_fluxes(self._h, solver_id(solver), dt, state_ptr,
                                                      fluxes_ptr, lambda_ptr, stream))

    def adjoint_steps(self, solver, dts, state_ptrs, flux_ptrs, lambda_ptr, stream=None):
        """Several reverse steps in one call, in the order they run (the
        latest step first), of consecutive steps of one rollout (step q + 1
        ended at state_ptrs[q]); flux_ptrs entries may be None
        (recomputed)."""
        k = len(dts)
        if len(state_ptrs) != k or (flux_ptrs is not None and len(flux_ptrs) != k):
            raise ValueError("dts, state_ptrs and flux_ptrs differ in length")
        d = (ctypes.c_double * k)(*dts)
        sp = (ctypes.c_void_p * k)(*state_ptrs)
        fp = None if flux_ptrs is None else (ctypes.c_void_p * k)(*flux_ptrs)
        self._check(self._L.pamhd_adjoint_steps(self._h, solver_id(solver), k, d, sp, fp,
                                                lambda_ptr, stream))

    STEP_MODES = {"auto": -1, "passes": 0, "fused": 1, "accumulate": 2}

    def set_fused_step(self, mode="auto", planes_per_chunk=0):
        """HLLD's step structure (identical results): "passes" (per-direction
        flux passes and an update pass), "fused" (one pass, then the
        face-field update), "accumulate" (per-direction marches that update
        the state, then the face-field update) or "auto" (accumulate);
        True / False select fused / passes.  planes_per_chunk: planes per fused block (1..32) or cells
        per accumulating march (1..32); 0 keeps the current value."""
        if mode is True or mode is False:
            m = 1 if mode else 0
        elif isinstance(mode, str):
            m = self.STEP_MODES[mode]
        else:
            m = int(mode)
        self._check(self._L.pamhd_set_fused_step(self._h, m, int(planes_per_chunk)))

    def set_slabs(self, slabs):
        """z-slab pipelining of the flux pass and the update (1 = off)."""
        self._check(self._L.pamhd_set_slabs(self._h, int(slabs)))

    def set_adjoint_jacobian(self, reverse=True):
        """the adjoint's flux Jacobian: reverse accumulation (default) or
        forward dual numbers"""
        self._check(self._L.pamhd_set_adjoint_jacobian(self._h, 1 if reverse else 0))

    def set_poisson_smoother(self, fused):
        self._check(self._L.pamhd_set_poisson_smoother(self._h, 1 if fused else 0))

    def max_signal_speeds(self):
        v = (ctypes.c_double * 3)()
        self._check(self._L.pamhd_max_signal_speeds(self._h, v))
        return tuple(v)


def face_flux_batch(solver, L, R, bg=None, adiabatic_index=5.0 / 3.0,
                    vacuum_permeability=1.0, precision="fp64"):
    """Batched Riemann solver on torch CUDA tensors (8, n) in the face frame.

    Returns (F (8, n), smax (n,), status (n,)) as torch tensors on the same
    device.
    """
    import torch
    Lib = _lib.load()
    L = L.contiguous().to(torch.float64)
    R = R.contiguous().to(torch.float64)
    n = L.shape[1]
    F = torch.empty_like(L)
    s = torch.empty(n, dtype=torch.float64, device=L.device)
    st = torch.empty(n, dtype=torch.int32, device=L.device)
    bgp = None
    if bg is not None:
        bg = bg.contiguous().to(torch.float64)
        bgp = bg.data_ptr()
    prec = {"fp64": _lib.FP64, "mixed": _lib.MIXED}[precision]
    stream = torch.cuda.current_stream(L.device).cuda_stream
    _lib.check(Lib.pamhd_face_flux_batch(solver_id(solver), prec, n, L.data_ptr(),
                                         R.data_ptr(), bgp, adiabatic_index,
                                         vacuum_permeability, F.data_ptr(),
                                         s.data_ptr(), st.data_ptr(), stream))
    return F, s, st


def plain1d_run(solver, state, cell_size, adiabatic_index, vacuum_permeability, t_end,
                cfl=0.5, max_steps=1 << 40, stream=None):
    """The 1-D plain method of tests/mhd/reference.cpp on the GPU
    (pamhd_plain1d_run): state is a float64 array (8, ncells) {rho, m, E, B}
    with cell-centred B, advanced in place from t = 0 to t_end.  Returns the
    number of steps."""
    Lib = _lib.load()
    if not (isinstance(state, np.ndarray) and state.dtype == np.float64
            and state.flags.c_contiguous and state.ndim == 2 and state.shape[0] == 8):
        raise ValueError("state must be a C-contiguous float64 array of shape (8, ncells)")
    steps = ctypes.c_long(0)
    _lib.check(Lib.pamhd_plain1d_run(state.shape[1], cell_size, adiabatic_index,
                                     vacuum_permeability, cfl, solver_id(solver), t_end,
                                     max_steps, state.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                     ctypes.byref(steps), stream))
    return steps.value
