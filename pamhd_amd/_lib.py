"""ctypes binding of the C ABI in include/pamhd_mi355x.h.

The library is built in-tree (pamhd_amd/lib/libpamhd_mi355x.so) by
``python -m pamhd_amd.build``.  There is no CPU fallback: every entry point
runs the HIP kernels, and loading fails loudly when the library is missing.
"""
import ctypes
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "lib", "libpamhd_mi355x.so")

# reference solver numbering, source/mhd/common.hpp:54-61
RUSANOV, RUSANOV_STAGGERED, HLL_ATHENA, HLLD_ATHENA, ROE_ATHENA, HYBRID = range(6)
FP64, MIXED = 0, 1
PAMHD_E_POISSON = 9
PAMHD_E_SUBSTEP = 10
PAMHD_E_ARG = 16
PAMHD_E_UNSUPPORTED = 17
# pamhd_grid_kind
GRID_KINDS = ("periodic", "bounded", "solar wind box")

EXPORTS = (
    "pamhd_abi_version", "pamhd_create", "pamhd_destroy", "pamhd_upload_state",
    "pamhd_download_state", "pamhd_state_device_ptr", "pamhd_set_background_faces",
    "pamhd_timestep", "pamhd_apply_boundaries", "pamhd_step", "pamhd_run",
    "pamhd_check", "pamhd_div_face2volume", "pamhd_clean_divergence",
    "pamhd_set_poisson_smoother", "pamhd_max_signal_speeds",
    "pamhd_face_flux_batch", "pamhd_autotune", "pamhd_set_launch", "pamhd_enable_timing", "pamhd_kernel_time",
    "pamhd_last_error", "pamhd_create_solar_wind_box", "pamhd_state_components",
    "pamhd_create_bounded", "pamhd_set_boundary_values", "pamhd_set_adjoint_jacobian",
    "pamhd_grid_kind", "pamhd_plain1d_run", "pamhd_plain1d_run_noalloc",
    "pamhd_box_cell_types", "pamhd_box_background", "pamhd_ensemble_create",
    "pamhd_ensemble_destroy", "pamhd_ensemble_timestep", "pamhd_ensemble_run",
    "pamhd_ensemble_check", "pamhd_ensemble_set_barrier_polls", "pamhd_ensemble_download", "pamhd_ensemble_cells",
    "pamhd_ensemble_set_tile", "pamhd_ensemble_autotune", "pamhd_ensemble_last_error",
    "pamhd_step_fixed", "pamhd_step_fixed_from", "pamhd_set_state_device", "pamhd_get_state_device",
    "pamhd_adjoint_step", "pamhd_get_magnetic_fluxes", "pamhd_adjoint_step_fluxes", "pamhd_adjoint_steps", "pamhd_set_slabs", "pamhd_set_fused_step", "pamhd_set_substepping", "pamhd_set_substep_walk", "pamhd_set_substepping_cells",
    "pamhd_last_max_substep", "pamhd_last_flux_calcs", "pamhd_last_substep_periods",
)


class GridDesc(ctypes.Structure):
    _fields_ = [("cells", ctypes.c_int32 * 3), ("start", ctypes.c_double * 3),
                ("cell_length", ctypes.c_double * 3), ("periodic", ctypes.c_int32 * 3)]


class Physics(ctypes.Structure):
    _fields_ = [("adiabatic_index", ctypes.c_double),
                ("vacuum_permeability", ctypes.c_double),
                ("proton_mass", ctypes.c_double)]


MAX_DIPOLES = 8


class SolarWindDesc(ctypes.Structure):
    """pamhd_solar_wind_desc of include/pamhd_mi355x.h"""
    _fields_ = [("direction", ctypes.c_int32),
                ("number_density", ctypes.c_double),
                ("velocity", ctypes.c_double * 3),
                ("pressure", ctypes.c_double),
                ("magnetic_field", ctypes.c_double * 3),
                ("inner_radius", ctypes.c_double),
                ("inner_number_density", ctypes.c_double),
                ("inner_velocity", ctypes.c_double * 3),
                ("inner_pressure", ctypes.c_double),
                ("bg_value", ctypes.c_double * 3),
                ("bg_min_distance", ctypes.c_double),
                ("n_dipoles", ctypes.c_int32),
                ("dipole_moment", (ctypes.c_double * 3) * MAX_DIPOLES),
                ("dipole_position", (ctypes.c_double * 3) * MAX_DIPOLES)]


class PamhdError(RuntimeError):
    def __init__(self, code, message):
        super().__init__("pamhd error %d: %s" % (code, message))
        self.code = code


_lib = None
_dp = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p


def load():
    """Load the HIP library; raises if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("pamhd_amd HIP library not built: %s "
                          "(run python -m pamhd_amd.build)" % LIB_PATH)
    L = ctypes.CDLL(LIB_PATH)
    H = ctypes.c_void_p
    L.pamhd_abi_version.restype = ctypes.c_int
    L.pamhd_create.argtypes = [ctypes.POINTER(GridDesc), ctypes.POINTER(Physics),
                               ctypes.c_int, ctypes.POINTER(H)]
    _bp = ctypes.POINTER(ctypes.c_byte)
    _dpp = ctypes.POINTER(ctypes.c_double)
    L.pamhd_create_bounded.argtypes = [ctypes.POINTER(GridDesc), ctypes.POINTER(Physics),
                                       _bp, _bp, _bp, ctypes.POINTER(H)]
    L.pamhd_set_boundary_values.argtypes = [H, ctypes.POINTER(ctypes.c_ubyte), _dpp, _dpp,
                                            _dpp, _dpp]
    L.pamhd_set_adjoint_jacobian.argtypes = [H, ctypes.c_int]
    L.pamhd_grid_kind.argtypes = [H, ctypes.POINTER(ctypes.c_int)]
    L.pamhd_plain1d_run.argtypes = [ctypes.c_size_t, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                    ctypes.c_double, ctypes.c_long, _dpp,
                                    ctypes.POINTER(ctypes.c_long), ctypes.c_void_p]
    L.pamhd_plain1d_run_noalloc.argtypes = [ctypes.c_size_t, ctypes.c_double, ctypes.c_double,
                                            ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                            ctypes.c_double, ctypes.c_long, _vp,
                                            ctypes.POINTER(ctypes.c_long), ctypes.c_void_p]
    L.pamhd_destroy.argtypes = [H]
    L.pamhd_upload_state.argtypes = [H, _vp]
    L.pamhd_download_state.argtypes = [H, _vp]
    L.pamhd_state_device_ptr.argtypes = [H, ctypes.POINTER(_vp)]
    L.pamhd_set_background_faces.argtypes = [H, _vp]
    for name in ("pamhd_timestep", "pamhd_step"):
        getattr(L, name).argtypes = [H, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                     _dp, _vp]
    L.pamhd_apply_boundaries.argtypes = [H, ctypes.c_double, _vp]
    L.pamhd_run.argtypes = [H, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                            ctypes.c_double, ctypes.c_long,
                            ctypes.POINTER(ctypes.c_long), _dp, _vp]
    L.pamhd_check.argtypes = [H]
    L.pamhd_div_face2volume.argtypes = [H, _dp]
    L.pamhd_clean_divergence.argtypes = [H, ctypes.c_double, ctypes.c_int, _dp, _dp,
                                         ctypes.POINTER(ctypes.c_int), _vp]
    L.pamhd_set_poisson_smoother.argtypes = [H, ctypes.c_int]
    L.pamhd_max_signal_speeds.argtypes = [H, _dp]
    L.pamhd_face_flux_batch.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_size_t,
                                        _vp, _vp, _vp, ctypes.c_double, ctypes.c_double,
                                        _vp, _vp, _vp, _vp]
    L.pamhd_autotune.argtypes = [H, ctypes.c_int, ctypes.c_char_p, ctypes.c_size_t, _vp]
    L.pamhd_set_launch.argtypes = [H, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.pamhd_set_substepping.argtypes = [H, ctypes.c_int, ctypes.c_int]
    L.pamhd_set_substepping_cells.argtypes = [H, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
    L.pamhd_last_max_substep.argtypes = [H, ctypes.POINTER(ctypes.c_int)]
    L.pamhd_last_substep_periods.argtypes = [H, _vp]
    L.pamhd_last_flux_calcs.argtypes = [H, ctypes.POINTER(ctypes.c_ulonglong)]
    L.pamhd_set_slabs.argtypes = [H, ctypes.c_int]
    L.pamhd_set_fused_step.argtypes = [H, ctypes.c_int, ctypes.c_int]
    L.pamhd_enable_timing.argtypes = [H, ctypes.c_int]
    L.pamhd_kernel_time.argtypes = [H, ctypes.c_char_p, _dp, ctypes.POINTER(ctypes.c_long)]
    L.pamhd_last_error.argtypes = [H, ctypes.c_char_p, ctypes.c_size_t]
    L.pamhd_create_solar_wind_box.argtypes = [ctypes.POINTER(GridDesc), ctypes.POINTER(Physics),
                                              ctypes.POINTER(SolarWindDesc), ctypes.POINTER(H)]
    L.pamhd_state_components.argtypes = [H, ctypes.POINTER(ctypes.c_int)]
    L.pamhd_box_cell_types.argtypes = [H, _vp, _vp]
    L.pamhd_box_background.argtypes = [H, _vp]
    E = ctypes.c_void_p
    L.pamhd_ensemble_create.argtypes = [ctypes.c_int, ctypes.POINTER(GridDesc),
                                        ctypes.POINTER(Physics), ctypes.POINTER(SolarWindDesc),
                                        ctypes.POINTER(E)]
    L.pamhd_ensemble_destroy.argtypes = [E]
    L.pamhd_ensemble_timestep.argtypes = [E, ctypes.c_int, ctypes.c_double, ctypes.c_double, _vp]
    L.pamhd_ensemble_run.argtypes = [E, ctypes.c_int, ctypes.c_double, ctypes.c_long,
                                     ctypes.c_int, _vp]
    L.pamhd_ensemble_check.argtypes = [E, _vp, _vp]
    L.pamhd_ensemble_set_barrier_polls.argtypes = [E, ctypes.c_long]
    L.pamhd_ensemble_download.argtypes = [E, ctypes.c_int, _vp]
    L.pamhd_ensemble_cells.argtypes = [E, ctypes.c_int, ctypes.POINTER(ctypes.c_longlong)]
    L.pamhd_ensemble_set_tile.argtypes = [E, ctypes.c_int]
    L.pamhd_ensemble_autotune.argtypes = [E, ctypes.c_int, ctypes.c_double, ctypes.c_char_p,
                                          ctypes.c_size_t, _vp]
    L.pamhd_ensemble_last_error.argtypes = [E, ctypes.c_char_p, ctypes.c_size_t]
    L.pamhd_set_substep_walk.argtypes = [H, ctypes.c_int]
    L.pamhd_step_fixed.argtypes = [H, ctypes.c_int, ctypes.c_double, _vp]
    L.pamhd_step_fixed_from.argtypes = [H, ctypes.c_int, ctypes.c_double, _vp, _vp, _vp, _vp]
    L.pamhd_set_state_device.argtypes = [H, _vp, _vp]
    L.pamhd_get_state_device.argtypes = [H, _vp, _vp]
    L.pamhd_adjoint_step.argtypes = [H, ctypes.c_int, ctypes.c_double, _vp, _vp, _vp]
    L.pamhd_get_magnetic_fluxes.argtypes = [H, _vp, _vp]
    L.pamhd_adjoint_step_fluxes.argtypes = [H, ctypes.c_int, ctypes.c_double, _vp, _vp, _vp, _vp]
    L.pamhd_adjoint_steps.argtypes = [H, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                      ctypes.POINTER(_vp), ctypes.POINTER(_vp), _vp, _vp]
    for name in EXPORTS:
        getattr(L, name).restype = ctypes.c_int
    _lib = L
    return L


def last_error(handle=None):
    buf = ctypes.create_string_buffer(512)
    load().pamhd_last_error(handle, buf, 512)
    return buf.value.decode()


def check(code, handle=None):
    if code != 0:
        raise PamhdError(code, last_error(handle))
    return code


def ensemble_error(ens=None):
    buf = ctypes.create_string_buffer(512)
    load().pamhd_ensemble_last_error(ens, buf, 512)
    return buf.value.decode()


def check_ensemble(code, ens=None):
    if code != 0:
        raise PamhdError(code, ensemble_error(ens))
    return code
