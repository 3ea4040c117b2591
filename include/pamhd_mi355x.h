/*
C ABI of the MI355X-native PAMHD ideal-MHD finite-volume step.

The library owns a uniform Cartesian grid of cells in HBM in structure-of-
arrays form and advances it with hand-written HIP kernels for gfx950.  It is
a drop-in for the reference's MHD step:

  reference entry point                          replaced by
  ---------------------------------------------  --------------------------------
  pamhd::mhd::get_flux_rusanov / athena::         pamhd_face_flux_batch
    get_flux_hll / get_flux_hlld / get_flux_roe
    (source/mhd/rusanov.hpp:59-73,
     hll_athena.hpp:57-71, hlld_athena.hpp:54-68,
     roe_athena.hpp:567-581)
  pamhd::mhd::timestep                            pamhd_timestep
    (source/mhd/solve.hpp:1837-1924)
  timestep + apply_boundaries + sync_magnetic_    pamhd_step
    field of the main loop
    (tests/mhd/test.cpp:514-568)
  pamhd::mhd::apply_boundaries                    pamhd_apply_boundaries
    (source/mhd/boundaries.hpp:1249-1297)
  pamhd::math::get_divergence_face2volume         pamhd_div_face2volume
    (source/math/nabla.hpp:58-113)
  pamhd::divergence::remove                       pamhd_clean_divergence
    (source/divergence/remove.hpp:609-769)

Conventions
  * All functions return 0 on success or a PAMHD_E* code; no exception
    crosses the ABI.  pamhd_last_error() gives a message for the last failure
    of a handle (physics errors name the face and the failed check, like the
    diagnostic of source/mhd/solve.hpp:156-181).
  * State arrays exchanged with the host are component-major double arrays of
    8 * n values, n = nx * ny * nz, cell index i + nx * (j + ny * k):
        rho, m_x, m_y, m_z, E, B_x(i-1/2,j,k), B_y(i,j-1/2,k), B_z(i,j,k-1/2)
    i.e. the conserved variables of the cell (Mass_Density, Momentum_Density,
    Total_Energy_Density) and the face magnetic field of the cell's three
    negative-side faces (the reference's Face_Magnetic_Field(-dir), stored
    once per face).  The cell-centred field is the average of the two faces,
    as update_vol_B computes it (source/mhd/solve.hpp:1661-1663).
  * `stream` is a hipStream_t (NULL = the default stream).  Calls are
    stream-ordered; only calls that return host values synchronise.
*/
#ifndef PAMHD_MI355X_H
#define PAMHD_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Riemann solver numbering of source/mhd/common.hpp:54-61 */
enum pamhd_solver {
	PAMHD_RUSANOV = 0,
	PAMHD_RUSANOV_STAGGERED = 1, /* not a separate solver in the reference's dispatch */
	PAMHD_HLL_ATHENA = 2,
	PAMHD_HLLD_ATHENA = 3,
	PAMHD_ROE_ATHENA = 4,
	PAMHD_HYBRID = 5             /* MHD solution not used: fluxes are skipped */
};

/* arithmetic of the face fluxes and wave speeds; the cell state is FP64 in both */
enum pamhd_precision {
	PAMHD_FP64 = 0,   /* bit-compatible with the FP64 CPU path */
	PAMHD_MIXED = 1   /* FP32 fluxes/speeds, FP64 state and pressure recovery */
};

enum pamhd_status {
	PAMHD_OK = 0,
	PAMHD_E_DENSITY = 1,       /* non-positive or non-normal mass density */
	PAMHD_E_PRESSURE = 2,      /* non-positive or non-normal thermal pressure */
	PAMHD_E_SPEED = 3,         /* invalid wave speed */
	PAMHD_E_FLUX = 4,          /* non-finite momentum or flux */
	PAMHD_E_FIELD = 5,         /* non-finite total magnetic field */
	PAMHD_E_ROE_DENSITY = 6,   /* Roe intermediate-state density */
	PAMHD_E_ROE_PRESSURE = 7,  /* Roe intermediate-state pressure */
	PAMHD_E_SOLVER = 8,        /* unknown solver */
	PAMHD_E_POISSON = 9,       /* divergence cleaning: tolerance not reached */
	PAMHD_E_SUBSTEP = 10,      /* substep-min above a cell's allowed level, or a level above 30 */
	PAMHD_E_ARG = 16,          /* invalid argument */
	PAMHD_E_UNSUPPORTED = 17,  /* valid request this build does not implement */
	PAMHD_E_DEVICE = 18,       /* HIP runtime failure (allocation, launch) */
	PAMHD_E_NOLIB = 19
};

typedef struct pamhd_grid_desc {
	int32_t cells[3];        /* nx, ny, nz (grid-options "cells") */
	double start[3];         /* grid-options "start" */
	double cell_length[3];   /* volume / cells, tests/mhd/test.cpp:339-346 */
	int32_t periodic[3];     /* grid-options "periodic" */
} pamhd_grid_desc;

typedef struct pamhd_physics {
	double adiabatic_index;      /* "adiabatic-index" */
	double vacuum_permeability;  /* "vacuum-permeability" */
	double proton_mass;          /* "proton-mass" */
} pamhd_physics;

/*
Solar-wind box description, the options of tests/mhd/solar_wind_box.cpp
(source/solar_wind_box_options.hpp:50-390 and the "background-magnetic-field"
object of source/background_magnetic_field.hpp): one solar-wind state, the
inner (planet) boundary and a constant field plus point dipoles.
*/
#define PAMHD_MAX_DIPOLES 8
typedef struct pamhd_solar_wind_desc {
	int32_t direction;             /* "direction": -1 "-x", +1 "+x", -2, +2, -3, +3 */
	double number_density;         /* "number-density" (1/m^3) */
	double velocity[3];            /* "velocity" (m/s) */
	double pressure;               /* "pressure" (Pa) */
	double magnetic_field[3];      /* "magnetic-field" (T); component along the
	                                  inflow direction must be 0 */
	double inner_radius;           /* "inner-boundary": "radius" (m) */
	double inner_number_density;
	double inner_velocity[3];
	double inner_pressure;
	double bg_value[3];            /* "background-magnetic-field": "value" */
	double bg_min_distance;        /* "min-distance" */
	int32_t n_dipoles;             /* <= PAMHD_MAX_DIPOLES */
	double dipole_moment[PAMHD_MAX_DIPOLES][3];
	double dipole_position[PAMHD_MAX_DIPOLES][3];
} pamhd_solar_wind_desc;

typedef struct pamhd_handle pamhd_handle;
typedef struct pamhd_ensemble pamhd_ensemble;

/* Library/ABI version, for loaders: returns 1. */
int pamhd_abi_version(void);

/* Allocates the device state for a grid.  Fully periodic grids only in this
   version (PAMHD_E_UNSUPPORTED otherwise). */
int pamhd_create(const pamhd_grid_desc* grid, const pamhd_physics* physics,
	int precision, pamhd_handle** out);
int pamhd_destroy(pamhd_handle* h);

/* Host <-> device copies of the 8 * n state described above (synchronous). */
int pamhd_upload_state(pamhd_handle* h, const double* state);
int pamhd_download_state(pamhd_handle* h, double* state);

/* Device pointer of the current 8 * n state (valid until the next step). */
int pamhd_state_device_ptr(pamhd_handle* h, double** ptr);

/* Background field on each cell's positive-side faces (the reference's
   Bg_B(cell)(+dir), source/common_variables.hpp:553-559), 9 * n host doubles:
   x-face components x,y,z, then y-face, then z-face; NULL clears it. */
int pamhd_set_background_faces(pamhd_handle* h, const double* bg);

/*
One step of pamhd::mhd::timestep with global time stepping: the step length
is min(lagged CFL dt, max_time_step) following
source/substepping.hpp:331-338, where the lagged CFL dt is
time_step_factor * cell length / max signal speed of the previous flux pass
(0 before the first pass, or if any face speed was <= 0).  Computes all face
fluxes, the constrained-transport face-field update, the conserved update and
the two constant-pressure energy refits of the reference's timestep.

dt_taken (may be NULL) receives the step length; when non-NULL the call
synchronises and reports device-side physics errors.
*/
int pamhd_timestep(pamhd_handle* h, int solver, double max_time_step,
	double time_step_factor, double* dt_taken, void* stream);

/* Boundary update of pamhd::mhd::apply_boundaries: for a periodic grid the
   constant-pressure energy refit of all cells. */
int pamhd_apply_boundaries(pamhd_handle* h, double simulation_time, void* stream);

/* Fused main-loop step: pamhd_timestep + pamhd_apply_boundaries in one pass
   (three energy refits), as tests/mhd/test.cpp:523-568 does per step. */
int pamhd_step(pamhd_handle* h, int solver, double max_time_step,
	double time_step_factor, double* dt_taken, void* stream);

/*
Program loop of tests/mhd/test.cpp:446-610 on the device-resident state:
apply_boundaries, the zero-length "final init" timestep, then steps of
min(lagged CFL dt, t_end - t) until t >= t_end or max_steps steps.
*/
int pamhd_run(pamhd_handle* h, int solver, double time_step_factor,
	double t_start, double t_end, long max_steps, long* steps_taken,
	double* t_final, void* stream);

/* Synchronises and returns the first device-side physics error, if any. */
int pamhd_check(pamhd_handle* h);

/* Mean |div B| of the face field over all cells (synchronous). */
int pamhd_div_face2volume(pamhd_handle* h, double* avg_div);

/*
Divergence cleaning of the face field, the role of pamhd::divergence::remove
(source/divergence/remove.hpp:609-769) on a periodic grid: solves the
periodic 7-point Poisson problem L phi = div(B_face) by multigrid V-cycles
until the residual's 2-norm (mean removed) is <= tol times the right-hand
side's, or at the roundoff floor 8 eps ||sum_d 2|B_d|/len_d||_2, or
max_cycles cycles; then B_face -= face gradient of phi, so the new face
divergence is the Poisson residual, and E is refit at constant thermal
pressure (update_vol_B, source/mhd/solve.hpp:1646-1672).  Outputs (each may
be NULL): div_before = mean |div B| before cleaning (what remove() returns),
rel_residual, cycles.  Returns PAMHD_E_POISSON when the tolerance was not
reached (the field is still projected).  Synchronises the stream.
The reference computes the divergence of the cell-centred field with central
differences and delegates the solve to dccrg's Poisson solver; this build's
state holds face fields, so it projects those (the constrained-transport
step then keeps the face divergence at roundoff).
*/
int pamhd_clean_divergence(pamhd_handle* h, double tol, int max_cycles,
	double* div_before, double* rel_residual, int* cycles, void* stream);

/* How pamhd_adjoint_step forms the transposed Jacobian of the HLL flux:
   1 (default) by hand-written reverse accumulation, 0 by forward dual
   numbers (16 seeded directions per face; kept as the check of the
   former).  Both give the exact derivative up to rounding.  (No
   counterpart in the reference.) */
int pamhd_set_adjoint_jacobian(pamhd_handle* h, int reverse);

/* Smoother of the Poisson solve: 1 = LDS-tiled fused red-black sweeps where
   every dimension of a level is even and >= 4 (half the bytes of the two
   colour passes, but latency bound: slower on MI355X at 256^3), 0 = one
   launch per colour (default).  Both give identical results.  (No
   counterpart in the reference.) */
int pamhd_set_poisson_smoother(pamhd_handle* h, int fused);

/* Per-cell 2^N substepping of timestep(), source/mhd/solve.hpp:1837-1924
   with source/substepping.hpp:54-399 (options "substep-min" and
   "substep-max" of source/simulation_options.hpp, integer form).  With
   substep_max > 0 pamhd_timestep / pamhd_step give every cell a period of
   2^N substeps from its own CFL limit (at most one level above any face or
   edge neighbour), solve each face at the pace of the smallest period
   around its owner and return the step length (largest period x substep
   length); substep_max = 0 is the global time step.  FP64 periodic grids;
   set it before the zero-dt final init (the per-face speeds it needs are
   measured from then on).  The reference's defaults are substep_min 0 and
   substep_max 999 (simulation_options.hpp:55-56); a computed level above
   30, or a cell whose substep_min exceeds its level (the reference aborts,
   substepping.hpp:376-384), fails the step with PAMHD_E_SUBSTEP. */
int pamhd_set_substepping(pamhd_handle* h, int substep_min, int substep_max);

/* Launch shape of the substepped step: the passes of a substep walk lists
   of the 64 x 4-cell blocks that have work from this many workgroups (0:
   the default, 4096; fewer blocks than that: one workgroup per block).  All
   give identical results.  (No counterpart in the reference.) */
int pamhd_set_substep_walk(pamhd_handle* h, int workgroups);

/* The same with each cell's own options (n ints each, x fastest), as the
   reference evaluates string-valued "substep-min" / "substep-max" at the
   cell centres and the current time before every step
   (set_minmax_substepping_period(time, ...), substepping.hpp:197-255): the
   caller sets them before each step.  Copies the arrays (synchronous).
   From the first call on, steps run on the substepping path whatever the
   options (all periods 1 give the global step), so that the per-face
   speeds behind the lagged time step carry over when the options change. */
int pamhd_set_substepping_cells(pamhd_handle* h, const int* substep_min, const int* substep_max);

/*
Uniform grids with value and copy boundaries and closed dimensions
(tests/mhd/test.cpp with source/boundaries/ and source/mhd/boundaries.hpp).
The caller classifies the cells (Boundaries::classify, boundaries.hpp:
100-226; set_solver_info, mhd/boundaries.hpp:67-245; classify_faces,
:1039-1217) and passes, for n cells in x-fastest order: var_type (4 n:
number density, velocity, pressure, magnetic field; 1 normal, 2 not
solved, 3 value cell, 4 copy cell), cell_type (n: 1, 0, -1) and face_type
(6 n, faces -x, +x, -y, +y, -z, +z of each cell: 1, 0, -1); grid->periodic
gives the periodic dimensions, the others end.  FP64; the state holds 14
doubles per cell (rho, m (3), E, Vol_B (3), the cell's own copies of its
six faces' B), read and written by pamhd_upload_state /
pamhd_download_state.  pamhd_timestep is timestep() with per-cell
substepping over the cell types (pamhd_set_substepping* give the options;
(0, 0) by default: the global step); pamhd_apply_boundaries and the fused
pamhd_step apply the boundaries (apply_boundaries, mhd/boundaries.hpp:
1261-1297) and sync_magnetic_field.  Divergence cleaning, the adjoint,
autotuning and device-pointer state access are not available on such
grids (PAMHD_E_UNSUPPORTED).
*/
int pamhd_create_bounded(const pamhd_grid_desc* grid, const pamhd_physics* physics,
	const signed char* var_type, const signed char* cell_type, const signed char* face_type,
	pamhd_handle** out);

/* What a handle steps: PAMHD_GRID_PERIODIC (pamhd_create),
   PAMHD_GRID_BOUNDED (pamhd_create_bounded) or PAMHD_GRID_SOLAR_WIND_BOX
   (pamhd_create_solar_wind_box). */
enum { PAMHD_GRID_PERIODIC = 0, PAMHD_GRID_BOUNDED = 1, PAMHD_GRID_SOLAR_WIND_BOX = 2 };
int pamhd_grid_kind(pamhd_handle* h, int* kind);

/* The value boundaries for the next pamhd_apply_boundaries / pamhd_step,
   resolved per cell (the last boundary covering a cell wins; the caller
   evaluates time-dependent values at the time the boundaries are applied,
   Value_Boundary::get_data, source/boundaries/value_boundary.hpp): mask
   (4 n, 1 where a variable has a value boundary), number density (n, the
   value before the proton mass factor), velocity (3 n), pressure (n) and
   face field (6 n, the component normal to each face at its centre, used
   on boundary faces).  Synchronous copies. */
int pamhd_set_boundary_values(pamhd_handle* h, const unsigned char* mask,
	const double* number_density, const double* velocity, const double* pressure,
	const double* face_field);

/* Largest period (substeps) of the last step: timestep()'s second return
   value (1 without substepping). */
int pamhd_last_max_substep(pamhd_handle* h, int* max_substep);

/* Each cell's substep period (2^level) of the last substepped step, n ints
   on the host (the reference's per-cell Substep variable,
   source/substepping.hpp:54-96); 1 everywhere without substepping.
   Synchronous. */
int pamhd_last_substep_periods(pamhd_handle* h, int* periods);

/* Faces solved by the last pamhd_timestep / pamhd_step (both cells of type
   >= 0; with substepping each face once per substep it is due):
   timestep()'s third return value flux_calcs (source/mhd/solve.hpp:1918,
   summed by tests/mhd/test.cpp:612-627).  Call after the step has
   completed (a step given dt_taken synchronises). */
int pamhd_last_flux_calcs(pamhd_handle* h, unsigned long long* flux_calcs);

/* Largest face signal speed per dimension of the last flux pass (synchronous). */
int pamhd_max_signal_speeds(pamhd_handle* h, double smax[3]);

/*
Batched Riemann solver on device buffers: L, R (8 * n), bg (3 * n or NULL),
F (8 * n), smax (n), status (n, may be NULL), component-major, states already
rotated into the face frame.  precision selects FP64 or FP32 arithmetic.
Faces with a physics error get status != 0 and zero flux.
*/
int pamhd_face_flux_batch(int solver, int precision, size_t n, const double* L,
	const double* R, const double* bg, double adiabatic_index,
	double vacuum_permeability, double* F, double* smax, int* status,
	void* stream);

/*
The 1-D plain method of tests/mhd/reference.cpp: solve_mhd
(reference.cpp:284-440; fluxes of the faces between neighbouring cells,
the upwind electric field on each cell's +x edge, reference.cpp:393-404)
with apply_fluxes (mhd/common.hpp:707-743) and the main loop
(reference.cpp:830-877: lagged dt = cfl * min(cell_size / max signal
speed), the first step of length 0, until t >= t_end), on the GPU.  The two
end cells receive no fluid flux and the last cell no field change, as in
the reference.  state: host, 8 * ncells FP64 component-major {rho, m (3),
E, B (3)} (cell-centred field), read and overwritten.  At most max_steps
steps (the reference has no such limit).  Errors: the first failing face's
solver error (the state is then that before the failing step) or the first
failing cell's density / pressure check (the state after the step).
(Replaces the reference program's loop; no handle.)
*/
int pamhd_plain1d_run(size_t ncells, double cell_size, double adiabatic_index,
	double vacuum_permeability, double cfl, int solver, double t_end, long max_steps,
	double* state, long* steps_taken, void* stream);

/* Most cells pamhd_plain1d_run_noalloc takes (its scratch: 7 x (cells - 1)
   doubles of face terms and a 256-byte status record, 14.7 MB per GPU). */
#define PAMHD_PLAIN1D_SCRATCH_CELLS 262144

/*
pamhd_plain1d_run without any allocation, for callers that cannot allocate
device or pinned memory (code under HIP graph capture, serving loops,
harnesses that forbid hipMalloc).  No counterpart in the reference.
state_device: device buffer of 8 * ncells FP64 (the layout of
pamhd_plain1d_run), read and overwritten in place on `stream`; the face
terms and the status record live in a fixed scratch array of the library
(one per GPU, held until the process ends), so ncells is limited to
PAMHD_PLAIN1D_SCRATCH_CELLS (larger: PAMHD_E_ARG).  Every call on one GPU
shares that GPU's scratch: calls that overlap on one GPU (other host threads
or streams) race and must be serialised by the caller.  The status record
is reset by a kernel at every step; the step loop reads it back to the host
(a synchronisation per step, as pamhd_plain1d_run).  Same results and
errors as pamhd_plain1d_run.
*/
int pamhd_plain1d_run_noalloc(size_t ncells, double cell_size, double adiabatic_index,
	double vacuum_permeability, double cfl, int solver, double t_end, long max_steps,
	double* state_device, long* steps_taken, void* stream);

/*
Autotuner: times the launch variants of the flux and update passes that give
bit-identical results (block order, per-cell vs z-marching update and its
planes per block) on the handle's current state with the given solver, keeps
the fastest for later steps and writes a report to buf (may be NULL).  The
state and the lagged time step are not changed.  Until it (or
pamhd_set_launch) is called a handle uses the variants it picks for the
256^3 HLLD step: XCD-aware flux block order, HLLD's per-cell part carried
along 16-cell marches in all three directions, the z-marching update with 8
planes per 128 x 2 tile (grids of fewer than 16 planes: per-face flux
kernels and the per-cell update).  (No counterpart in the reference.)
*/
int pamhd_autotune(pamhd_handle* h, int solver, char* buf, size_t buflen,
	void* stream);

/* Sets the launch variants directly: flux_remap bit 0 = XCD-aware block
   order of the flux pass, bit 1 = planes fastest in the block order, bit 2 =
   flux kernels built for 4 blocks per CU, bit 3 = HLLD's per-cell part
   computed once per cell and shared through LDS, bit 4 = that part carried
   along marches of 16 cells for the y- and z-faces (bit 5: 4 cells), bit 6
   = as bit 4 and the x-faces marched along y too, each lane's part moved to
   the next lane (bits 3-6 for HLLD, bits 4-6 also for HLL and Rusanov
   on grids without a background field; bit 1 is then ignored);
   update_remap bits 0 and 1 the same
   orders for the update pass (bit 0 only for the z-marching update) and,
   for the z-marching update, bits 2-4 its tile (0: 64 x 4, 1: 128 x 2,
   2: 32 x 8, 3: 128 x 4, 4: 64 x 8 threads); update_kz = 0 for the per-cell update or the planes
   per block of the z-marching update.  All give identical results. */
int pamhd_set_launch(pamhd_handle* h, int flux_remap, int update_kz, int update_remap);

/* HLLD's step structure (no counterpart in the reference; all give
   identical results): fused = 0 per-direction flux passes that store every
   face flux, then an update pass; 1 one pass computing every face flux and
   the conserved update of each cell without a flux array; 2 per-direction
   marches, each adding its faces' flux differences to the new state in
   place (no flux array); 1 and 2 keep only the two field components of
   each face the constrained-transport edge terms use (6 doubles per cell)
   and finish with a pass that updates the face field and refits the
   energy.  -1 (the default) picks 2; pamhd_autotune times them all.
   planes_per_chunk (0 = keep) is the z-extent of mode 1's blocks (1..32)
   or the cells per march of mode 2 (1..32). */
int pamhd_set_fused_step(pamhd_handle* h, int fused, int planes_per_chunk);

/* Pipelines the flux pass and the per-cell update over `slabs` z-slabs on
   two streams (1 = no pipelining; capped at nz / 2; ignored while kernel
   timing is on or with the z-marching update).  Identical results.
   (No counterpart in the reference.) */
int pamhd_set_slabs(pamhd_handle* h, int slabs);

/*
Per-kernel device-time accounting: when enabled, each launch of the handle is
bracketed by HIP events on its stream.  pamhd_kernel_time synchronises and
returns the accumulated milliseconds and launch count of one kernel:
"dt", "flux_pass", "update" or "refit".  Enabling resets the totals.
(No counterpart in the reference, which has no timers; used by the
benchmark for the roofline figure.)
*/
int pamhd_enable_timing(pamhd_handle* h, int enable);
int pamhd_kernel_time(pamhd_handle* h, const char* kernel, double* total_ms,
	long* launches);

/* ---- frozen-time-step stepping and its adjoint ---- */

/* One main-loop step (as pamhd_step) of exactly dt, without the CFL
   limit: the frozen time-step schedule a gradient is taken along.
   (No counterpart in the reference.) */
int pamhd_step_fixed(pamhd_handle* h, int solver, double dt, void* stream);

/*
pamhd_step_fixed from the caller's device state state_in (8 * n) into
state_out (8 * n, not overlapping it); the grid's own state is left as it
was.  With fluxes non-NULL the step also writes its magnetic fluxes there
(6 * n, the layout of pamhd_get_magnetic_fluxes) from the update pass, for
which the step must keep its FP64 flux array (not mixed precision, not
HLLD's marches).  For recomputing a checkpoint segment in a reverse pass
without copies.  Periodic grids only (not solar-wind boxes or grids with
boundaries); like pamhd_step_fixed it never substeps.  (No counterpart in
the reference.)
*/
int pamhd_step_fixed_from(pamhd_handle* h, int solver, double dt, const double* state_in,
	double* state_out, double* fluxes, void* stream);

/* Stream-ordered device-to-device copies of the 8 * n state. */
int pamhd_set_state_device(pamhd_handle* h, const double* state, void* stream);
int pamhd_get_state_device(pamhd_handle* h, double* state, void* stream);

/*
Reverse step of pamhd_step_fixed: given the state U^n (device, 8 * n) the
forward step of length dt started from and lambda = dL/dU^{n+1} (device,
8 * n, same layout), overwrites lambda with dL/dU^n.  HLL fluxes, FP64,
periodic grid without background field.  Uses about 14 * n doubles of
scratch (plus the tiles' boundary terms) allocated on first use.  (No counterpart in the reference, which has no
derivatives.)
*/
int pamhd_adjoint_step(pamhd_handle* h, int solver, double dt, const double* state,
	double* lambda, void* stream);

/*
The magnetic fluxes of the last step, for the adjoint of that step: copies
the six transverse field components of dt * F (x-faces: B_y, B_z; y-faces:
B_z, B_x; z-faces: B_x, B_y; 6 * n doubles, the cell owning the face on its
negative side) into `fluxes` (device memory).  Valid after an FP64 step
that kept the step's flux array (Rusanov, HLL, Roe without substepping or
boundaries); PAMHD_E_UNSUPPORTED otherwise.  (No counterpart in the
reference: an adjoint helper.)
*/
int pamhd_get_magnetic_fluxes(pamhd_handle* h, double* fluxes, void* stream);

/*
pamhd_adjoint_step with the step's magnetic fluxes given (the output of
pamhd_get_magnetic_fluxes after the forward step from `state`) instead of
recomputed from the state: the same lambda bit for bit, one kernel fewer.
(No counterpart in the reference.)
*/
int pamhd_adjoint_step_fluxes(pamhd_handle* h, int solver, double dt, const double* state,
	const double* fluxes, double* lambda, void* stream);

/*
nsteps reverse steps in one call, in the order they run (the latest step of
the rollout first): dts, states and fluxes are host arrays of nsteps entries,
states[q] the device state the step q started from, fluxes[q] its magnetic
fluxes (pamhd_get_magnetic_fluxes) or NULL (recomputed; fluxes itself may be
NULL).  The steps are consecutive steps of one rollout: step q + 1 ended at
states[q], whose face field is that step's constrained-transport result.
Where the next step's fluxes are given, the last kernel of a step also makes
the next step's first-stage terms, one launch fewer per step, reading its new
faces from states[q].
The same lambda as nsteps calls of pamhd_adjoint_step, bit for bit.  (No
counterpart in the reference.)
*/
int pamhd_adjoint_steps(pamhd_handle* h, int solver, int nsteps, const double* dts,
	const double* const* states, const double* const* fluxes, double* lambda, void* stream);

/* ---- solar-wind box (tests/mhd/solar_wind_box.cpp) ---- */

/*
Creates a non-periodic uniform grid (grid->periodic all 0, >= 4 cells per
dimension, the outer layer of cells being the boundary) set up as the
reference's solar-wind program does before its final-init step: cell
classification (prepare_grid, source/grid/solar_wind_box.hpp:193-406),
background field at face centres, initialize_plasma
(source/mhd/solar_wind_box.hpp:62-205), update_vol_B at constant pressure,
apply_boundaries_sw_box and sync_magnetic_field
(tests/mhd/solar_wind_box.cpp:430-461).  FP64.  On the handle:
  pamhd_timestep        timestep() (source/mhd/solve.hpp:1837-1924)
  pamhd_step            timestep() + apply_boundaries_sw_box + sync, the
                        main-loop step of tests/mhd/solar_wind_box.cpp:514-553
  pamhd_apply_boundaries  apply_boundaries_sw_box + sync
  pamhd_run             final-init timestep(0) then pamhd_step to t_end
  pamhd_upload/download_state  14 * n doubles: rho, m (3), E, Vol_B (3),
                        the cell's own copies of its -x, +x, -y, +y, -z, +z
                        face B (the reference keeps one copy per cell)
The time step is global ("substep-max": 0).
*/
int pamhd_create_solar_wind_box(const pamhd_grid_desc* grid, const pamhd_physics* physics,
	const pamhd_solar_wind_desc* solar_wind, pamhd_handle** out);

/* values per cell of pamhd_upload/download_state: 8 (periodic) or 14 (box) */
int pamhd_state_components(pamhd_handle* h, int* components);

/* Cell_Type (1 normal, 0 boundary, -1 not solved) and the boundary rule of
   each cell (1..27: copy from the neighbour at offset (o0, o1, o2) with code
   1 + (o0+1) + 3 (o1+1) + 9 (o2+1); 28 planet; 29 inflow); either may be NULL */
int pamhd_box_cell_types(pamhd_handle* h, signed char* cell_type, unsigned char* rule);

/* Bg_B(cell)(+d) of a box, 9 * n doubles as pamhd_set_background_faces */
int pamhd_box_background(pamhd_handle* h, double* bg);

/*
Ensemble of solar-wind boxes of different sizes advanced together: every
stage of a step is one launch over all boxes, each box with its own lagged
time step.  (No counterpart in the reference, which runs one box per
program.)
  create       as pamhd_create_solar_wind_box for each box
  timestep     timestep() of every box (the final init uses max_dt = 0)
  run          `steps` main-loop steps of every box (no time limit); with
               use_graph = 1 the steps are replayed from a captured HIP
               graph, with use_graph = 2 they all run in one persistent
               cooperative launch (stages separated by grid-wide barriers
               whose waits are bounded: a timeout returns PAMHD_E_DEVICE;
               synchronous); every mode gives the same results, and
               the steps of one call run as chains whose steps leave
               their last face sync to the next step's update
  check        synchronises; per-box time step of the last step and the
               first physics error
  autotune     picks the fastest tile shape of the cell kernels (64x4, 32x8,
               16x16) on a copy of the state; set_tile sets it
*/
int pamhd_ensemble_create(int nbox, const pamhd_grid_desc* grids, const pamhd_physics* physics,
	const pamhd_solar_wind_desc* solar_winds, pamhd_ensemble** out);
int pamhd_ensemble_destroy(pamhd_ensemble* e);
int pamhd_ensemble_timestep(pamhd_ensemble* e, int solver, double max_time_step,
	double time_step_factor, void* stream);
int pamhd_ensemble_run(pamhd_ensemble* e, int solver, double time_step_factor, long steps,
	int use_graph, void* stream);
int pamhd_ensemble_check(pamhd_ensemble* e, double* dt_per_box, void* stream);
/* How long a persistent run's barrier waits (use_graph = 2): polls before
   a workgroup gives up, sets the run's error flag and leaves (0: the
   default, 2^24 polls, about a second); every workgroup that sees the flag
   leaves too and the run returns PAMHD_E_DEVICE.  For callers that bound
   their wait, and for testing that the run gives up.  (No counterpart in
   the reference.) */
int pamhd_ensemble_set_barrier_polls(pamhd_ensemble* e, long polls);
int pamhd_ensemble_download(pamhd_ensemble* e, int box, double* state);
int pamhd_ensemble_cells(pamhd_ensemble* e, int box, long long* cells);
int pamhd_ensemble_set_tile(pamhd_ensemble* e, int tile);
int pamhd_ensemble_autotune(pamhd_ensemble* e, int solver, double time_step_factor,
	char* buf, size_t buflen, void* stream);
/* last error of e, or of the last failed pamhd_ensemble_create when e is NULL */
int pamhd_ensemble_last_error(pamhd_ensemble* e, char* buf, size_t n);

/* Copies the last error message of h (or of the library when h is NULL). */
int pamhd_last_error(pamhd_handle* h, char* buf, size_t n);

#ifdef __cplusplus
}
#endif

#endif
