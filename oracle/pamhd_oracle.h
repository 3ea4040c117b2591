/*
FP64 CPU oracle for the PAMHD ideal-MHD finite-volume step.

TEST INFRASTRUCTURE ONLY.  Nothing in the product path (the HIP library,
its C ABI, the Python package) may link, load or call this code; only the
tests, the smoke check and the CPU-baseline leg of the benchmark use it,
and only as the checker.

Every function restates the reference's arithmetic in the same operation
order (no FMA contraction: build with -ffp-contract=off) so that the GPU's
FP64 path can be compared to it bit for bit.  Each function cites the
reference file:line it follows (paths relative to the reference checkout).

State vectors in the face-normal frame are 8 doubles:
    {rho, m0, m1, m2, E, B0, B1, B2}   (index 0 is the face normal)
*/
#ifndef PAMHD_ORACLE_H
#define PAMHD_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* solver numbering of source/mhd/common.hpp:54-61 */
enum {
	OR_RUSANOV = 0,
	OR_RUSANOV_STAGGERED = 1,
	OR_HLL_ATHENA = 2,
	OR_HLLD_ATHENA = 3,
	OR_ROE_ATHENA = 4,
	OR_HYBRID = 5
};

/* error classes of the std::domain_error throw sites of the solvers */
enum {
	OR_OK = 0,
	OR_ERR_DENSITY = 1,      /* non-positive / non-normal mass density */
	OR_ERR_PRESSURE = 2,     /* non-positive / non-normal thermal pressure */
	OR_ERR_SPEED = 3,        /* invalid wave speed, squared speeds or to_sqrt */
	OR_ERR_FLUX = 4,         /* non-finite momentum or flux */
	OR_ERR_FIELD = 5,        /* non-finite total magnetic field */
	OR_ERR_ROE_DENSITY = 6,  /* Roe intermediate state density */
	OR_ERR_ROE_PRESSURE = 7, /* Roe intermediate state pressure */
	OR_ERR_SOLVER = 8,       /* unknown solver */
	OR_ERR_SUBSTEP = 10      /* a cell's substep-min above its allowed level, or a level above 30 */
};

/* ---- physics helpers, source/mhd/common.hpp ---- */
int or_get_pressure(const double U[8], double gamma, double mu0, double* p);
double or_total_energy(double rho, const double v[3], double p, const double B[3],
	double gamma, double mu0);
int or_fast_speed(const double U[8], const double bg[3], double gamma, double mu0,
	double* cf);
int or_physical_flux(const double U[8], const double bg[3], double gamma,
	double mu0, double F[8]);

/* ---- Riemann solvers: L = negative side, R = positive side ---- */
int or_flux_rusanov(const double L[8], const double R[8], const double bg[3],
	double gamma, double mu0, double F[8], double* smax);
int or_flux_hll(const double L[8], const double R[8], const double bg[3],
	double gamma, double mu0, double F[8], double* smax);
int or_flux_hlld(const double L[8], const double R[8], const double bg[3],
	double gamma, double mu0, double F[8], double* smax);
int or_flux_roe(const double L[8], const double R[8], const double bg[3],
	double gamma, double mu0, double F[8], double* smax);
int or_flux(int solver, const double L[8], const double R[8], const double bg[3],
	double gamma, double mu0, double F[8], double* smax);

/* Roe eigensystem, exposed for the L*R = I check */
void or_roe_eigensystem(double d, double v1, double v2, double v3, double h,
	double b1, double b2, double b3, double x, double y, double gamma,
	double ev[7], double rem[7][7], double lem[7][7]);

/*
Batched face flux in structure-of-arrays form: L, R are 8*n (component-major),
bg is 3*n or NULL, F is 8*n, smax and status are n.  Returns number of faces
whose status is non-zero.
*/
size_t or_face_flux_batch(int solver, size_t n, const double* L, const double* R,
	const double* bg, double gamma, double mu0, double* F, double* smax,
	int* status);

/* ---- uniform periodic staggered grid (Appendix A of the restatement) ---- */
typedef struct {
	int nx, ny, nz;
	double dx, dy, dz;
	double gamma, mu0;
	/* background field per face owned by the cell, 9*n (face x comps, face y
	   comps, face z comps; each component-major) or NULL for zero */
	const double* bg;
} or_grid;

/*
Cell arrays, 8*n component-major:
    rho, mx, my, mz, E, Bx(i-1/2,j,k), By(i,j-1/2,k), Bz(i,j,k-1/2)
with cell index i + nx*(j + ny*k).
*/

/*
One time step of length dt on a fully periodic grid of normal cells:
face fluxes, flux slots, constrained-transport face dB, conserved update,
and `refits` constant-pressure energy refits (the reference makes three per
main-loop step).  Writes the new state to out (may not alias in).

smax_dim[d] receives the largest signal speed over d-faces and zero_dim[d]
the number of d-faces whose speed is <= 0 (either may be NULL).
Returns 0 or the first error code; *bad_face gets its flat face index
(dir*n + cell) when not NULL.
*/
int or_step_periodic(const or_grid* g, int solver, double dt, int refits,
	const double* in, double* out, double smax_dim[3], long zero_dim[3],
	long* bad_face);

/* lagged CFL time step: min over dims of cfl*len/smax, 0 if any speed <= 0 */
double or_cfl_dt(const or_grid* g, double cfl, const double smax_dim[3],
	const long zero_dim[3]);

/* volume-averaged |div B| over all cells, source/math/nabla.hpp:58-113 */
double or_div_face2volume(const or_grid* g, const double* state);

/*
Program driver equivalent to tests/mhd/test.cpp:446-610 for a periodic grid:
one refit for apply_boundaries, the zero-dt final-init step, then steps of
dt = min(lagged CFL dt, t_end - t) until t >= t_end or max_steps steps.
state is updated in place.  Returns 0 or an error code; *steps_taken and
*t_final receive the step count and time.
*/
int or_run_periodic(const or_grid* g, int solver, double cfl, double t_start,
	double t_end, long max_steps, double* state, long* steps_taken,
	double* t_final);

/*
Per-cell 2^N substepping (timestep() of source/mhd/solve.hpp:1837-1924 with
source/substepping.hpp).  or_substep_periods computes each cell's period
(2^N), the smallest period among each cell and its face and edge neighbours
(minsub), the substep length and the largest period from the per-face
signal speeds maxv (3 * n, face on each cell's negative side; <= 0 = not
yet known).  or_step_substep advances `state` (face-once layout) by one
step of *dt_out = max period * substep length, updating maxv; it makes the
reference's in-step refits (one per update of a cell, one after the face
sync), not apply_boundaries' (or_refit_periodic).  or_run_periodic_substep
is or_run_periodic with substepping.  substep_min and substep_max hold each
cell's values of the options "substep-min" and "substep-max" (n ints; the
run loop takes one value for all cells).
*/
int or_substep_periods(const or_grid* g, double time_step_factor, double max_dt,
	const int* substep_min, const int* substep_max, const double* maxv, int* period,
	int* minsub, double* sub_dt, int* max_period);
int or_step_substep(const or_grid* g, int solver, double time_step_factor, double max_dt,
	const int* substep_min, const int* substep_max, double* state, double* maxv,
	double* dt_out, int* max_period_out, long* bad_face);
void or_refit_periodic(const or_grid* g, double* state);
int or_run_periodic_substep(const or_grid* g, int solver, double cfl, int substep_min,
	int substep_max, double t_start, double t_end, long max_steps, double* state,
	long* steps_taken, double* t_final, int* last_max_period);

/* ---- 1-D plain method of tests/mhd/reference.cpp ---- */
typedef struct {
	size_t ncells;           /* including the two end cells */
	double cell_size;
	double gamma, mu0;
	double cfl;              /* 0.5 in the reference */
} or_plain1d;

/*
state: 8*ncells component-major {rho, m0, m1, m2, E, B0, B1, B2}.
Runs from t = 0 until t >= t_end with the reference's lagged dt.
*/
int or_plain1d_run(const or_plain1d* p, int solver, double t_end, double* state,
	long* steps_taken);

/*
Periodic 7-point Poisson solve L phi = rhs by geometric multigrid V-cycles
until ||residual - mean||_2 <= max(tol ||rhs - mean||_2, atol) (or
max_cycles); rhs must have zero mean up to roundoff (a face divergence does).
Divergence cleaning passes atol = 8 eps ||sum_d 2|B_d|/h_d||_2, the roundoff
floor of the face divergence.
Returns 0, or OR_ERR_SOLVER when the tolerance was not reached.
*/
int or_poisson_periodic(const int dims[3], const double h[3], const double* rhs,
	double* phi, double tol, double atol, int max_cycles, double* rel_residual,
	int* cycles);

/*
Divergence cleaning of the face field (in place): Poisson solve on the face
divergence, face-gradient projection, constant-pressure energy refit.
*/
int or_clean_divergence(const or_grid* g, double* state, double tol, int max_cycles,
	double* rel_residual, int* cycles);

/* ---- solar-wind box (tests/mhd/solar_wind_box.cpp with a uniform grid) ---- */

/* boundary rule of a cell: 1..27 = copy from the neighbour at offset
   (o0, o1, o2), code 1 + (o0 + 1) + 3 (o1 + 1) + 9 (o2 + 1), mirrored across
   the walls with a non-zero offset (face, edge and vertex cells of
   source/mhd/solar_wind_box.hpp:386-520); planet and inflow cells */
enum { OR_RULE_NONE = 0, OR_RULE_PLANET = 28, OR_RULE_SW = 29 };

typedef struct {
	int nx, ny, nz;             /* cells, including the one-cell boundary layer */
	double start[3], len[3];    /* grid start and cell size */
	double gamma, mu0, mp;
	int sw_dir;                 /* wall of the inflow cells: +-1, +-2, +-3 */
	double sw_n, sw_v[3], sw_p, sw_B[3];           /* solar-wind state */
	double inner_radius, inner_n, inner_v[3], inner_p;
	const double* bg;           /* 9 * n Bg_B(cell)(+d) or NULL */
	const signed char* ctype;   /* n: 1 normal, 0 boundary, -1 not solved */
	const unsigned char* rule;  /* n: OR_RULE_* or a copy code */
} or_box;

/*
Box cell state, 14 * n component-major: rho, m (3), E, Vol_B (3), then the
cell's own copies of its faces' B: -x, +x, -y, +y, -z, +z.
*/

/* prepare_grid's classification, source/grid/solar_wind_box.hpp:262-406 */
int or_box_classify(const or_box* b, signed char* ctype, unsigned char* rule);

/* Bg_B(cell)(+d) at face centres from a constant field and point dipoles
   (moment, position: 3 * ndip each), background_magnetic_field.hpp:331-494 */
void or_box_background(const or_box* b, const double constant[3], double min_distance,
	int ndip, const double* moment, const double* position, double* bg);

/* initialize_plasma + update_vol_B + apply_boundaries_sw_box +
   sync_magnetic_field, tests/mhd/solar_wind_box.cpp:430-461 */
int or_box_init(const or_box* b, double* state);

/* one timestep() body of length dt (source/mhd/solve.hpp:1837-1924, global
   time step): fluxes, update of normal cells, refit, face sync, refit */
int or_box_step(const or_box* b, int solver, double dt, double* state,
	double smax_dim[3], long zero_dim[3], long* bad_face);

/* apply_boundaries_sw_box, source/mhd/solar_wind_box.hpp:306-577 */
int or_box_boundaries(const or_box* b, double* state);

/* sync_magnetic_field, source/common_functions.hpp:184-269 */
void or_box_sync(const or_box* b, double* state);

/* final-init step then main loop of tests/mhd/solar_wind_box.cpp:464-560
   (state after or_box_init) */
int or_box_run(const or_box* b, int solver, double cfl, double t_start, double t_end,
	long max_steps, double* state, long* steps_taken, double* t_final);

/* ---- uniform grids with value and copy boundaries ----

A grid of nx*ny*nz cells, periodic or closed in each dimension, whose state
keeps each cell's own copies of its six faces (14 doubles per cell,
component-major: rho, m (3), E, Vol_B (3), face B of the -x, +x, -y, +y,
-z, +z faces).  Neighbours are those of a neighbourhood of length 1 in the
order z slowest, x fastest (offsets -1..1); periodic dimensions wrap (a
one-cell periodic dimension makes a cell its own neighbour), closed ones
end.

or_bnd_classify restates Boundaries::classify (source/boundaries/
boundaries.hpp:100-226) for each of the four variables (number density,
velocity, pressure, magnetic field), set_solver_info (source/mhd/
boundaries.hpp:67-245) and classify_faces (:1039-1217).  Inputs per
variable v: in_value[v*n + c] / in_copy[v*n + c] = 1 where cell c overlaps
a geometry of one of v's value / copy boundaries.  Outputs: vtype[v*n + c]
(1 normal, 2 not solved, 3 value boundary, 4 copy boundary with at least
one normal neighbour), ctype[c] (1 / 0 / -1) and finfo[f*n + c] for the six
faces f = -x, +x, -y, +y, -z, +z (1 / 0 / -1).  Returns OR_ERR_SOLVER when
the variables' not-solved cells differ (the reference throws).

or_bnd_apply restates apply_boundaries (source/mhd/boundaries.hpp:
1261-1297): the value and copy boundaries of number density, velocity and
pressure (apply_fluid_boundaries, :738-1026), of the face field
(apply_magnetic_field_boundaries_staggered, :250-640, uniform grids), and
the constant-pressure refit of every cell of type >= 0.  Value boundaries
come resolved per cell (the last boundary covering a cell wins):
val_mask[v*n + c], n_val (n), v_val (3n), p_val (n), b_val (6n, the field
component normal to each face at its centre).  or_bnd_sync is
sync_magnetic_field (source/common_functions.hpp:184-269).

or_bnd_step is timestep() (source/mhd/solve.hpp:1837-1924) with per-cell
substepping over the cell types: cells of type >= 0 get periods, faces
between two of them fluxes, normal cells the update; every cell of type
>= 0 refits at its period; the face copies are synced at the end. */

typedef struct or_bgrid {
	int nx, ny, nz;
	double dx, dy, dz;
	double gamma, mu0, proton_mass;
	int periodic[3];
	const signed char* ctype;   /* n */
	const signed char* finfo;   /* 6 n */
	const signed char* vtype;   /* 4 n */
	const double* bg;           /* 9 n or NULL: Bg_B(cell)(+d), as or_grid */
} or_bgrid;

int or_bnd_classify(const or_bgrid* g, const unsigned char* in_value,
	const unsigned char* in_copy, signed char* vtype, signed char* ctype, signed char* finfo);
int or_bnd_apply(const or_bgrid* g, const unsigned char* val_mask, const double* n_val,
	const double* v_val, const double* p_val, const double* b_val, double* state);
void or_bnd_sync(const or_bgrid* g, double* state);
int or_bnd_step(const or_bgrid* g, int solver, double time_step_factor, double max_dt,
	const int* substep_min, const int* substep_max, double* state, double* maxv,
	double* dt_out, int* max_period_out, long* bad_face);

/* OpenMP threads of the grid step (n > 0 sets it); returns the count in use */
int or_set_threads(int n);
/* faces solved by the last or_step_periodic, or_step_substep, or_bnd_step
   or or_box_step: timestep()'s flux_calcs (source/mhd/solve.hpp:1918) */
unsigned long long or_last_flux_calcs(void);

/* ---- initial conditions used by the goldens and the benchmark configs ---- */

/* 7-point Gauss-Legendre average over [a, b] of f (textbook nodes/weights) */
double or_gl7_average(double (*f)(double, void*), void* ctx, double a, double b);

#ifdef __cplusplus
}
#endif

#endif
