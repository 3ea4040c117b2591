/*
Solar-wind boxes on the device: uniform non-periodic grids with a one-cell
boundary layer (inflow wall, outflow copies), planet cells and a dipole
background field, advanced with the global time step.  A BoxSet holds any
number of boxes of different sizes in concatenated arrays, so every stage
of a step is one launch for all of them (an ensemble); a single box is a
set of one.
*/
#ifndef PAMHD_SWBOX_HPP
#define PAMHD_SWBOX_HPP

#include <hip/hip_runtime.h>

#include <string>

#include "pamhd_mi355x.h"

namespace pamhd_dev {

struct BoxSet;

/* boxes[i] describes box i; all share the physics */
int boxset_create(BoxSet** out, int nbox, const pamhd_grid_desc* grids,
	const pamhd_physics* phys, const pamhd_solar_wind_desc* sw, std::string& err);
void boxset_destroy(BoxSet* s);

int boxset_count(const BoxSet* s);
long long boxset_cells(const BoxSet* s, int box);
/* faces a step of box `box` solves (both cells of type >= 0): timestep()'s
   flux_calcs, source/mhd/solve.hpp:253-313 */
long long boxset_faces(const BoxSet* s, int box);

/* tile shape of the cell kernels: 0 = 64x4, 1 = 32x8, 2 = 16x16 */
int boxset_set_tile(BoxSet* s, int tile);
int boxset_tile(const BoxSet* s);

/*
Launches one timestep() of every box (lagged CFL step from the previous flux
pass, bounded by max_dt): fluxes, update, refit, face sync, refit.  with_
boundaries adds apply_boundaries_sw_box and the face sync of the main loop.
No synchronisation.
*/
int boxset_launch_step(BoxSet* s, int solver, double max_dt, double cfl, bool with_boundaries,
	hipStream_t st, std::string& err);

/*
`steps` main-loop steps of every box (boxset_launch_step with boundaries
and max_dt unbounded, each step's last face sync done in the next step's
update, the last step's in its own pass); results as from that many
boxset_launch_step calls.  No synchronisation.
*/
int boxset_launch_steps(BoxSet* s, int solver, double cfl, int steps, hipStream_t st, std::string& err);

/*
`steps` main-loop steps of every box (boxset_launch_step with boundaries,
max_dt unbounded) in one persistent cooperative launch; the stages are
separated by grid-wide barriers whose waits are bounded (a timeout returns
PAMHD_E_DEVICE).  Synchronous.
*/
int boxset_run_persistent(BoxSet* s, int solver, double cfl, long steps, hipStream_t st,
	std::string& err);

/* polls of a persistent run's barrier wait before it gives up (0: the
   default, 2^24, about a second) */
void boxset_set_barrier_polls(BoxSet* s, unsigned polls);

/* apply_boundaries_sw_box + sync_magnetic_field only */
int boxset_launch_boundaries(BoxSet* s, hipStream_t st);

/* synchronises; per-box time step of the last step (may be NULL) and the
   first device-side physics error */
int boxset_readback(BoxSet* s, hipStream_t st, double* dt_per_box, std::string& err);

hipError_t boxset_download(BoxSet* s, int box, double* state);
hipError_t boxset_upload(BoxSet* s, int box, const double* state);
hipError_t boxset_cell_types(BoxSet* s, int box, signed char* ctype, unsigned char* rule);
hipError_t boxset_background(BoxSet* s, int box, double* bg);

/*
Times `steps` full steps with each tile shape (after one warm-up step) on a
copy of the current state, restores the state and the lagged-step status,
keeps the fastest shape and appends a report.  Synchronous.
*/
int boxset_autotune(BoxSet* s, int solver, double cfl, int steps, hipStream_t st,
	std::string& report, std::string& err);

/* mean |div B| over the normal cells of one box (synchronous) */
hipError_t boxset_div_face2volume(BoxSet* s, int box, double* avg);

} // namespace pamhd_dev

#endif
