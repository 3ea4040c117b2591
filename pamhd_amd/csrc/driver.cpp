/*
Host driver of the MHD program loop, tests/mhd/test.cpp:446-610, for a
periodic grid resident on the device:

    apply_boundaries (energy refit)                    test.cpp:446-453
    timestep(max_dt = 0)  "final init", fills speeds   test.cpp:462-470
    while t < t_end:
        dt = timestep(max_dt = t_end - t)              test.cpp:522-537
        apply_boundaries, sync_magnetic_field           test.cpp:559-568

timestep + apply_boundaries run as one fused pass (pamhd_step); the face
field is stored once per face, so sync_magnetic_field has nothing to average.
*/
#include <cstdio>

#include "pamhd_mi355x.h"

extern "C" int pamhd_run(pamhd_handle* h, int solver, double time_step_factor,
	double t_start, double t_end, long max_steps, long* steps_taken,
	double* t_final, void* stream)
{
	if (!h) return PAMHD_E_ARG;
	int kind = PAMHD_GRID_PERIODIC;
	int e = pamhd_grid_kind(h, &kind);
	if (e) return e;
	/* a solar-wind box is created with its boundaries applied
	   (tests/mhd/solar_wind_box.cpp:448-461); the others get them here
	   (a grid with boundaries: the values last given to
	   pamhd_set_boundary_values) */
	if (kind != PAMHD_GRID_SOLAR_WIND_BOX) {
		e = pamhd_apply_boundaries(h, t_start, stream);
		if (e) return e;
	}
	double dt = 0;
	e = pamhd_timestep(h, solver, 0.0, time_step_factor, &dt, stream);
	if (e) return e;
	double t = t_start;
	long steps = 0;
	while (t < t_end && steps < max_steps) {
		const double until_end = t_end - t;
		e = pamhd_step(h, solver, until_end, time_step_factor, &dt, stream);
		if (e) return e;
		t += dt;
		steps++;
	}
	if (steps_taken) *steps_taken = steps;
	if (t_final) *t_final = t;
	return PAMHD_OK;
}
