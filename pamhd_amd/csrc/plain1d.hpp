/* The 1-D plain method of tests/mhd/reference.cpp (plain1d.hip). */
#pragma once
#include <hip/hip_runtime.h>
#include <string>
#include "pamhd_mi355x.h"

namespace pamhd_dev {

/* state: host, 8 * n doubles {rho, m (3), E, B (3)}, read and overwritten;
   returns a PAMHD_* code with a message in err */
int plain1d_run(long long n, double cell_size, double gamma, double mu0, double cfl, int solver,
	double t_end, long max_steps, double* state, long* steps_taken, hipStream_t stream,
	std::string& err);

/* the same on a device state (8 * n doubles, in place) with the library's
   fixed scratch array: no allocation; n <= PAMHD_PLAIN1D_SCRATCH_CELLS */
int plain1d_run_device(long long n, double cell_size, double gamma, double mu0, double cfl,
	int solver, double t_end, long max_steps, double* state, long* steps_taken, hipStream_t stream,
	std::string& err);

} // namespace pamhd_dev
