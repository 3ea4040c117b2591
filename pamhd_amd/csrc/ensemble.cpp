/*
Ensemble of solar-wind boxes behind the C ABI: every box of the set is
advanced by the same launches (one per stage of the step), and a run of
main-loop steps replays a HIP graph holding a chain of 16 (or 2)
consecutive steps (the lagged-step status alternates between two buffers,
so an even number of steps brings the launch arguments back to where they
started; inside a chain each step's last face sync is done by the next
step's update), or runs every step in one persistent cooperative launch.
*/
#include <hip/hip_runtime.h>

#include <cstdio>
#include <string>

#include "pamhd_mi355x.h"
#include "swbox.hpp"

using namespace pamhd_dev;

struct pamhd_ensemble {
	BoxSet* set = nullptr;
	hipStream_t own = nullptr;
	std::string err;
};

namespace {

int efail(pamhd_ensemble* e, int code, const std::string& m)
{
	if (e) e->err = m;
	return code;
}

std::string& create_error()
{
	static thread_local std::string m;
	return m;
}

hipStream_t pick(pamhd_ensemble* e, void* stream)
{
	return stream ? (hipStream_t)stream : e->own;
}

} // namespace

extern "C" {

int pamhd_ensemble_create(int nbox, const pamhd_grid_desc* grids, const pamhd_physics* physics,
	const pamhd_solar_wind_desc* sw, pamhd_ensemble** out)
{
	if (!out) return PAMHD_E_ARG;
	*out = nullptr;
	pamhd_ensemble* e = new pamhd_ensemble();
	std::string err;
	const int rc = boxset_create(&e->set, nbox, grids, physics, sw, err);
	if (rc) {
		delete e;
		create_error() = err;
		return rc;
	}
	if (hipStreamCreate(&e->own) != hipSuccess) {
		boxset_destroy(e->set);
		delete e;
		return PAMHD_E_DEVICE;
	}
	*out = e;
	return PAMHD_OK;
}

int pamhd_ensemble_destroy(pamhd_ensemble* e)
{
	if (!e) return PAMHD_OK;
	if (e->own) {
		(void)hipStreamSynchronize(e->own);
		(void)hipStreamDestroy(e->own);
	}
	boxset_destroy(e->set);
	delete e;
	return PAMHD_OK;
}

int pamhd_ensemble_timestep(pamhd_ensemble* e, int solver, double max_time_step,
	double time_step_factor, void* stream)
{
	if (!e) return PAMHD_E_ARG;
	const int rc = boxset_launch_step(e->set, solver, max_time_step, time_step_factor, false,
		pick(e, stream), e->err);
	return rc ? efail(e, rc, e->err) : PAMHD_OK;
}

int pamhd_ensemble_run(pamhd_ensemble* e, int solver, double time_step_factor, long steps,
	int use_graph, void* stream)
{
	if (!e) return PAMHD_E_ARG;
	if (steps < 0) return efail(e, PAMHD_E_ARG, "steps must be >= 0");
	hipStream_t st = pick(e, stream);
	int rc = PAMHD_OK;
	if (use_graph == 2) {
		/* one persistent launch for all the steps */
		rc = boxset_run_persistent(e->set, solver, time_step_factor, steps, st, e->err);
		return rc ? efail(e, rc, e->err) : PAMHD_OK;
	}
	/* chains of at most 1024 steps (the launch count of one call is an int) */
	if (!use_graph || steps < 4) {
		for (long done = 0; done < steps && rc == PAMHD_OK; done += 1024) {
			const int n = (int)(steps - done < 1024 ? steps - done : 1024);
			rc = boxset_launch_steps(e->set, solver, time_step_factor, n, st, e->err);
		}
		return rc ? efail(e, rc, e->err) : PAMHD_OK;
	}
	const int chain = steps >= 32 ? 16 : 2;
	if (steps % chain) rc = boxset_launch_steps(e->set, solver, time_step_factor, (int)(steps % chain), st, e->err);
	if (rc) return efail(e, rc, e->err);
	/* capture one chain (st is never the default stream here) */
	hipStream_t cap = st;
	hipGraph_t graph = nullptr;
	hipGraphExec_t exec = nullptr;
	hipError_t he = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
	if (he != hipSuccess) return efail(e, PAMHD_E_DEVICE, std::string("begin capture: ") + hipGetErrorString(he));
	const int r2 = boxset_launch_steps(e->set, solver, time_step_factor, chain, cap, e->err);
	he = hipStreamEndCapture(cap, &graph);
	if (r2 || he != hipSuccess) {
		if (graph) (void)hipGraphDestroy(graph);
		return efail(e, r2 ? r2 : PAMHD_E_DEVICE, r2 ? e->err : std::string("end capture: ") + hipGetErrorString(he));
	}
	he = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
	if (he != hipSuccess) {
		(void)hipGraphDestroy(graph);
		return efail(e, PAMHD_E_DEVICE, std::string("instantiate: ") + hipGetErrorString(he));
	}
	const long reps = steps / chain;
	for (long r = 0; r < reps && he == hipSuccess; r++) he = hipGraphLaunch(exec, cap);
	/* the executable graph may be destroyed once its launches complete */
	(void)hipStreamSynchronize(cap);
	(void)hipGraphExecDestroy(exec);
	(void)hipGraphDestroy(graph);
	if (he != hipSuccess) return efail(e, PAMHD_E_DEVICE, std::string("graph launch: ") + hipGetErrorString(he));
	return PAMHD_OK;
}

int pamhd_ensemble_set_barrier_polls(pamhd_ensemble* e, long polls)
{
	if (!e) return PAMHD_E_ARG;
	if (polls < 0 || polls > 0xffffffffL) return efail(e, PAMHD_E_ARG, "polls must be in 0..2^32-1");
	boxset_set_barrier_polls(e->set, (unsigned)polls);
	return PAMHD_OK;
}

int pamhd_ensemble_check(pamhd_ensemble* e, double* dt_per_box, void* stream)
{
	if (!e) return PAMHD_E_ARG;
	const int rc = boxset_readback(e->set, pick(e, stream), dt_per_box, e->err);
	return rc ? efail(e, rc, e->err) : PAMHD_OK;
}

int pamhd_ensemble_download(pamhd_ensemble* e, int box, double* state)
{
	if (!e || !state || box < 0 || box >= boxset_count(e->set)) return efail(e, PAMHD_E_ARG, "invalid box or null state");
	(void)hipStreamSynchronize(e->own);
	const hipError_t he = boxset_download(e->set, box, state);
	return he == hipSuccess ? PAMHD_OK : efail(e, PAMHD_E_DEVICE, hipGetErrorString(he));
}

int pamhd_ensemble_cells(pamhd_ensemble* e, int box, long long* cells)
{
	if (!e || !cells || box < 0 || box >= boxset_count(e->set)) return efail(e, PAMHD_E_ARG, "invalid box");
	*cells = boxset_cells(e->set, box);
	return PAMHD_OK;
}

int pamhd_ensemble_set_tile(pamhd_ensemble* e, int tile)
{
	if (!e) return PAMHD_E_ARG;
	const int rc = boxset_set_tile(e->set, tile);
	return rc ? efail(e, rc, "invalid tile " + std::to_string(tile)) : PAMHD_OK;
}

int pamhd_ensemble_autotune(pamhd_ensemble* e, int solver, double time_step_factor, char* buf,
	size_t buflen, void* stream)
{
	if (!e) return PAMHD_E_ARG;
	std::string report;
	const int rc = boxset_autotune(e->set, solver, time_step_factor, 3, pick(e, stream), report, e->err);
	if (buf && buflen) snprintf(buf, buflen, "%s", report.c_str());
	return rc ? efail(e, rc, e->err) : PAMHD_OK;
}

int pamhd_ensemble_last_error(pamhd_ensemble* e, char* buf, size_t n)
{
	if (!buf || n == 0) return PAMHD_E_ARG;
	snprintf(buf, n, "%s", e ? e->err.c_str() : create_error().c_str());
	return PAMHD_OK;
}

} // extern "C"
