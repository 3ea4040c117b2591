/*
The 1-D plain method of tests/mhd/reference.cpp on MI355X (gfx950): a line
of cells along x with cell-centred fields, the reference's correctness
program for the flux functions.  One step (solve_mhd, reference.cpp:284-440,
then apply_fluxes, source/mhd/common.hpp:707-743) is two launches:

  plain_face_kernel  one thread per face between cells f and f + 1: the
                     solver's flux times dt (mass, momentum, energy), the
                     upwind electric field on the face (the cell's +x edge,
                     reference.cpp:393-404) and min(cell size / max signal
                     speed) by an ordered-bits atomicMin
  plain_cell_kernel  one thread per cell: the flux sums of its two faces in
                     the reference's accumulation order, U += sum / volume,
                     then the density and pressure checks

The two end cells receive no fluid flux (the reference's loop skips them)
and the last cell no field change, exactly as reference.cpp does.  The host
runs the main loop (reference.cpp:830-877): lagged dt, the first step of
length 0, until t >= t_end.  Data in HBM, component-major: S 8 x n
{rho, m (3), E, B (3)}, face terms 7 x (n - 1).
*/
#include <hip/hip_runtime.h>
#include <cfloat>
#include <cstring>
#include <mutex>
#include <string>
#include "riemann.hpp"
#include "plain1d.hpp"

namespace pamhd_dev {
namespace {

struct PlainStatus {
	unsigned long long err_face;  /* (face << 8) | code of the first failing face; ~0 none */
	unsigned long long err_cell;  /* (cell << 8) | code of the first failing cell; ~0 none */
	unsigned long long dt_bits;   /* min cell_size / speed (non-negative values), as bits */
};

enum { PF_RHO = 0, PF_M = 1, PF_E = 4, PF_E1 = 5, PF_E2 = 6, PF_N = 7 };

template <int SOLVER>
__global__ __launch_bounds__(256) void plain_face_kernel(long long n, const double* __restrict__ S,
	double* __restrict__ PF, PlainStatus* st, Phys ph, double dt, double cell_size)
{
	const long long f = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	const long long nf = n - 1;
	if (f >= nf) return;
	double L[8], R[8];
#pragma unroll
	for (int q = 0; q < 8; q++) {
		L[q] = S[q * n + f];
		R[q] = S[q * n + f + 1];
	}
	const double zero[3] = {0, 0, 0};
	double F[8], vel = -1;
	const int e = riemann<SOLVER, double>(L, R, zero, ph, F, vel);
	if (e) {
		atomicMin(&st->err_face, ((unsigned long long)f << 8) | (unsigned long long)e);
		return;
	}
	/* face_area * dt, reference.cpp:373-381 */
	const double fa = 1.0 * dt;
	PF[PF_RHO * nf + f] = F[RHO] * fa;
	for (int q = 0; q < 3; q++) PF[(PF_M + q) * nf + f] = F[M0 + q] * fa;
	PF[PF_E * nf + f] = F[NRJ] * fa;
	const double m = cell_size / vel;
	if (m >= 0) atomicMin(&st->dt_bits, (unsigned long long)__double_as_longlong(m));
	/* upwind -v x B on the +x edge of cell f */
	const double ic = 1 / L[RHO], in = 1 / R[RHO];
	const double vc[3] = {L[M0] * ic, L[M1] * ic, L[M2] * ic};
	const double vn[3] = {R[M0] * in, R[M1] * in, R[M2] * in};
	double e1, e2;
	if (vc[0] + vn[0] > 0) {
		e1 = vc[0] * L[B2] - vc[2] * L[B0];
		e2 = vc[1] * L[B0] - vc[0] * L[B1];
	} else {
		e1 = vn[0] * R[B2] - vn[2] * L[B0];
		e2 = vn[1] * L[B0] - vn[0] * R[B1];
	}
	PF[PF_E1 * nf + f] = e1;
	PF[PF_E2 * nf + f] = e2;
}

/* a cell reads only its two faces' terms, so it updates S in place */
__global__ __launch_bounds__(256) void plain_cell_kernel(long long n, double* __restrict__ S,
	const double* __restrict__ PF, PlainStatus* st, Phys ph, double dt, double inverse_volume)
{
	const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (c >= n) return;
	/* a face error stops the step before any cell changes */
	if (st->err_face != ~0ull) return;
	const long long nf = n - 1;
	/* fluid: += the flux of face c - 1, then -= that of face c, from 0, for
	   the cells 1 .. n - 2 (reference.cpp:376-391) */
	double fr = 0, fm[3] = {0, 0, 0}, fe = 0;
	if (c >= 1 && c <= n - 2) {
		fr += PF[PF_RHO * nf + c - 1];
		for (int q = 0; q < 3; q++) fm[q] = fm[q] + PF[(PF_M + q) * nf + c - 1];
		fe += PF[PF_E * nf + c - 1];
		fr -= PF[PF_RHO * nf + c];
		for (int q = 0; q < 3; q++) fm[q] = fm[q] - PF[(PF_M + q) * nf + c];
		fe -= PF[PF_E * nf + c];
	}
	/* field: the edge terms of the cell's +x and -x edges, times dt, for
	   the cells 0 .. n - 2 (reference.cpp:405-420) */
	double fb[3] = {0, 0, 0};
	if (c <= n - 2) {
		fb[1] = PF[PF_E2 * nf + c];
		fb[2] = -PF[PF_E1 * nf + c];
		if (c > 0) {
			fb[1] -= PF[PF_E2 * nf + c - 1];
			fb[2] += PF[PF_E1 * nf + c - 1];
		}
		for (int q = 0; q < 3; q++) fb[q] = fb[q] * dt;
	}
	double U[8];
	U[RHO] = S[c] + fr * inverse_volume;
	for (int q = 0; q < 3; q++) U[M0 + q] = S[(M0 + q) * n + c] + fm[q] * inverse_volume;
	U[NRJ] = S[NRJ * n + c] + fe * inverse_volume;
	for (int q = 0; q < 3; q++) U[B0 + q] = S[(B0 + q) * n + c] + fb[q] * inverse_volume;
#pragma unroll
	for (int q = 0; q < 8; q++) S[q * n + c] = U[q];
	int e = 0;
	if (U[RHO] <= 0) e = ERR_DENSITY;
	else if (pressure64(U, ph.gm1, ph.half_inv_mu0) <= 0) e = ERR_PRESSURE;
	if (e) atomicMin(&st->err_cell, ((unsigned long long)c << 8) | (unsigned long long)e);
}

template <int SOLVER>
void launch_plain_step(long long n, double* S, double* PF, PlainStatus* st, const Phys& ph,
	double dt, double cell_size, double inverse_volume, hipStream_t stream)
{
	const unsigned bf = (unsigned)((n - 1 + 255) / 256), bc = (unsigned)((n + 255) / 256);
	plain_face_kernel<SOLVER><<<bf, 256, 0, stream>>>(n, S, PF, st, ph, dt, cell_size);
	plain_cell_kernel<<<bc, 256, 0, stream>>>(n, S, PF, st, ph, dt, inverse_volume);
}

double smin_host(double a, double b) { return (b < a) ? b : a; }

/*
Scratch of the allocation-free entry point (plain1d_run_device): the status
record in the first 256 bytes, then the face terms, 7 x (n - 1) doubles for
at most PAMHD_PLAIN1D_SCRATCH_CELLS cells (14.7 MB).  Each GPU the library
runs on holds its own copy until the process ends; host code reaches it only
through the address hipGetSymbolAddress returns for the current device.
*/
constexpr long long SCRATCH_STATUS_DOUBLES = 32;
constexpr long long SCRATCH_DOUBLES =
	SCRATCH_STATUS_DOUBLES + (long long)PF_N * (PAMHD_PLAIN1D_SCRATCH_CELLS - 1);
static_assert(sizeof(PlainStatus) <= SCRATCH_STATUS_DOUBLES * sizeof(double), "status fits");
__device__ __align__(256) double scratch[SCRATCH_DOUBLES];

constexpr int SCRATCH_MAX_DEVICES = 64;
std::mutex scratch_mutex;
double* scratch_ptr[SCRATCH_MAX_DEVICES] = {};

/* the current device's copy of scratch, looked up once per device */
int scratch_address(double** out, std::string& err)
{
	int dev = -1;
	hipError_t he = hipGetDevice(&dev);
	if (he != hipSuccess) {
		err = std::string("hipGetDevice: ") + hipGetErrorString(he);
		return PAMHD_E_DEVICE;
	}
	if (dev < 0 || dev >= SCRATCH_MAX_DEVICES) {
		err = "device index " + std::to_string(dev) + " beyond the scratch table";
		return PAMHD_E_DEVICE;
	}
	std::lock_guard<std::mutex> lock(scratch_mutex);
	if (!scratch_ptr[dev]) {
		void* p = nullptr;
		he = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
		if (he != hipSuccess || !p) {
			err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(he);
			return PAMHD_E_DEVICE;
		}
		scratch_ptr[dev] = (double*)p;
	}
	*out = scratch_ptr[dev];
	return PAMHD_OK;
}

/* the status record at the start of a step, written by a kernel (no memset
   or host copy, so the entry point can be captured) */
__global__ void plain_status_init_kernel(PlainStatus* st)
{
	if (threadIdx.x == 0 && blockIdx.x == 0) {
		st->err_face = ~0ull;
		st->err_cell = ~0ull;
		st->dt_bits = (unsigned long long)__double_as_longlong(DBL_MAX);
	}
}

/* the main loop of reference.cpp:830-877 over device buffers: S is the
   caller's state, PF and st are work buffers, host_st host memory for the
   status read back after each step */
int plain_loop(long long n, double cell_size, const Phys& ph, double cfl, int solver, double t_end,
	long max_steps, double* S, double* PF, PlainStatus* st, PlainStatus* host_st, long* steps_taken,
	hipStream_t stream, bool init_by_kernel, std::string& err)
{
	const double inverse_volume = 1.0 / (1.0 * cell_size);
	double max_dt = 0, t = 0;
	long steps = 0;
	int code = PAMHD_OK;
	hipError_t he;
	auto hip_err = [&](hipError_t e, const char* what) {
		err = std::string(what) + ": " + hipGetErrorString(e);
		return PAMHD_E_DEVICE;
	};
	while (t < t_end && steps < max_steps) {
		const double until_end = t_end - t;
		const double time_step = smin_host(cfl * max_dt, until_end);
		if (init_by_kernel) {
			plain_status_init_kernel<<<1, 64, 0, stream>>>(st);
		} else {
			PlainStatus init;
			init.err_face = init.err_cell = ~0ull;
			const double big = DBL_MAX;
			memcpy(&init.dt_bits, &big, sizeof big);
			*host_st = init;
			if ((he = hipMemcpyAsync(st, host_st, sizeof(PlainStatus), hipMemcpyHostToDevice, stream)) != hipSuccess)
				return hip_err(he, "hipMemcpyAsync");
		}
		switch (solver) {
		case RUSANOV: launch_plain_step<RUSANOV>(n, S, PF, st, ph, time_step, cell_size, inverse_volume, stream); break;
		case HLL_ATHENA: launch_plain_step<HLL_ATHENA>(n, S, PF, st, ph, time_step, cell_size, inverse_volume, stream); break;
		case HLLD_ATHENA: launch_plain_step<HLLD_ATHENA>(n, S, PF, st, ph, time_step, cell_size, inverse_volume, stream); break;
		default: launch_plain_step<ROE_ATHENA>(n, S, PF, st, ph, time_step, cell_size, inverse_volume, stream); break;
		}
		if ((he = hipGetLastError()) != hipSuccess) return hip_err(he, "kernel launch");
		if ((he = hipMemcpyAsync(host_st, st, sizeof(PlainStatus), hipMemcpyDeviceToHost, stream)) != hipSuccess)
			return hip_err(he, "hipMemcpyAsync");
		if ((he = hipStreamSynchronize(stream)) != hipSuccess) return hip_err(he, "hipStreamSynchronize");
		const bool face = host_st->err_face != ~0ull;
		if (face || host_st->err_cell != ~0ull) {
			const unsigned long long e = face ? host_st->err_face : host_st->err_cell;
			code = (int)(e & 0xff);
			err = std::string("step ") + std::to_string(steps) + (face ? ", face " : ", cell ") +
				std::to_string(e >> 8);
			break;
		}
		double next;
		memcpy(&next, &host_st->dt_bits, sizeof next);
		max_dt = smin_host(DBL_MAX, next);
		t += time_step;
		steps++;
	}
	if (steps_taken) *steps_taken = steps;
	return code;
}

int check_args(long long n, int solver, std::string& err)
{
	if (n < 2) {
		err = "the plain method needs at least two cells";
		return PAMHD_E_ARG;
	}
	if (solver != RUSANOV && solver != HLL_ATHENA && solver != HLLD_ATHENA && solver != ROE_ATHENA) {
		err = "unknown solver";
		return PAMHD_E_SOLVER;
	}
	return PAMHD_OK;
}

} // namespace

int plain1d_run(long long n, double cell_size, double gamma, double mu0, double cfl, int solver,
	double t_end, long max_steps, double* state, long* steps_taken, hipStream_t stream,
	std::string& err)
{
	int code = check_args(n, solver, err);
	if (code) return code;
	const Phys ph = make_phys(gamma, mu0);
	double* S = nullptr;
	double* PF = nullptr;
	PlainStatus* st = nullptr;
	PlainStatus* host = nullptr;
	auto release = [&]() {
		if (S) (void)hipFree(S);
		if (PF) (void)hipFree(PF);
		if (st) (void)hipFree(st);
		if (host) (void)hipHostFree(host);
	};
	auto hip_err = [&](hipError_t e, const char* what) {
		err = std::string(what) + ": " + hipGetErrorString(e);
		release();
		return PAMHD_E_DEVICE;
	};
	hipError_t he;
	if ((he = hipMalloc(&S, sizeof(double) * 8 * n)) != hipSuccess) return hip_err(he, "hipMalloc");
	if ((he = hipMalloc(&PF, sizeof(double) * PF_N * (n - 1))) != hipSuccess) return hip_err(he, "hipMalloc");
	if ((he = hipMalloc(&st, sizeof(PlainStatus))) != hipSuccess) return hip_err(he, "hipMalloc");
	if ((he = hipHostMalloc(&host, sizeof(PlainStatus), hipHostMallocDefault)) != hipSuccess)
		return hip_err(he, "hipHostMalloc");
	if ((he = hipMemcpyAsync(S, state, sizeof(double) * 8 * n, hipMemcpyHostToDevice, stream)) != hipSuccess)
		return hip_err(he, "hipMemcpyAsync");
	code = plain_loop(n, cell_size, ph, cfl, solver, t_end, max_steps, S, PF, st, host, steps_taken,
		stream, false, err);
	if (code == PAMHD_E_DEVICE) {
		release();
		return code;
	}
	if ((he = hipMemcpyAsync(state, S, sizeof(double) * 8 * n, hipMemcpyDeviceToHost, stream)) != hipSuccess)
		return hip_err(he, "hipMemcpyAsync");
	if ((he = hipStreamSynchronize(stream)) != hipSuccess) return hip_err(he, "hipStreamSynchronize");
	release();
	return code;
}

int plain1d_run_device(long long n, double cell_size, double gamma, double mu0, double cfl,
	int solver, double t_end, long max_steps, double* state, long* steps_taken, hipStream_t stream,
	std::string& err)
{
	int code = check_args(n, solver, err);
	if (code) return code;
	if (n > PAMHD_PLAIN1D_SCRATCH_CELLS) {
		err = "the allocation-free plain method takes at most " +
			std::to_string(PAMHD_PLAIN1D_SCRATCH_CELLS) + " cells (" + std::to_string(n) + " given)";
		return PAMHD_E_ARG;
	}
	double* base = nullptr;
	code = scratch_address(&base, err);
	if (code) return code;
	PlainStatus* st = (PlainStatus*)base;
	double* PF = base + SCRATCH_STATUS_DOUBLES;
	PlainStatus host;
	return plain_loop(n, cell_size, make_phys(gamma, mu0), cfl, solver, t_end, max_steps, state, PF,
		st, &host, steps_taken, stream, true, err);
}

} // namespace pamhd_dev
