/*
MI355X (gfx950) implementation of the PAMHD ideal-MHD finite-volume step
behind the C ABI of include/pamhd_mi355x.h.

Data layout in HBM (structure of arrays, FP64, component-major):
  state[2]  ping-pong buffers of 8 * n doubles:
            rho, m_x, m_y, m_z, E, B_x(i-1/2), B_y(j-1/2), B_z(k-1/2)
  flux      24 * n values: dt * F_q of the d-face owned by each cell
            (d-major, then component, then cell), written by the flux pass
            and read by the update pass (FP64, or FP32 in mixed precision)
  bg        optional 9 * n doubles of background field on the positive faces
  status[2] per-pass reductions (max signal speed per dimension, count of
            faces with speed <= 0, first physics error) feeding the lagged
            CFL time step of the next step

One step = dt_kernel (1 thread: lagged CFL step from the previous pass) ->
flux_pass_kernel (one thread per face, all faces of one dimension in a
z-slab of blocks) -> update_cell_kernel or update_column_kernel (one thread
per cell, or column tiles marching in z: constrained-transport face update,
conserved update, constant-pressure energy refits).  In mixed precision the flux array holds FP32 values.
*/
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "pamhd_mi355x.h"
#include "poisson.hpp"
#include "riemann.hpp"
#include "plain1d.hpp"
#include "swbox.hpp"

namespace pamhd_dev {
size_t adjoint_scratch_doubles(int nx, int ny, int nz);
void adjoint_launch(int nx, int ny, int nz, const double len[3], double dt, const Phys& ph,
	const double* state, double* lam, double* scratch, hipStream_t st, int reverse,
	const double* pe_given);
void adjoint_steps_launch(int nx, int ny, int nz, const double len[3], int nsteps, const double* dts,
	const double* const* states, const double* const* pes, const Phys& ph, double* lam,
	double* scratch, hipStream_t st, int reverse);
}

using namespace pamhd_dev;

namespace {

struct GridDev {
	int nx, ny, nz;
	long long n;
	double dx, dy, dz;
	Phys ph;
};

/* the flux pass spreads its per-block maxima over NSLOT addresses per
   dimension (slot = block id mod NSLOT), so 10^5 blocks do not all update
   one word; the dt kernel reduces the slots */
constexpr int NSLOT = 64;

struct StepStatus {
	unsigned long long smax_bits[3][NSLOT];
	unsigned int zero[3];
	int err;
	long long err_face;
	unsigned long long flux_calcs;   /* faces solved (substepping path) */
};

constexpr int BX = 64, BY = 4;      /* flux pass and cell kernels */
/* column update tiles (TX x TY threads): 64 x 4, 128 x 2, 32 x 8 */
constexpr int NXCD = 8;             /* MI355X: 8 XCDs, blocks dispatched round-robin */

__device__ __forceinline__ int wrap_dec(int i, int n) { return i == 0 ? n - 1 : i - 1; }
__device__ __forceinline__ int wrap_inc(int i, int n) { return i + 1 == n ? 0 : i + 1; }

/*
XCD-aware block order: the hardware hands consecutive block ids to the 8 XCDs
in turn, each with its own L2.  Remapping id -> (id % 8) * (nblocks / 8) +
id / 8 gives every XCD one contiguous range of logical blocks (x fastest,
then y, then z), so the neighbouring planes a block reads were touched by the
same XCD shortly before.
*/
__device__ __forceinline__ unsigned logical_block(unsigned id, unsigned nblocks)
{
	if (nblocks % NXCD) return id;
	return (id % NXCD) * (nblocks / NXCD) + id / NXCD;
}

/*
Tile (bx, by) and plane k of logical block lb of a launch over gx * gy
column tiles and gridDim / (gx * gy) planes from k0.  Bit 1 of `order`
walks the planes fastest (consecutive blocks take the same column tile in
consecutive planes, so the neighbouring planes a block reads were just
fetched into the same XCD's L2 when bit 0, the XCD-aware order, is set),
otherwise x tiles are fastest.
*/
__device__ __forceinline__ void tile_of_block(unsigned lb, unsigned gx, unsigned gy, unsigned nblocks,
	int order, int k0, unsigned& bx, unsigned& by, unsigned& k)
{
	if (order & 2) {
		const unsigned nk = nblocks / (gx * gy);
		const unsigned r = lb / nk;
		k = k0 + lb % nk;
		bx = r % gx;
		by = r / gx;
	} else {
		bx = lb % gx;
		by = (lb / gx) % gy;
		k = k0 + lb / (gx * gy);
	}
}

__device__ __forceinline__ void record_error(StepStatus* st, int code, long long face)
{
	if (atomicCAS(&st->err, 0, code) == 0) {
		st->err_face = face;
	}
}

/* block-wide max of s and count of s <= 0, one atomic per block */
__device__ __forceinline__ void reduce_speed(StepStatus* st, int d, double s, unsigned zero)
{
	__shared__ double ms[BX * BY / 64];
	__shared__ unsigned mz[BX * BY / 64];
	for (int off = 32; off > 0; off >>= 1) {
		const double o = __shfl_xor(s, off, 64);
		s = smax(s, o);
		zero += __shfl_xor(zero, off, 64);
	}
	const int tid = threadIdx.x + blockDim.x * threadIdx.y;
	const int wave = tid / 64, lane = tid % 64;
	if (lane == 0) {
		ms[wave] = s;
		mz[wave] = zero;
	}
	__syncthreads();
	if (tid == 0) {
		double m = ms[0];
		unsigned z = mz[0];
		for (int w = 1; w < (BX * BY) / 64; w++) {
			m = smax(m, ms[w]);
			z += mz[w];
		}
		const unsigned long long bits = m > 0 ? (unsigned long long)__double_as_longlong(m) : 0ull;
		atomicMax(&st->smax_bits[d][blockIdx.x % NSLOT], bits);
		if (z) atomicAdd(&st->zero[d], z);
	}
}

/*
Flux pass: get_fluxes + get_flux of source/mhd/solve.hpp:68-188, :304-393.
One launch per face direction D (compile-time, so the rotation of
get_rotated_vector is register renaming); one thread per face: the thread's
cell (i, j, k) owns the D-face on its negative side, between L = cell - e_D
and R = cell.  Stores dt * F (global frame) as PT.  Offsets are 32-bit
(n < 2^28 cells) from per-array base pointers.
*/
struct StateArrays {
	const double* q[8];
};

template <int D>
__device__ __forceinline__ void load_face_state_d(const StateArrays& S, const GridDev& g,
	int i, int j, int k, double U[8])
{
	const int c = i + g.nx * (j + g.ny * k);
	const int cx = wrap_inc(i, g.nx) + g.nx * (j + g.ny * k);
	const int cy = i + g.nx * (wrap_inc(j, g.ny) + g.ny * k);
	const int cz = i + g.nx * (j + g.ny * wrap_inc(k, g.nz));
	const double m[3] = {S.q[1][c], S.q[2][c], S.q[3][c]};
	const double B[3] = {
		0.5 * (S.q[5][c] + S.q[5][cx]),
		0.5 * (S.q[6][c] + S.q[6][cy]),
		0.5 * (S.q[7][c] + S.q[7][cz])};
	U[RHO] = S.q[0][c];
	U[NRJ] = S.q[4][c];
	U[M0] = m[D];
	U[M1] = m[(D + 1) % 3];
	U[M2] = m[(D + 2) % 3];
	U[B0] = B[D];
	U[B1] = B[(D + 1) % 3];
	U[B2] = B[(D + 2) % 3];
}

/* MINB > 1 asks the compiler for MINB blocks per CU (at most 128 VGPRs per
   lane for 256-thread blocks with MINB = 4), an autotuned occupancy variant */
template <int SOLVER, class T, class PT, int D, int MINB>
__global__ __launch_bounds__(BX * BY, MINB) void flux_pass_kernel(GridDev g, StateArrays S,
	const double* __restrict__ bg, PT* __restrict__ P, const double* __restrict__ dtp,
	StepStatus* st, unsigned gx, unsigned gy, int remap, int k0)
{
	const unsigned lb = (remap & 1) ? logical_block(blockIdx.x, gridDim.x) : blockIdx.x;
	unsigned bx, by, k;
	tile_of_block(lb, gx, gy, gridDim.x, remap, k0, bx, by, k);
	const int i = bx * BX + threadIdx.x;
	const int j = by * BY + threadIdx.y;
	const int n = (int)g.n;
	double sp = 0;
	unsigned zero = 0;
	if (i < g.nx && j < g.ny) {
		const int c = i + g.nx * (j + g.ny * (int)k);
		double L[8], R[8];
		load_face_state_d<D>(S, g, D == 0 ? wrap_dec(i, g.nx) : i, D == 1 ? wrap_dec(j, g.ny) : j,
			D == 2 ? wrap_dec((int)k, g.nz) : (int)k, L);
		load_face_state_d<D>(S, g, i, j, (int)k, R);
		T b[3] = {0, 0, 0};
		if (bg) {
			/* Bg_B(L)(+D): the face is the positive D-face of L */
			const int cl = D == 0 ? wrap_dec(i, g.nx) + g.nx * (j + g.ny * (int)k)
				: D == 1 ? i + g.nx * (wrap_dec(j, g.ny) + g.ny * (int)k)
				: i + g.nx * (j + g.ny * wrap_dec((int)k, g.nz));
			const double* bgd = bg + (size_t)(3 * D) * g.n;
			b[0] = (T)bgd[(size_t)D * g.n + cl];
			b[1] = (T)bgd[(size_t)((D + 1) % 3) * g.n + cl];
			b[2] = (T)bgd[(size_t)((D + 2) % 3) * g.n + cl];
		}
		/* dt * F rotated back to the global frame, in the flux precision
		   (FP64: dt * F of the reference's flux slots; the mixed HLLD's
		   FP64 fluxes are scaled in FP64) */
		using FT = std::conditional_t<SOLVER == HLLD_ATHENA, double, T>;
		FT F[8], Fg[8], sT = 0;
		int e;
		if constexpr (SOLVER == HLLD_ATHENA && std::is_same_v<T, float>) {
			double sd = 0;
			e = riemann_d<SOLVER, T>(L, R, b, g.ph, F, sd);
			sT = sd;
		} else {
			e = riemann<SOLVER, T>(L, R, b, g.ph, F, sT);
		}
		if (e) {
			record_error(st, e, (long long)D * g.n + c);
			for (int q = 0; q < 8; q++) Fg[q] = 0;
			sT = 0;
		} else {
			const FT dt = (FT)*dtp;
			Fg[RHO] = dt * F[RHO];
			Fg[NRJ] = dt * F[NRJ];
			Fg[M0 + D] = dt * F[M0];
			Fg[M0 + (D + 1) % 3] = dt * F[M1];
			Fg[M0 + (D + 2) % 3] = dt * F[M2];
			Fg[B0 + D] = dt * F[B0];
			Fg[B0 + (D + 1) % 3] = dt * F[B1];
			Fg[B0 + (D + 2) % 3] = dt * F[B2];
		}
		PT* Pd = P + (size_t)D * 8 * g.n;
#pragma unroll
		for (int q = 0; q < 8; q++) (Pd + (size_t)q * g.n)[c] = (PT)Fg[q];
		sp = (double)sT;
		zero = sp <= 0 ? 1u : 0u;
		(void)n;
	}
	reduce_speed(st, D, sp, zero);
}

/*
HLLD flux pass with the per-cell part shared through LDS.  A block's
BX x BY threads own a tile of cells: BX along x by BY along y in plane k
for D = 0 and 1, BX along x by BY along z in row j for D = 2.  Each thread
computes hlld_side of its cell in the D-frame once (thermal pressure,
1 / rho, fast speed, check stage) and stages it in LDS; the cells just
outside the tile's -D edge are done by the first wave in a second pass.
The thread's face (between the cell at -e_D and its own) then takes the
left side from LDS, reloads the left state (a neighbour loaded it, so an
L1/L2 hit) and runs HLLD's per-face part.  Every value equals
flux_pass_kernel's: the same functions of the same inputs, the same first
error per face (first_side_error).  k_end bounds the planes of a z-slab.
*/
template <class T, class PT, int D, int MINB>
__global__ __launch_bounds__(BX * BY, MINB) void flux_hlld_shared_kernel(GridDev g, StateArrays S,
	PT* __restrict__ P, const double* __restrict__ dtp, StepStatus* st, unsigned gx, unsigned gy,
	int remap, int k0, int k_end)
{
	constexpr int C = D == 0 ? BX + 1 : BX;       /* LDS row length */
	constexpr int NS = D == 0 ? BY * (BX + 1) : (BY + 1) * BX;
	__shared__ T lp[NS], lir[NS], lcf[NS];
	__shared__ double lpt[NS];                     /* FP64 total pressure (FP32 sides) */
	__shared__ int lsc[NS];                        /* stage | code << 4 */
	const unsigned lb = (remap & 1) ? logical_block(blockIdx.x, gridDim.x) : blockIdx.x;
	const int tx = threadIdx.x, ty = threadIdx.y;
	int x0, y0, z0;                                /* tile origin */
	if (D < 2) {
		unsigned bx, by, kk;
		tile_of_block(lb, gx, gy, gridDim.x, remap & 1, k0, bx, by, kk);
		x0 = bx * BX;
		y0 = by * BY;
		z0 = (int)kk;
	} else {
		const unsigned bx = lb % gx, r = lb / gx;
		x0 = bx * BX;
		y0 = (int)(r % (unsigned)g.ny);
		z0 = k0 + (int)(r / (unsigned)g.ny) * BY;
	}
	const int i = x0 + tx;
	const int j = D < 2 ? y0 + ty : y0;
	const int k = D < 2 ? z0 : z0 + ty;
	const bool active = i < g.nx && j < g.ny && k < k_end;
	const int own = D == 0 ? ty * C + tx + 1 : (ty + 1) * C + tx;
	const int left = ty * C + tx;

	double U[8];
	Side<T> sd;
	T cf = 0;
	int code = 0, stage = 0;
	if (active) {
		load_face_state_d<D>(S, g, i, j, k, U);
		stage = hlld_side<T>(U, g.ph, sd, cf, code);
		lp[own] = sd.p;
		lir[own] = sd.ir;
		lcf[own] = cf;
		lpt[own] = sd.pt64;
		lsc[own] = stage | code << 4;
	}
	/* the cells on the -D side of the tile */
	if (ty == 0) {
		int hi, hj, hk, slot;
		bool h_active;
		if (D == 0) {
			hi = wrap_dec(x0, g.nx);
			hj = y0 + tx;
			hk = z0;
			slot = tx * C;
			h_active = tx < BY && hj < g.ny;
		} else if (D == 1) {
			hi = x0 + tx;
			hj = wrap_dec(y0, g.ny);
			hk = z0;
			slot = tx;
			h_active = hi < g.nx;
		} else {
			hi = x0 + tx;
			hj = y0;
			hk = wrap_dec(z0, g.nz);
			slot = tx;
			h_active = hi < g.nx && z0 < k_end;
		}
		if (h_active) {
			double H[8];
			Side<T> hs;
			T hcf = 0;
			int hcode;
			load_face_state_d<D>(S, g, hi, hj, hk, H);
			const int hst = hlld_side<T>(H, g.ph, hs, hcf, hcode);
			lp[slot] = hs.p;
			lir[slot] = hs.ir;
			lcf[slot] = hcf;
			lpt[slot] = hs.pt64;
			lsc[slot] = hst | hcode << 4;
		}
	}
	__syncthreads();

	double sp = 0;
	unsigned zero = 0;
	if (active) {
		const int c = i + g.nx * (j + g.ny * k);
		const int sc = lsc[left];
		const int stage_l = sc & 15, code_l = sc >> 4;
		T F[8], sT = 0;
		int e;
		if (stage_l | stage) {
			e = first_side_error(stage_l, code_l, stage, code);
		} else {
			double Ld[8];
			load_face_state_d<D>(S, g, D == 0 ? wrap_dec(i, g.nx) : i, D == 1 ? wrap_dec(j, g.ny) : j,
				D == 2 ? wrap_dec(k, g.nz) : k, Ld);
			Side<T> Ls;
#pragma unroll
			for (int q = 0; q < 8; q++) Ls.u[q] = (T)Ld[q];
			Ls.p = lp[left];
			Ls.ir = lir[left];
			Ls.v[0] = Ls.u[M0] * Ls.ir;
			Ls.v[1] = Ls.u[M1] * Ls.ir;
			Ls.v[2] = Ls.u[M2] * Ls.ir;
			Ls.p64 = 0;
			Ls.pt64 = lpt[left];
			Ls.rho_pos = true;
			const T cfl = lcf[left];
			e = flux_hlld_face<T>(Ls, sd, cfl, cf, Ld, U, g.ph, F, sT);
		}
		T Fg[8];
		if (e) {
			record_error(st, e, (long long)D * g.n + c);
			for (int q = 0; q < 8; q++) Fg[q] = 0;
			sT = 0;
		} else {
			const T dt = (T)*dtp;
			Fg[RHO] = dt * F[RHO];
			Fg[NRJ] = dt * F[NRJ];
			Fg[M0 + D] = dt * F[M0];
			Fg[M0 + (D + 1) % 3] = dt * F[M1];
			Fg[M0 + (D + 2) % 3] = dt * F[M2];
			Fg[B0 + D] = dt * F[B0];
			Fg[B0 + (D + 1) % 3] = dt * F[B1];
			Fg[B0 + (D + 2) % 3] = dt * F[B2];
		}
		PT* Pd = P + (size_t)D * 8 * g.n;
#pragma unroll
		for (int q = 0; q < 8; q++) (Pd + (size_t)q * g.n)[c] = (PT)Fg[q];
		sp = (double)sT;
		zero = sp <= 0 ? 1u : 0u;
	}
	reduce_speed(st, D, sp, zero);
}

/*
HLLD flux pass for the y- (D = 1) and z-faces (D = 2) with the per-cell part
carried along a march: thread (i, t) of a BX x BY block walks km cells of
its line along D (t is k for D = 1, j for D = 2), computing hlld_side once
per cell; the side of the cell just walked becomes the left side of the
next face, so a line of km faces needs km + 1 per-cell parts instead of
2 km.  No barrier: each thread's march is independent.  Every value equals
flux_pass_kernel's (the same functions of the same inputs, the same first
error per face); the block's max signal speed is reduced once at the end.
Lines along D are cut into chunks of km cells; [k0, k_end) bounds the
planes of a z-slab.
*/
template <class T, class PT, int D, int MINB>
__global__ __launch_bounds__(BX * BY, MINB) void flux_hlld_march_kernel(GridDev g, StateArrays S,
	PT* __restrict__ P, const double* __restrict__ dtp, StepStatus* st, unsigned gx, int km,
	int remap, int k0, int k_end)
{
	static_assert(D == 1 || D == 2, "march along y or z");
	const unsigned lb = (remap & 1) ? logical_block(blockIdx.x, gridDim.x) : blockIdx.x;
	/* blocks: x tiles fastest, then BY-line groups, then chunks along D */
	const int nt = D == 1 ? k_end - k0 : g.ny;           /* lines across */
	const unsigned gt = (unsigned)((nt + BY - 1) / BY);
	const unsigned bx = lb % gx, r = lb / gx;
	const int t = (int)(r % gt) * BY + threadIdx.y;
	const int chunk = (int)(r / gt);
	const int i = bx * BX + threadIdx.x;
	const int n_d = D == 1 ? g.ny : k_end;                /* march end */
	const int a0 = (D == 1 ? 0 : k0) + chunk * km;        /* first face */
	const int a1 = min(a0 + km, n_d);
	double sp = 0;
	unsigned zero = 0;
	if (i < g.nx && t < nt && a0 < a1) {
		const int jj = D == 1 ? 0 : t;                     /* fixed coordinates */
		const int kk = D == 1 ? k0 + t : 0;
		const int nd = D == 1 ? g.ny : g.nz;
		const T dt = (T)*dtp;
		/* left side of the first face */
		Side<T> Ls;
		T cfl = 0;
		int code_l;
		int stage_l;
		{
			double U[8];
			const int am = wrap_dec(a0, nd);
			load_face_state_d<D>(S, g, i, D == 1 ? am : jj, D == 1 ? kk : am, U);
			stage_l = hlld_side<T>(U, g.ph, Ls, cfl, code_l);
		}
		for (int a = a0; a < a1; a++) {
			const int j = D == 1 ? a : jj;
			const int k = D == 1 ? kk : a;
			double U[8];
			load_face_state_d<D>(S, g, i, j, k, U);
			Side<T> Rs;
			T cfr = 0;
			int code_r;
			const int stage_r = hlld_side<T>(U, g.ph, Rs, cfr, code_r);
			T F[8], sT = 0;
			int e;
			if (stage_l | stage_r) {
				e = first_side_error(stage_l, code_l, stage_r, code_r);
			} else {
				e = flux_hlld_face<T>(Ls, Rs, cfl, cfr, Ls.u, Rs.u, g.ph, F, sT);
			}
			const int c = i + g.nx * (j + g.ny * k);
			T Fg[8];
			if (e) {
				record_error(st, e, (long long)D * g.n + c);
				for (int q = 0; q < 8; q++) Fg[q] = 0;
				sT = 0;
			} else {
				Fg[RHO] = dt * F[RHO];
				Fg[NRJ] = dt * F[NRJ];
				Fg[M0 + D] = dt * F[M0];
				Fg[M0 + (D + 1) % 3] = dt * F[M1];
				Fg[M0 + (D + 2) % 3] = dt * F[M2];
				Fg[B0 + D] = dt * F[B0];
				Fg[B0 + (D + 1) % 3] = dt * F[B1];
				Fg[B0 + (D + 2) % 3] = dt * F[B2];
			}
			PT* Pd = P + (size_t)D * 8 * g.n;
#pragma unroll
			for (int q = 0; q < 8; q++) (Pd + (size_t)q * g.n)[c] = (PT)Fg[q];
			const double s_face = (double)sT;
			sp = smax(sp, s_face);
			zero += s_face <= 0 ? 1u : 0u;
			Ls = Rs;
			cfl = cfr;
			stage_l = stage_r;
			code_l = code_r;
		}
	}
	reduce_speed(st, D, sp, zero);
}

/*
Moves one 32-bit word of a wave's per-cell values one lane up: lane l
receives lane l - 1's r, lane 0 receives b.  Every caller's wave is one
64-lane row of a tile.  lane_up_u32 is a DPP wavefront shift (a VALU move
reading the neighbour lane's register, no LDS round trip; lane 0 has no
source lane and keeps the old value b), used for the FP64 path's per-cell
record; lane_up_u32_xbar goes through the LDS crossbar (ds_bpermute), used
for the mixed path's records, where it measured faster (with DPP moves the
256^3 HLLD FP64 step took 0.4-0.7 % less time, the mixed step 3 % more).
*/
constexpr int DPP_WAVE_SHR1 = 0x138;

__device__ __forceinline__ unsigned lane_up_u32(unsigned r, unsigned b)
{
	return (unsigned)__builtin_amdgcn_update_dpp((int)b, (int)r, DPP_WAVE_SHR1, 0xf, 0xf, false);
}

/* lane l receives lane l + 1's r, lane 63 receives b: the DPP wavefront
   shift the other way, for a double */
__device__ __forceinline__ double lane_down_dpp(double r, double b)
{
	constexpr int DPP_WAVE_SHL1 = 0x130;
	const unsigned long long rv = (unsigned long long)__double_as_longlong(r);
	const unsigned long long bv = (unsigned long long)__double_as_longlong(b);
	const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp((int)(unsigned)bv, (int)(unsigned)rv,
		DPP_WAVE_SHL1, 0xf, 0xf, false);
	const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp((int)(unsigned)(bv >> 32), (int)(unsigned)(rv >> 32),
		DPP_WAVE_SHL1, 0xf, 0xf, false);
	return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

__device__ __forceinline__ unsigned lane_up_u32_xbar(unsigned r, unsigned b)
{
	const unsigned up = (unsigned)__shfl_up((int)r, 1u, 64);
	return threadIdx.x == 0 ? b : up;
}

__device__ __forceinline__ double lane_up(double r, double b)
{
	const unsigned long long rv = (unsigned long long)__double_as_longlong(r);
	const unsigned long long bv = (unsigned long long)__double_as_longlong(b);
	const unsigned lo = lane_up_u32((unsigned)rv, (unsigned)bv);
	const unsigned hi = lane_up_u32((unsigned)(rv >> 32), (unsigned)(bv >> 32));
	return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

__device__ __forceinline__ float lane_up(float r, float b)
{
	return __uint_as_float(lane_up_u32(__float_as_uint(r), __float_as_uint(b)));
}

__device__ __forceinline__ int lane_up(int r, int b)
{
	return (int)lane_up_u32((unsigned)r, (unsigned)b);
}

/* a per-cell HLLD part with its fast speed and check stage, moved as above */
template <class T> struct SideState {
	Side<T> s;
	T cf;
	int stage, code;
};

template <class T> __device__ __forceinline__ SideState<T> lane_up(const SideState<T>& r,
	const SideState<T>& b)
{
	SideState<T> o;
#pragma unroll
	for (int q = 0; q < 8; q++) o.s.u[q] = lane_up(r.s.u[q], b.s.u[q]);
	o.s.p = lane_up(r.s.p, b.s.p);
#pragma unroll
	for (int q = 0; q < 3; q++) o.s.v[q] = lane_up(r.s.v[q], b.s.v[q]);
	o.s.ir = lane_up(r.s.ir, b.s.ir);
	o.s.p64 = lane_up(r.s.p64, b.s.p64);
	o.s.pt64 = lane_up(r.s.pt64, b.s.pt64);
	o.s.rho_pos = lane_up((int)r.s.rho_pos, (int)b.s.rho_pos) != 0;
	o.cf = lane_up(r.cf, b.cf);
	o.stage = lane_up(r.stage, b.stage);
	o.code = lane_up(r.code, b.code);
	return o;
}

/*
HLLD flux pass for the x-faces with the per-cell part computed once per
cell: a wave is the 64 lanes of one x tile (lane l at i = i0 + l), and each
thread walks km cells of its line along y (plane k0 + t, the y-march's
blocks).  At each step every lane computes hlld_side of its own cell; the
left side of lane l's face is lane l - 1's, moved one lane up, and lane 0's
is that of the cell at i0 - 1 (periodic), computed for all km rows of the
march before it starts, one row per lane, and kept in LDS (the wave's own
slots: written and read by the same wave, in order, no block barrier).
Lanes past the grid's end compute the side of the last cell (in bounds, not
used) so that every lane takes part in the moves.  A line of 64 faces needs
64 per-cell parts plus 1 / km for the tile's left edge instead of 128.
Every value equals flux_pass_kernel's (the same functions of the same
inputs, the same first error per face).
*/
template <class T, class PT, int MINB>
__global__ __launch_bounds__(BX * BY, MINB) void flux_hlld_xmarch_kernel(GridDev g, StateArrays S,
	PT* __restrict__ P, const double* __restrict__ dtp, StepStatus* st, unsigned gx, int km,
	int remap, int k0, int k_end)
{
	static_assert(BX == 64, "a wave is one x tile");
	const unsigned lb = (remap & 1) ? logical_block(blockIdx.x, gridDim.x) : blockIdx.x;
	const int nt = k_end - k0;
	const unsigned gt = (unsigned)((nt + BY - 1) / BY);
	const unsigned bx = lb % gx, r = lb / gx;
	const int t = (int)(r % gt) * BY + threadIdx.y;
	const int chunk = (int)(r / gt);
	const int i0 = (int)bx * BX;
	const int i = i0 + threadIdx.x;
	const int a0 = chunk * km;
	const int a1 = min(a0 + km, g.ny);
	double sp = 0;
	unsigned zero = 0;
	/* t, a0 and a1 are the same in every lane of the wave */
	if (t < nt && a0 < a1) {
		const int kk = k0 + t;
		const int ic = min(i, g.nx - 1);
		const T dt = (T)*dtp;
		__shared__ SideState<T> edge[BY][BX];
		{
			SideState<T> e{};
			double U[8];
			load_face_state_d<0>(S, g, wrap_dec(i0, g.nx), a0 + min((int)threadIdx.x, a1 - a0 - 1),
				kk, U);
			e.stage = hlld_side<T>(U, g.ph, e.s, e.cf, e.code);
			edge[threadIdx.y][threadIdx.x] = e;
		}
		__builtin_amdgcn_wave_barrier();
		for (int a = a0; a < a1; a++) {
			double U[8];
			load_face_state_d<0>(S, g, ic, a, kk, U);
			SideState<T> R;
			R.cf = 0;
			R.stage = hlld_side<T>(U, g.ph, R.s, R.cf, R.code);
			const SideState<T> L = lane_up(R, edge[threadIdx.y][a - a0]);
			if (i < g.nx) {
				T F[8], sT = 0;
				int e;
				if (L.stage | R.stage) {
					e = first_side_error(L.stage, L.code, R.stage, R.code);
				} else {
					e = flux_hlld_face<T>(L.s, R.s, L.cf, R.cf, L.s.u, R.s.u, g.ph, F, sT);
				}
				const int c = i + g.nx * (a + g.ny * kk);
				T Fg[8];
				if (e) {
					record_error(st, e, (long long)c);
					for (int q = 0; q < 8; q++) Fg[q] = 0;
					sT = 0;
				} else {
					Fg[RHO] = dt * F[RHO];
					Fg[NRJ] = dt * F[NRJ];
					Fg[M0] = dt * F[M0];
					Fg[M0 + 1] = dt * F[M1];
					Fg[M0 + 2] = dt * F[M2];
					Fg[B0] = dt * F[B0];
					Fg[B0 + 1] = dt * F[B1];
					Fg[B0 + 2] = dt * F[B2];
				}
#pragma unroll
				for (int q = 0; q < 8; q++) (P + (size_t)q * g.n)[c] = (PT)Fg[q];
				const double s_face = (double)sT;
				sp = smax(sp, s_face);
				zero += s_face <= 0 ? 1u : 0u;
			}
		}
	}
	reduce_speed(st, 0, sp, zero);
}

/* any per-cell record moved one lane up, word by word (lane_up_u32_xbar) */
template <class C> __device__ __forceinline__ C lane_up_words(const C& r, const C& b)
{
	static_assert(sizeof(C) % 4 == 0, "whole 32-bit words");
	constexpr int W = sizeof(C) / 4;
	unsigned rw[W], bw[W], ow[W];
	__builtin_memcpy(rw, &r, sizeof(C));
	__builtin_memcpy(bw, &b, sizeof(C));
#pragma unroll
	for (int q = 0; q < W; q++) ow[q] = lane_up_u32_xbar(rw[q], bw[q]);
	C o;
	__builtin_memcpy(&o, ow, sizeof(C));
	return o;
}

/* dt * F of one face rotated back to the global frame and stored (zero
   flux and an error record when e != 0); returns the face's signal speed */
template <class T, class PT, int D>
__device__ __forceinline__ double store_face_flux(const GridDev& g, PT* __restrict__ P,
	StepStatus* st, int c, int e, const T F[8], T sT, T dt)
{
	T Fg[8];
	if (e) {
		record_error(st, e, (long long)D * g.n + c);
		for (int q = 0; q < 8; q++) Fg[q] = 0;
		sT = 0;
	} else {
		Fg[RHO] = dt * F[RHO];
		Fg[NRJ] = dt * F[NRJ];
		Fg[M0 + D] = dt * F[M0];
		Fg[M0 + (D + 1) % 3] = dt * F[M1];
		Fg[M0 + (D + 2) % 3] = dt * F[M2];
		Fg[B0 + D] = dt * F[B0];
		Fg[B0 + (D + 1) % 3] = dt * F[B1];
		Fg[B0 + (D + 2) % 3] = dt * F[B2];
	}
	PT* Pd = P + (size_t)D * 8 * g.n;
#pragma unroll
	for (int q = 0; q < 8; q++) (Pd + (size_t)q * g.n)[c] = (PT)Fg[q];
	return (double)sT;
}

/*
HLL and Rusanov on grids without a background field, with the per-cell
part (flux_cell: checks, pressure, fast speed, physical flux) computed once
per cell: y- and z-faces (D = 1, 2) along the marches of
flux_hlld_march_kernel, x-faces (D = 0) along y with each lane's part moved
to the next lane as in flux_hlld_xmarch_kernel (the tile's left-edge parts
in LDS; marches of at most 32 rows).  Every value equals flux_pass_kernel's.
*/
template <int SOLVER, class T, class PT, int D, int MINB>
__global__ __launch_bounds__(BX * BY, MINB) void flux_cell_march_kernel(GridDev g, StateArrays S,
	PT* __restrict__ P, const double* __restrict__ dtp, StepStatus* st, unsigned gx, int km,
	int remap, int k0, int k_end)
{
	static_assert(BX == 64, "a wave is one x tile");
	const unsigned lb = (remap & 1) ? logical_block(blockIdx.x, gridDim.x) : blockIdx.x;
	/* D = 0, 1: lines across are the planes of the slab, marches along y;
	   D = 2: lines across are the rows, marches along z */
	const int nt = D == 2 ? g.ny : k_end - k0;
	const unsigned gt = (unsigned)((nt + BY - 1) / BY);
	const unsigned bx = lb % gx, r = lb / gx;
	const int t = (int)(r % gt) * BY + threadIdx.y;
	const int chunk = (int)(r / gt);
	const int i0 = (int)bx * BX;
	const int i = i0 + threadIdx.x;
	const int n_d = D == 2 ? k_end : g.ny;
	const int a0 = (D == 2 ? k0 : 0) + chunk * km;
	const int a1 = min(a0 + km, n_d);
	double sp = 0;
	unsigned zero = 0;
	/* D = 0: t, a0 and a1 are the same in every lane of the wave */
	if (t < nt && a0 < a1 && (D == 0 || i < g.nx)) {
		const int jj = D == 2 ? t : 0;
		const int kk = D == 2 ? 0 : k0 + t;
		const int ic = min(i, g.nx - 1);
		const T dt = (T)*dtp;
		/* left-edge parts of up to 32 rows per wave */
		__shared__ FluxCell<T> edge[D == 0 ? BY : 1][D == 0 ? 32 : 1];
		FluxCell<T> Lc;
		if constexpr (D == 0) {
			FluxCell<T> e;
			double U[8];
			load_face_state_d<0>(S, g, wrap_dec(i0, g.nx), a0 + min((int)threadIdx.x, a1 - a0 - 1),
				kk, U);
			flux_cell<SOLVER, T>(U, g.ph, e);
			if (threadIdx.x < 32) edge[threadIdx.y][threadIdx.x] = e;
			__builtin_amdgcn_wave_barrier();
		} else {
			double U[8];
			const int am = wrap_dec(a0, D == 1 ? g.ny : g.nz);
			load_face_state_d<D>(S, g, i, D == 1 ? am : jj, D == 1 ? kk : am, U);
			flux_cell<SOLVER, T>(U, g.ph, Lc);
		}
		for (int a = a0; a < a1; a++) {
			const int j = D == 2 ? jj : a;
			const int k = D == 2 ? a : kk;
			double U[8];
			load_face_state_d<D>(S, g, D == 0 ? ic : i, j, k, U);
			FluxCell<T> Rc;
			flux_cell<SOLVER, T>(U, g.ph, Rc);
			if constexpr (D == 0) Lc = lane_up_words(Rc, edge[threadIdx.y][a - a0]);
			if (D != 0 || i < g.nx) {
				T F[8], sT = 0;
				const int e = flux_face<SOLVER, T>(Lc, Rc, F, sT);
				const int c = i + g.nx * (j + g.ny * k);
				const double s_face = store_face_flux<T, PT, D>(g, P, st, c, e, F, sT, dt);
				sp = smax(sp, s_face);
				zero += s_face <= 0 ? 1u : 0u;
			}
			if constexpr (D != 0) Lc = Rc;
		}
	}
	reduce_speed(st, D, sp, zero);
}

/* constant-thermal-pressure energy refit of update_vol_B,
   source/mhd/solve.hpp:1646-1672 */
__device__ __forceinline__ void refit_energy(double rho, const double m[3], double& E,
	const double vb_old[3], const double vb_new[3], const Phys& ph)
{
	if (rho <= 0) return;
	const double U[8] = {rho, m[0], m[1], m[2], E, vb_old[0], vb_old[1], vb_old[2]};
	const double p = pressure64(U, ph.gm1, ph.half_inv_mu0);
	const double inv = 1 / rho;
	const double v[3] = {m[0] * inv, m[1] * inv, m[2] * inv};
	const double
		kinetic = 0.5 * rho * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]),
		magnetic = ph.half_inv_mu0 * (vb_new[0] * vb_new[0] + vb_new[1] * vb_new[1] + vb_new[2] * vb_new[2]);
	E = p / ph.gm1 + kinetic + magnetic;
}

/*
Constrained-transport face changes.  Each face sums the 12 edge terms of
assign_face_dBs_f{x,y,z} (source/mhd/solve.hpp:610-613, :926-928,
:1223-1225) that touch its edges, in a fixed order (the face's own
direction, then the next two cyclically), from dt * F of the faces owned by
the cells around it:
  x-face: Ez = +F_By dt dz/4, Ey = -F_Bz dt dy/4
  y-face: Ex = +F_Bz dt dx/4, Ez = -F_Bx dt dz/4
  z-face: Ex = -F_By dt dx/4, Ey = +F_Bx dt dy/4
*/
template <class PT> struct EdgeTerms {
	const PT* P;
	long long n;
	double dx, dy, dz;
	__device__ __forceinline__ double ezx(long long c) const { return +((double)P[6 * n + c] * dz / 4); }
	__device__ __forceinline__ double eyx(long long c) const { return -((double)P[7 * n + c] * dy / 4); }
	__device__ __forceinline__ double exy(long long c) const { return +((double)P[15 * n + c] * dx / 4); }
	__device__ __forceinline__ double ezy(long long c) const { return -((double)P[13 * n + c] * dz / 4); }
	__device__ __forceinline__ double exz(long long c) const { return -((double)P[22 * n + c] * dx / 4); }
	__device__ __forceinline__ double eyz(long long c) const { return +((double)P[21 * n + c] * dy / 4); }
};

struct Nb {
	int nx, ny;
	__device__ __forceinline__ long long operator()(int i, int j, int k) const
	{
		return i + (long long)nx * (j + (long long)ny * k);
	}
};

/* the same over a grid with closed dimensions: a coordinate of -1 (beyond
   a closed end) names no cell, and its flux terms are those of a zero
   flux */
struct EdgeTermsB {
	const double* P;
	long long n;
	double dx, dy, dz;
	__device__ __forceinline__ double at(int q, long long c) const { return c < 0 ? 0.0 : P[q * n + c]; }
	__device__ __forceinline__ double ezx(long long c) const { return +(at(6, c) * dz / 4); }
	__device__ __forceinline__ double eyx(long long c) const { return -(at(7, c) * dy / 4); }
	__device__ __forceinline__ double exy(long long c) const { return +(at(15, c) * dx / 4); }
	__device__ __forceinline__ double ezy(long long c) const { return -(at(13, c) * dz / 4); }
	__device__ __forceinline__ double exz(long long c) const { return -(at(22, c) * dx / 4); }
	__device__ __forceinline__ double eyz(long long c) const { return +(at(21, c) * dy / 4); }
};

struct NbB {
	int nx, ny;
	__device__ __forceinline__ long long operator()(int i, int j, int k) const
	{
		return (i < 0 || j < 0 || k < 0) ? -1 : i + (long long)nx * (j + (long long)ny * k);
	}
};

/*
The 12 edge terms of one face change, summed in a fixed order; the
arguments are the terms in that order (see dBx_at, dBy_at, dBz_at).
*/
__device__ __forceinline__ double dBx_sum(double a0, double a1, double a2, double a3, double a4,
	double a5, double a6, double a7, double a8, double a9, double a10, double a11)
{
	double r = 0;
	r += a0; r -= a1; r -= a2; r += a3; r += a4; r += a5;
	r -= a6; r -= a7; r -= a8; r -= a9; r += a10; r += a11;
	return r;
}

__device__ __forceinline__ double dBy_sum(double a0, double a1, double a2, double a3, double a4,
	double a5, double a6, double a7, double a8, double a9, double a10, double a11)
{
	double r = 0;
	r += a0; r -= a1; r -= a2; r += a3; r += a4; r += a5;
	r -= a6; r -= a7; r -= a8; r += a9; r -= a10; r += a11;
	return r;
}

/* dBz_sum has the signs of dBx_sum */
__device__ __forceinline__ double dBz_sum(double a0, double a1, double a2, double a3, double a4,
	double a5, double a6, double a7, double a8, double a9, double a10, double a11)
{
	return dBx_sum(a0, a1, a2, a3, a4, a5, a6, a7, a8, a9, a10, a11);
}

/* dB of the x-face owned by (i, j, k); im = i-1, jm/jp = j-+1, km/kp = k-+1 */
template <class ET, class NB> __device__ __forceinline__ double dBx_at(const ET& E, const NB& L,
	int i, int im, int j, int jm, int jp, int k, int km, int kp)
{
	return dBx_sum(
		E.ezx(L(i, jp, k)), E.ezx(L(i, jm, k)), E.eyx(L(i, j, kp)), E.eyx(L(i, j, km)),
		E.ezy(L(i, jp, k)), E.ezy(L(im, jp, k)), E.ezy(L(i, j, k)), E.ezy(L(im, j, k)),
		E.eyz(L(i, j, kp)), E.eyz(L(im, j, kp)), E.eyz(L(i, j, k)), E.eyz(L(im, j, k)));
}

template <class ET, class NB> __device__ __forceinline__ double dBy_at(const ET& E, const NB& L,
	int i, int im, int ip, int j, int jm, int k, int km, int kp)
{
	return dBy_sum(
		E.exy(L(i, j, kp)), E.exy(L(i, j, km)), E.ezy(L(ip, j, k)), E.ezy(L(im, j, k)),
		E.exz(L(i, j, kp)), E.exz(L(i, jm, kp)), E.exz(L(i, j, k)), E.exz(L(i, jm, k)),
		E.ezx(L(ip, j, k)), E.ezx(L(i, j, k)), E.ezx(L(ip, jm, k)), E.ezx(L(i, jm, k)));
}

template <class ET, class NB> __device__ __forceinline__ double dBz_at(const ET& E, const NB& L,
	int i, int im, int ip, int j, int jm, int jp, int k, int km)
{
	return dBz_sum(
		E.eyz(L(ip, j, k)), E.eyz(L(im, j, k)), E.exz(L(i, jp, k)), E.exz(L(i, jm, k)),
		E.eyx(L(ip, j, k)), E.eyx(L(ip, j, km)), E.eyx(L(i, j, k)), E.eyx(L(i, j, km)),
		E.exy(L(i, jp, k)), E.exy(L(i, jp, km)), E.exy(L(i, j, k)), E.exy(L(i, j, km)));
}

/* the six transverse magnetic fluxes (dt * F) of the three faces cell c
   owns, in the layout of magnetic_fluxes_kernel: slot 2 d + q - 1 (q = 1, 2)
   holds global field component 5 + (d + q) % 3 of the d-face; lo(d, comp)
   reads the flux array's d-face value of component comp at c */
template <class LO>
__device__ __forceinline__ void store_magnetic_fluxes(double* __restrict__ pe, long long n, long long c,
	const LO& lo)
{
#pragma unroll
	for (int d = 0; d < 3; d++)
#pragma unroll
	for (int q = 1; q < 3; q++) pe[(2 * d + q - 1) * n + c] = lo(d, B0 + (d + q) % 3);
}

/*
Update pass, column-marching variant: update_mhd_state + update_vol_B of
source/mhd/solve.hpp:1503-1683 with thread (i, j) of a TX x TY column
tile walking the planes [k0, k0 + kz) of its column.  Of the edge terms the
six face changes around a cell use, those of plane k + 1 are loaded once,
at plane k, and carried in registers through the next two planes; the
+z flux row of plane k + 1 becomes the next cell's own -z row, and the new
+z face field the next cell's -z face.  Per plane only the 12 terms of the
edges parallel to z are loaded fresh (their x and y neighbours are loaded
by the adjacent threads at the same time).  Every sum has the order of
update_cell_kernel, so both kernels give identical results.
*/
template <class PT, int TX, int TY>
__global__ __launch_bounds__(TX * TY) void update_column_kernel(GridDev g,
	const double* __restrict__ s, double* __restrict__ out,
	const PT* __restrict__ P, int refits, int kz, unsigned gx, unsigned gy, int remap,
	double* __restrict__ pe)
{
	const unsigned lb = remap ? logical_block(blockIdx.x, gridDim.x) : blockIdx.x;
	const unsigned bx = lb % gx, by = (lb / gx) % gy, bz = lb / (gx * gy);
	const int i = bx * TX + threadIdx.x;
	const int j = by * TY + threadIdx.y;
	if (i >= g.nx || j >= g.ny) return;
	const long long n = g.n;
	const int im = wrap_dec(i, g.nx), ip = wrap_inc(i, g.nx);
	const int jm = wrap_dec(j, g.ny), jp = wrap_inc(j, g.ny);
	const Nb L{g.nx, g.ny};
	const EdgeTerms<PT> E{P, n, g.dx, g.dy, g.dz};
	const double axy = g.dx * g.dy, axz = g.dx * g.dz, ayz = g.dy * g.dz;
	const double len[3] = {g.dx, g.dy, g.dz};
	const PT* P0 = P;
	const PT* P1 = P + 8 * n;
	const PT* P2 = P + 16 * n;
	const int k0 = (int)bz * kz;
	const int k1 = min(k0 + kz, g.nz);
	if (k0 >= k1) return;

	/* terms of planes k - 1 (suffix m) and k (suffix c) for k = k0; p / m
	   before the plane suffix = the cell at +1 / -1 in x (eyx, eyz) or y
	   (exy, exz) */
	const int kb = wrap_dec(k0, g.nz);
	double eyx_m = E.eyx(L(i, j, kb)), eyx_pm = E.eyx(L(ip, j, kb));
	double exy_m = E.exy(L(i, j, kb)), exy_pm = E.exy(L(i, jp, kb));
	double eyx_c = E.eyx(L(i, j, k0)), eyx_pc = E.eyx(L(ip, j, k0));
	double exy_c = E.exy(L(i, j, k0)), exy_pc = E.exy(L(i, jp, k0));
	double eyz_c = E.eyz(L(i, j, k0)), eyz_mc = E.eyz(L(im, j, k0)), eyz_pc = E.eyz(L(ip, j, k0));
	double exz_c = E.exz(L(i, j, k0)), exz_mc = E.exz(L(i, jm, k0)), exz_pc = E.exz(L(i, jp, k0));
	const long long cb = L(i, j, k0);
	double s7 = s[7 * n + cb];
	/* new -z face of the first cell (dBz_at at k0) */
	double bz_new = s7 + dBz_sum(eyz_pc, eyz_mc, exz_pc, exz_mc, eyx_pc, eyx_pm, eyx_c, eyx_m,
		exy_pc, exy_pm, exy_c, exy_m) / axy;
	double lo2[8];
#pragma unroll
	for (int q = 0; q < 8; q++) lo2[q] = (double)P2[(long long)q * n + cb];

	for (int k = k0; k < k1; k++) {
		const int kp = wrap_inc(k, g.nz);
		const long long c = L(i, j, k), cxp = L(ip, j, k), cyp = L(i, jp, k), czp = L(i, j, kp);
		/* edges parallel to z in this plane */
		const double
			ezx_c = E.ezx(c), ezx_jm = E.ezx(L(i, jm, k)), ezx_jp = E.ezx(cyp),
			ezx_p = E.ezx(cxp), ezx_pjm = E.ezx(L(ip, jm, k)), ezx_pjp = E.ezx(L(ip, jp, k)),
			ezy_c = E.ezy(c), ezy_m = E.ezy(L(im, j, k)), ezy_p = E.ezy(cxp),
			ezy_jp = E.ezy(cyp), ezy_mjp = E.ezy(L(im, jp, k)), ezy_pjp = E.ezy(L(ip, jp, k));
		/* plane k + 1 */
		const double
			eyx_n = E.eyx(czp), eyx_pn = E.eyx(L(ip, j, kp)),
			exy_n = E.exy(czp), exy_pn = E.exy(L(i, jp, kp)),
			eyz_n = E.eyz(czp), eyz_mn = E.eyz(L(im, j, kp)), eyz_pn = E.eyz(L(ip, j, kp)),
			exz_n = E.exz(czp), exz_mn = E.exz(L(i, jm, kp)), exz_pn = E.exz(L(i, jp, kp));

		const double s7n = s[7 * n + czp];
		const double bx_new = s[5 * n + c] + dBx_sum(ezx_jp, ezx_jm, eyx_n, eyx_m, ezy_jp, ezy_mjp,
			ezy_c, ezy_m, eyz_n, eyz_mn, eyz_c, eyz_mc) / ayz;
		const double by_new = s[6 * n + c] + dBy_sum(exy_n, exy_m, ezy_p, ezy_m, exz_n, exz_mn,
			exz_c, exz_mc, ezx_p, ezx_c, ezx_pjm, ezx_jm) / axz;
		const double bxp_new = s[5 * n + cxp] + dBx_sum(ezx_pjp, ezx_pjm, eyx_pn, eyx_pm, ezy_pjp,
			ezy_jp, ezy_p, ezy_c, eyz_pn, eyz_n, eyz_pc, eyz_c) / ayz;
		const double byp_new = s[6 * n + cyp] + dBy_sum(exy_pn, exy_pm, ezy_pjp, ezy_mjp, exz_pn,
			exz_n, exz_pc, exz_c, ezx_pjp, ezx_jp, ezx_p, ezx_c) / axz;
		const double bzp_new = s7n + dBz_sum(eyz_pn, eyz_mn, exz_pn, exz_mn, eyx_pn, eyx_pc, eyx_n,
			eyx_c, exy_pn, exy_pc, exy_n, exy_c) / axy;
		out[5 * n + c] = bx_new;
		out[6 * n + c] = by_new;
		out[7 * n + c] = bz_new;

		double hi2[8];
#pragma unroll
		for (int q = 0; q < 8; q++) hi2[q] = (double)P2[(long long)q * n + czp];
		auto lo = [&](int d, int q) -> double {
			return d == 0 ? (double)P0[(long long)q * n + c]
				: d == 1 ? (double)P1[(long long)q * n + c] : lo2[q];
		};
		auto hi = [&](int d, int q) -> double {
			return d == 0 ? (double)P0[(long long)q * n + cxp]
				: d == 1 ? (double)P1[(long long)q * n + cyp] : hi2[q];
		};
		double rho = s[c];
		double m[3] = {s[n + c], s[2 * n + c], s[3 * n + c]};
		double En = s[4 * n + c];
		double vb[3] = {
			0.5 * (s[5 * n + c] + s[5 * n + cxp]),
			0.5 * (s[6 * n + c] + s[6 * n + cyp]),
			0.5 * (s7 + s7n)};
#pragma unroll
		for (int d = 0; d < 3; d++) rho += ((0.0 + lo(d, RHO)) - (0.0 + hi(d, RHO))) / len[d];
#pragma unroll
		for (int d = 0; d < 3; d++) {
			const double il = 1 / len[d];
#pragma unroll
			for (int q = 0; q < 3; q++) m[q] = m[q] + ((0.0 + lo(d, M0 + q)) - (0.0 + hi(d, M0 + q))) * il;
		}
#pragma unroll
		for (int d = 0; d < 3; d++) En += ((0.0 + lo(d, NRJ)) - (0.0 + hi(d, NRJ))) / len[d];
#pragma unroll
		for (int d = 0; d < 3; d++) {
			const double il = 1 / len[d];
#pragma unroll
			for (int q = 0; q < 3; q++) vb[q] = vb[q] + ((0.0 + lo(d, B0 + q)) - (0.0 + hi(d, B0 + q))) * il;
		}
		const double vb_new[3] = {0.5 * (bx_new + bxp_new), 0.5 * (by_new + byp_new),
			0.5 * (bz_new + bzp_new)};
		for (int r = 0; r < refits; r++) {
			refit_energy(rho, m, En, r == 0 ? vb : vb_new, vb_new, g.ph);
		}
		out[c] = rho;
		out[n + c] = m[0];
		out[2 * n + c] = m[1];
		out[3 * n + c] = m[2];
		out[4 * n + c] = En;
		if (pe) store_magnetic_fluxes(pe, n, c, lo);

		/* plane k + 1 becomes the current plane */
		eyx_m = eyx_c; eyx_pm = eyx_pc; exy_m = exy_c; exy_pm = exy_pc;
		eyx_c = eyx_n; eyx_pc = eyx_pn; exy_c = exy_n; exy_pc = exy_pn;
		eyz_c = eyz_n; eyz_mc = eyz_mn; eyz_pc = eyz_pn;
		exz_c = exz_n; exz_mc = exz_mn; exz_pc = exz_pn;
		bz_new = bzp_new;
		s7 = s7n;
#pragma unroll
		for (int q = 0; q < 8; q++) lo2[q] = hi2[q];
	}
}

/*
Per-cell update variant (one thread per cell, no marching): the same sums in
the same order as update_column_kernel, recomputing the positive-side faces'
dB instead of carrying them between planes; also the kernel of the z-slab
pipeline.
*/
template <class PT>
__global__ __launch_bounds__(BX * BY) void update_cell_kernel(GridDev g,
	const double* __restrict__ s, double* __restrict__ out,
	const PT* __restrict__ P, int refits, unsigned gx, unsigned gy, int remap, int k0,
	double* __restrict__ pe)
{
	const unsigned lb = (remap & 1) ? logical_block(blockIdx.x, gridDim.x) : blockIdx.x;
	unsigned bx, by, k;
	tile_of_block(lb, gx, gy, gridDim.x, remap, k0, bx, by, k);
	const int i = bx * BX + threadIdx.x;
	const int j = by * BY + threadIdx.y;
	if (i >= g.nx || j >= g.ny) return;
	const long long n = g.n;
	const int im = wrap_dec(i, g.nx), ip = wrap_inc(i, g.nx);
	const int jm = wrap_dec(j, g.ny), jp = wrap_inc(j, g.ny);
	const int kk = (int)k, km = wrap_dec(kk, g.nz), kp = wrap_inc(kk, g.nz), kpp = wrap_inc(kp, g.nz);
	const Nb L{g.nx, g.ny};
	const EdgeTerms<PT> E{P, n, g.dx, g.dy, g.dz};
	const double axy = g.dx * g.dy, axz = g.dx * g.dz, ayz = g.dy * g.dz;
	const double len[3] = {g.dx, g.dy, g.dz};
	const long long c = L(i, j, kk), cxp = L(ip, j, kk), cyp = L(i, jp, kk), czp = L(i, j, kp);
	const double bx_new = s[5 * n + c] + dBx_at(E, L, i, im, j, jm, jp, kk, km, kp) / ayz;
	const double by_new = s[6 * n + c] + dBy_at(E, L, i, im, ip, j, jm, kk, km, kp) / axz;
	const double bz_new = s[7 * n + c] + dBz_at(E, L, i, im, ip, j, jm, jp, kk, km) / axy;
	const double bxp_new = s[5 * n + cxp] + dBx_at(E, L, ip, i, j, jm, jp, kk, km, kp) / ayz;
	const double byp_new = s[6 * n + cyp] + dBy_at(E, L, i, im, ip, jp, j, kk, km, kp) / axz;
	const double bzp_new = s[7 * n + czp] + dBz_at(E, L, i, im, ip, j, jm, jp, kp, kk) / axy;
	out[5 * n + c] = bx_new;
	out[6 * n + c] = by_new;
	out[7 * n + c] = bz_new;
	(void)kpp;

	const long long cp[3] = {cxp, cyp, czp};
	double rho = s[c];
	double m[3] = {s[n + c], s[2 * n + c], s[3 * n + c]};
	double En = s[4 * n + c];
	double vb[3] = {
		0.5 * (s[5 * n + c] + s[5 * n + cxp]),
		0.5 * (s[6 * n + c] + s[6 * n + cyp]),
		0.5 * (s[7 * n + c] + s[7 * n + czp])};
	for (int d = 0; d < 3; d++) {
		const PT* Pd = P + (long long)d * 8 * n;
		rho += ((0.0 + (double)Pd[c]) - (0.0 + (double)Pd[cp[d]])) / len[d];
	}
	for (int d = 0; d < 3; d++) {
		const PT* Pd = P + (long long)d * 8 * n;
		const double il = 1 / len[d];
		for (int q = 0; q < 3; q++) {
			const long long o = (long long)(M0 + q) * n;
			m[q] = m[q] + ((0.0 + (double)Pd[o + c]) - (0.0 + (double)Pd[o + cp[d]])) * il;
		}
	}
	for (int d = 0; d < 3; d++) {
		const PT* Pd = P + (long long)d * 8 * n;
		const long long o = (long long)NRJ * n;
		En += ((0.0 + (double)Pd[o + c]) - (0.0 + (double)Pd[o + cp[d]])) / len[d];
	}
	for (int d = 0; d < 3; d++) {
		const PT* Pd = P + (long long)d * 8 * n;
		const double il = 1 / len[d];
		for (int q = 0; q < 3; q++) {
			const long long o = (long long)(B0 + q) * n;
			vb[q] = vb[q] + ((0.0 + (double)Pd[o + c]) - (0.0 + (double)Pd[o + cp[d]])) * il;
		}
	}
	const double vb_new[3] = {0.5 * (bx_new + bxp_new), 0.5 * (by_new + byp_new), 0.5 * (bz_new + bzp_new)};
	for (int r = 0; r < refits; r++) {
		refit_energy(rho, m, En, r == 0 ? vb : vb_new, vb_new, g.ph);
	}
	out[c] = rho;
	out[n + c] = m[0];
	out[2 * n + c] = m[1];
	out[3 * n + c] = m[2];
	out[4 * n + c] = En;
	if (pe) {
		store_magnetic_fluxes(pe, n, c, [&](int d, int q) -> double {
			return (double)P[((long long)d * 8 + q) * n + c];
		});
	}
}

/*
Fused step.  Pass 1 (fused_flux_kernel) computes every face flux of the
step and finishes update_mhd_state's conserved update of each cell
(source/mhd/solve.hpp:1503-1614) in one pass over the state, with no flux
array: of each face only the two field components the constrained-transport
edge terms use are written (6 doubles per cell, `emf`).  Pass 2
(ct_refit_kernel) is the face-field update of assign_face_dBs_* and the
energy refits of update_vol_B (solve.hpp:593-1415, :1622-1683), reading
those 6 values per cell.  Both passes repeat the two-pass path's arithmetic
in its order (update_column_kernel), so the results are identical.

Pass 1 block: 64 lanes along x (lane l is cell i0 + l of the x tile) by NW
waves along y (wave w is row j0 - 1 + w: waves 1 .. NW-2 are the tile's
rows, waves 0 and NW-1 the rows below and above it), each thread marching
along z over a chunk of KC planes.  Per plane:
  * x-faces: the left side is the previous lane's per-cell part (moved one
    lane up), lane 0's that of cell i0 - 1, computed for all planes of the
    chunk before the march; the right neighbour's face flux comes one lane
    down, lane 63's from the face of cell i0 + 64, also computed before the
    march;
  * y-faces: per-cell parts of all rows go through LDS; the face flux of
    the row above comes back through LDS (two barriers per plane);
  * z-faces: the per-cell part of plane k is carried in registers to the
    face between k and k + 1.
All indices wrap periodically, so every access is in bounds for any grid
size; lanes and waves beyond the grid compute duplicates that are neither
written nor counted.  The conserved update is accumulated in the order of
the two-pass path: the x terms, then y, then z of each variable.
*/
constexpr int FLANES = 64;
constexpr int FKC_MAX = 32;      /* planes per chunk, at most */

template <int D> __device__ __forceinline__ void to_frame(const double Ug[8], double U[8])
{
	U[RHO] = Ug[RHO];
	U[NRJ] = Ug[NRJ];
	U[M0] = Ug[M0 + D];
	U[M1] = Ug[M0 + (D + 1) % 3];
	U[M2] = Ug[M0 + (D + 2) % 3];
	U[B0] = Ug[B0 + D];
	U[B1] = Ug[B0 + (D + 1) % 3];
	U[B2] = Ug[B0 + (D + 2) % 3];
}

/* HLLD flux of the face between parts L and R, dt-scaled and rotated back
   to the global frame (zero on error); returns the error, sets the speed.
   FP64: flux_hlld_face; mixed: flux_hlld_mixed_face (FP64 results) */
template <class T, int D> __device__ __forceinline__ int fused_face(const XPart<T>& L,
	const XPart<T>& R, const Phys& ph, double dt, double Fg[8], double& speed)
{
	const int sl = L.sc >> 4, sr = R.sc >> 4;
	int e;
	double F[8];
	double sT = 0;
	if (sl | sr) {
		e = first_side_error(sl, L.sc & 15, sr, R.sc & 15);
	} else if constexpr (std::is_same_v<T, float>) {
		e = flux_hlld_mixed_face(L, R, ph, F, sT);
	} else {
		Side<T> Ls, Rs;
		unpack_side<T>(L, ph, Ls);
		unpack_side<T>(R, ph, Rs);
		e = flux_hlld_face<T>(Ls, Rs, L.cf, R.cf, L.u, R.u, ph, F, sT);
	}
	if (e) {
#pragma unroll
		for (int q = 0; q < 8; q++) Fg[q] = 0;
		speed = 0;
		return e;
	}
	Fg[RHO] = dt * F[RHO];
	Fg[NRJ] = dt * F[NRJ];
	Fg[M0 + D] = dt * F[M0];
	Fg[M0 + (D + 1) % 3] = dt * F[M1];
	Fg[M0 + (D + 2) % 3] = dt * F[M2];
	Fg[B0 + D] = dt * F[B0];
	Fg[B0 + (D + 1) % 3] = dt * F[B1];
	Fg[B0 + (D + 2) % 3] = dt * F[B2];
	speed = (double)sT;
	return ERR_OK;
}

template <class T> __device__ __forceinline__ XPart<T> lane_up(const XPart<T>& r, const XPart<T>& b)
{
	return lane_up_words(r, b);
}

/* lane l receives lane l + 1's value, lane 63 receives b */
__device__ __forceinline__ unsigned lane_down_u32(unsigned r, unsigned b)
{
	const unsigned dn = (unsigned)__shfl_down((int)r, 1u, 64);
	return threadIdx.x == 63 ? b : dn;
}

template <class C> __device__ __forceinline__ C lane_down_words(const C& r, const C& b)
{
	static_assert(sizeof(C) % 4 == 0, "whole 32-bit words");
	constexpr int W = sizeof(C) / 4;
	unsigned rw[W], bw[W], ow[W];
	__builtin_memcpy(rw, &r, sizeof(C));
	__builtin_memcpy(bw, &b, sizeof(C));
#pragma unroll
	for (int q = 0; q < W; q++) ow[q] = lane_down_u32(rw[q], bw[q]);
	C o;
	__builtin_memcpy(&o, ow, sizeof(C));
	return o;
}

template <class T> struct Flux8 { T f[8]; };

/* block-wide max of s and count of s <= 0 over NWAVES waves, one atomic per block */
template <int NWAVES> __device__ __forceinline__ void reduce_speed_blk(StepStatus* st, int d,
	double s, unsigned zero, double* ms, unsigned* mz)
{
	for (int off = 32; off > 0; off >>= 1) {
		const double o = __shfl_xor(s, off, 64);
		s = smax(s, o);
		zero += __shfl_xor(zero, off, 64);
	}
	const int tid = threadIdx.x + blockDim.x * threadIdx.y;
	const int wave = tid / 64, lane = tid % 64;
	if (lane == 0) {
		ms[wave] = s;
		mz[wave] = zero;
	}
	__syncthreads();
	if (tid == 0) {
		double m = ms[0];
		unsigned z = mz[0];
		for (int w = 1; w < NWAVES; w++) {
			m = smax(m, ms[w]);
			z += mz[w];
		}
		const unsigned long long bits = m > 0 ? (unsigned long long)__double_as_longlong(m) : 0ull;
		atomicMax(&st->smax_bits[d][blockIdx.x % NSLOT], bits);
		if (z) atomicAdd(&st->zero[d], z);
	}
	__syncthreads();
}

/* the state of cell c in the global frame: conserved variables and the
   cell-centred field, the average of its two faces (update_vol_B) */
__device__ __forceinline__ void load_cell(const StateArrays& S, int c, int cx, int cy, int cz,
	double U[8])
{
	U[RHO] = S.q[0][c];
	U[M0] = S.q[1][c];
	U[M1] = S.q[2][c];
	U[M2] = S.q[3][c];
	U[NRJ] = S.q[4][c];
	U[B0] = 0.5 * (S.q[5][c] + S.q[5][cx]);
	U[B1] = 0.5 * (S.q[6][c] + S.q[6][cy]);
	U[B2] = 0.5 * (S.q[7][c] + S.q[7][cz]);
}

/* conserved update by one direction's two faces, update_column_kernel's
   order: (0 + lo - (0 + hi)) / len for rho and E, * (1 / len) for m and
   the cell-centred field */
template <class T> __device__ __forceinline__ void add_terms(double acc[8], const T lo[8],
	const T hi[8], double len, double il)
{
	acc[RHO] += ((0.0 + (double)lo[RHO]) - (0.0 + (double)hi[RHO])) / len;
#pragma unroll
	for (int q = 0; q < 3; q++)
		acc[M0 + q] = acc[M0 + q] + ((0.0 + (double)lo[M0 + q]) - (0.0 + (double)hi[M0 + q])) * il;
	acc[NRJ] += ((0.0 + (double)lo[NRJ]) - (0.0 + (double)hi[NRJ])) / len;
#pragma unroll
	for (int q = 0; q < 3; q++)
		acc[B0 + q] = acc[B0 + q] + ((0.0 + (double)lo[B0 + q]) - (0.0 + (double)hi[B0 + q])) * il;
}

template <class T, int NW> struct FusedShared {
	XPart<T> yside[NW][FLANES];
	Flux8<double> fy[NW][FLANES];
	XPart<T> edge_l[NW][FKC_MAX];
	Flux8<double> edge_r[NW][FKC_MAX];
	double ms[NW];
	unsigned mz[NW];
};

template <class T, int NW>
__global__ __launch_bounds__(FLANES * NW) void fused_flux_kernel(GridDev g, StateArrays S,
	double* __restrict__ out, double* __restrict__ emf, const double* __restrict__ dtp,
	StepStatus* st, unsigned gx, unsigned gy, int kc, int refits, int remap)
{
	__shared__ FusedShared<T, NW> sh;
	const unsigned lb = (remap & 1) ? logical_block(blockIdx.x, gridDim.x) : blockIdx.x;
	const unsigned bx = lb % gx, by = (lb / gx) % gy, chunk = lb / (gx * gy);
	const int lane = threadIdx.x, w = threadIdx.y;
	const int nx = g.nx, ny = g.ny, nz = g.nz, n = (int)g.n;
	const int i0 = (int)bx * FLANES, j0 = (int)by * (NW - 2);
	const int i = (i0 + lane) % nx;
	const int j = ((j0 - 1 + w) % ny + ny) % ny;
	const bool own_x = i0 + lane < nx;
	const bool inner = w >= 1 && w <= NW - 2 && j0 + w - 1 < ny;   /* a row of the tile */
	const bool own = own_x && inner;
	const int k0 = (int)chunk * kc, k1 = min(k0 + kc, nz);
	const double len[3] = {g.dx, g.dy, g.dz};
	const double il[3] = {1 / g.dx, 1 / g.dy, 1 / g.dz};
	const double dt = *dtp;
	const int nxy = nx * ny;
	auto cell = [&](int ii, int jj, int kk) { return ii + nx * jj + nxy * kk; };
	auto load = [&](int ii, int jj, int kk, double U[8]) {
		load_cell(S, cell(ii, jj, kk), cell(wrap_inc(ii, nx), jj, kk), cell(ii, wrap_inc(jj, ny), kk),
			cell(ii, jj, wrap_inc(kk, nz)), U);
	};
	double sp[3] = {0, 0, 0};
	unsigned zero[3] = {0, 0, 0};

	/* before the march: parts of the tile's left-edge cells and the face
	   fluxes at its right edge, one plane per lane */
	if (inner && lane < k1 - k0) {
		const int kk = k0 + lane;
		double Ug[8], U[8];
		load(wrap_dec(i0 % nx, nx), j, kk, Ug);
		to_frame<0>(Ug, U);
		hlld_part<T>(U, g.ph, sh.edge_l[w][lane]);
		if (i0 + FLANES <= nx) {
			const int a = (i0 + FLANES - 1) % nx, b = (i0 + FLANES) % nx;
			XPart<T> pa, pb;
			load(a, j, kk, Ug);
			to_frame<0>(Ug, U);
			hlld_part<T>(U, g.ph, pa);
			load(b, j, kk, Ug);
			to_frame<0>(Ug, U);
			hlld_part<T>(U, g.ph, pb);
			double s_unused;
			fused_face<T, 0>(pa, pb, g.ph, dt, sh.edge_r[w][lane].f, s_unused);
		}
	}
	/* the z-face owned by plane k0 */
	double Uc[8];               /* state of the current plane, global frame */
	XPart<T> zc;                /* its z-frame part */
	double fz_lo[8];            /* dt * F of its -z face */
	{
		double Ug[8], U[8];
		XPart<T> zm;
		load(i, j, wrap_dec(k0, nz), Ug);
		to_frame<2>(Ug, U);
		hlld_part<T>(U, g.ph, zm);
		load(i, j, k0, Uc);
		to_frame<2>(Uc, U);
		hlld_part<T>(U, g.ph, zc);
		if (inner) {
			double s_face;
			const int e = fused_face<T, 2>(zm, zc, g.ph, dt, fz_lo, s_face);
			if (own) {
				const int c = cell(i, j, k0);
				if (e) record_error(st, e, 2ll * n + c);
				sp[2] = smax(sp[2], s_face);
				zero[2] += s_face <= 0 ? 1u : 0u;
				emf[4 * n + c] = fz_lo[B1];
				emf[5 * n + c] = fz_lo[B0];
			}
		}
	}
	__syncthreads();

	for (int k = k0; k < k1; k++) {
		const int kn = wrap_inc(k, nz);
		double Un[8];
		load(i, j, kn, Un);
		/* y part of this plane, for the face of the row above */
		{
			double U[8];
			to_frame<1>(Uc, U);
			hlld_part<T>(U, g.ph, sh.yside[w][lane]);
		}
		double acc[8];
		double fz_hi[8];
		XPart<T> zn;
		if (inner) {
			/* x-face of this cell and of the next lane's */
			XPart<T> px;
			{
				double U[8];
				to_frame<0>(Uc, U);
				hlld_part<T>(U, g.ph, px);
			}
			const XPart<T> pl = lane_up(px, sh.edge_l[w][k - k0]);
			double fx[8];
			double s_face;
			const int e = fused_face<T, 0>(pl, px, g.ph, dt, fx, s_face);
			Flux8<double> lo, hi;
#pragma unroll
			for (int q = 0; q < 8; q++) lo.f[q] = fx[q];
			hi = lane_down_words(lo, sh.edge_r[w][k - k0]);
			const int c = cell(i, j, k);
			if (own) {
				if (e) record_error(st, e, (long long)c);
				sp[0] = smax(sp[0], s_face);
				zero[0] += s_face <= 0 ? 1u : 0u;
				emf[c] = fx[B1];
				emf[n + c] = fx[B2];
			}
#pragma unroll
			for (int q = 0; q < 8; q++) acc[q] = Uc[q];
			add_terms<double>(acc, lo.f, hi.f, len[0], il[0]);
			/* z-face between this plane and the next, owned by the next */
			{
				double U[8];
				to_frame<2>(Un, U);
				hlld_part<T>(U, g.ph, zn);
			}
			const int ez = fused_face<T, 2>(zc, zn, g.ph, dt, fz_hi, s_face);
			if (own && k + 1 < k1) {
				const int cn = cell(i, j, kn);
				if (ez) record_error(st, ez, 2ll * n + cn);
				sp[2] = smax(sp[2], s_face);
				zero[2] += s_face <= 0 ? 1u : 0u;
				emf[4 * n + cn] = fz_hi[B1];
				emf[5 * n + cn] = fz_hi[B0];
			}
		}
		__syncthreads();
		if (w >= 1) {
			double fyv[8];
			double s_face;
			const int e = fused_face<T, 1>(sh.yside[w - 1][lane], sh.yside[w][lane], g.ph, dt, fyv,
				s_face);
#pragma unroll
			for (int q = 0; q < 8; q++) sh.fy[w][lane].f[q] = fyv[q];
			if (own) {
				const int c = cell(i, j, k);
				if (e) record_error(st, e, (long long)n + c);
				sp[1] = smax(sp[1], s_face);
				zero[1] += s_face <= 0 ? 1u : 0u;
				emf[2 * n + c] = fyv[B2];
				emf[3 * n + c] = fyv[B0];
			}
		}
		__syncthreads();
		if (inner) {
			add_terms<double>(acc, sh.fy[w][lane].f, sh.fy[w + 1][lane].f, len[1], il[1]);
			add_terms<double>(acc, fz_lo, fz_hi, len[2], il[2]);
			if (own) {
				const int c = cell(i, j, k);
				out[c] = acc[RHO];
				out[n + c] = acc[M0];
				out[2 * n + c] = acc[M1];
				out[3 * n + c] = acc[M2];
				/* the thermal pressure the first energy refit recovers
				   (update_vol_B before the face update), or the energy
				   itself when no refit will use it */
				out[4 * n + c] = (refits > 0 && acc[RHO] > 0)
					? pressure64(acc, g.ph.gm1, g.ph.half_inv_mu0) : acc[NRJ];
			}
#pragma unroll
			for (int q = 0; q < 8; q++) fz_lo[q] = fz_hi[q];
			zc = zn;
		}
#pragma unroll
		for (int q = 0; q < 8; q++) Uc[q] = Un[q];
	}
	for (int d = 0; d < 3; d++) reduce_speed_blk<NW>(st, d, sp[d], zero[d], sh.ms, sh.mz);
}

/* constrained-transport edge terms from the 6 field components per cell of
   the fused pass (same values as EdgeTerms of the flux array) */
struct EdgeTerms6 {
	const double* P;
	long long n;
	double dx, dy, dz;
	__device__ __forceinline__ double ezx(long long c) const { return +(P[c] * dz / 4); }
	__device__ __forceinline__ double eyx(long long c) const { return -(P[n + c] * dy / 4); }
	__device__ __forceinline__ double exy(long long c) const { return +(P[2 * n + c] * dx / 4); }
	__device__ __forceinline__ double ezy(long long c) const { return -(P[3 * n + c] * dz / 4); }
	__device__ __forceinline__ double exz(long long c) const { return -(P[4 * n + c] * dx / 4); }
	__device__ __forceinline__ double eyz(long long c) const { return +(P[5 * n + c] * dy / 4); }
};

/* total energy from the thermal pressure at the new cell-centred field,
   the second half of refit_energy */
__device__ __forceinline__ double energy_at_pressure(double rho, const double m[3], double p,
	const double vb_new[3], const Phys& ph)
{
	const double inv = 1 / rho;
	const double v[3] = {m[0] * inv, m[1] * inv, m[2] * inv};
	const double
		kinetic = 0.5 * rho * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]),
		magnetic = ph.half_inv_mu0 * (vb_new[0] * vb_new[0] + vb_new[1] * vb_new[1] + vb_new[2] * vb_new[2]);
	return p / ph.gm1 + kinetic + magnetic;
}

/*
Pass 2 of the fused step: the face-field update of update_column_kernel
(the same edge terms, sums and carries along z) from the fused pass's 6
values per cell, then the energy refits: the first from the thermal
pressure pass 1 stored in the energy slot, the others as refit_energy.
Reads the old face field from s, rho, m and that pressure from out, and
writes the new face field and energy into out.
*/
template <int TX, int TY>
__global__ __launch_bounds__(TX * TY) void ct_refit_kernel(GridDev g,
	const double* __restrict__ s, double* __restrict__ out,
	const double* __restrict__ P, int refits, int kz, unsigned gx, unsigned gy, int remap)
{
	const unsigned lb = remap ? logical_block(blockIdx.x, gridDim.x) : blockIdx.x;
	const unsigned bx = lb % gx, by = (lb / gx) % gy, bz = lb / (gx * gy);
	const int i = bx * TX + threadIdx.x;
	const int j = by * TY + threadIdx.y;
	if (i >= g.nx || j >= g.ny) return;
	const long long n = g.n;
	const int im = wrap_dec(i, g.nx), ip = wrap_inc(i, g.nx);
	const int jm = wrap_dec(j, g.ny), jp = wrap_inc(j, g.ny);
	const Nb L{g.nx, g.ny};
	const EdgeTerms6 E{P, n, g.dx, g.dy, g.dz};
	const double axy = g.dx * g.dy, axz = g.dx * g.dz, ayz = g.dy * g.dz;
	const int k0 = (int)bz * kz;
	const int k1 = min(k0 + kz, g.nz);
	if (k0 >= k1) return;

	const int kb = wrap_dec(k0, g.nz);
	double eyx_m = E.eyx(L(i, j, kb)), eyx_pm = E.eyx(L(ip, j, kb));
	double exy_m = E.exy(L(i, j, kb)), exy_pm = E.exy(L(i, jp, kb));
	double eyx_c = E.eyx(L(i, j, k0)), eyx_pc = E.eyx(L(ip, j, k0));
	double exy_c = E.exy(L(i, j, k0)), exy_pc = E.exy(L(i, jp, k0));
	double eyz_c = E.eyz(L(i, j, k0)), eyz_mc = E.eyz(L(im, j, k0)), eyz_pc = E.eyz(L(ip, j, k0));
	double exz_c = E.exz(L(i, j, k0)), exz_mc = E.exz(L(i, jm, k0)), exz_pc = E.exz(L(i, jp, k0));
	const long long cb = L(i, j, k0);
	double s7 = s[7 * n + cb];
	double bz_new = s7 + dBz_sum(eyz_pc, eyz_mc, exz_pc, exz_mc, eyx_pc, eyx_pm, eyx_c, eyx_m,
		exy_pc, exy_pm, exy_c, exy_m) / axy;

	for (int k = k0; k < k1; k++) {
		const int kp = wrap_inc(k, g.nz);
		const long long c = L(i, j, k), cxp = L(ip, j, k), cyp = L(i, jp, k), czp = L(i, j, kp);
		const double
			ezx_c = E.ezx(c), ezx_jm = E.ezx(L(i, jm, k)), ezx_jp = E.ezx(cyp),
			ezx_p = E.ezx(cxp), ezx_pjm = E.ezx(L(ip, jm, k)), ezx_pjp = E.ezx(L(ip, jp, k)),
			ezy_c = E.ezy(c), ezy_m = E.ezy(L(im, j, k)), ezy_p = E.ezy(cxp),
			ezy_jp = E.ezy(cyp), ezy_mjp = E.ezy(L(im, jp, k)), ezy_pjp = E.ezy(L(ip, jp, k));
		const double
			eyx_n = E.eyx(czp), eyx_pn = E.eyx(L(ip, j, kp)),
			exy_n = E.exy(czp), exy_pn = E.exy(L(i, jp, kp)),
			eyz_n = E.eyz(czp), eyz_mn = E.eyz(L(im, j, kp)), eyz_pn = E.eyz(L(ip, j, kp)),
			exz_n = E.exz(czp), exz_mn = E.exz(L(i, jm, kp)), exz_pn = E.exz(L(i, jp, kp));

		const double s7n = s[7 * n + czp];
		const double bx_new = s[5 * n + c] + dBx_sum(ezx_jp, ezx_jm, eyx_n, eyx_m, ezy_jp, ezy_mjp,
			ezy_c, ezy_m, eyz_n, eyz_mn, eyz_c, eyz_mc) / ayz;
		const double by_new = s[6 * n + c] + dBy_sum(exy_n, exy_m, ezy_p, ezy_m, exz_n, exz_mn,
			exz_c, exz_mc, ezx_p, ezx_c, ezx_pjm, ezx_jm) / axz;
		const double bxp_new = s[5 * n + cxp] + dBx_sum(ezx_pjp, ezx_pjm, eyx_pn, eyx_pm, ezy_pjp,
			ezy_jp, ezy_p, ezy_c, eyz_pn, eyz_n, eyz_pc, eyz_c) / ayz;
		const double byp_new = s[6 * n + cyp] + dBy_sum(exy_pn, exy_pm, ezy_pjp, ezy_mjp, exz_pn,
			exz_n, exz_pc, exz_c, ezx_pjp, ezx_jp, ezx_p, ezx_c) / axz;
		const double bzp_new = s7n + dBz_sum(eyz_pn, eyz_mn, exz_pn, exz_mn, eyx_pn, eyx_pc, eyx_n,
			eyx_c, exy_pn, exy_pc, exy_n, exy_c) / axy;
		out[5 * n + c] = bx_new;
		out[6 * n + c] = by_new;
		out[7 * n + c] = bz_new;

		const double rho = out[c];
		const double m[3] = {out[n + c], out[2 * n + c], out[3 * n + c]};
		double En = out[4 * n + c];
		if (refits > 0 && rho > 0) {
			const double vb_new[3] = {0.5 * (bx_new + bxp_new), 0.5 * (by_new + byp_new),
				0.5 * (bz_new + bzp_new)};
			En = energy_at_pressure(rho, m, En, vb_new, g.ph);
			for (int r = 1; r < refits; r++) refit_energy(rho, m, En, vb_new, vb_new, g.ph);
		}
		out[4 * n + c] = En;

		eyx_m = eyx_c; eyx_pm = eyx_pc; exy_m = exy_c; exy_pm = exy_pc;
		eyx_c = eyx_n; eyx_pc = eyx_pn; exy_c = exy_n; exy_pc = exy_pn;
		eyz_c = eyz_n; eyz_mc = eyz_mn; eyz_pc = eyz_pn;
		exz_c = exz_n; exz_mc = exz_mn; exz_pc = exz_pn;
		bz_new = bzp_new;
		s7 = s7n;
	}
}

/*
Accumulating step (HLLD): the three per-direction marches of the two-pass
path, each finishing its direction's part of update_mhd_state's conserved
update (source/mhd/solve.hpp:1503-1614) instead of storing face fluxes:

  acc_x_kernel      x-faces, a wave per 64-cell x tile walking km rows along
                    y (the left part of lane l is lane l - 1's, the face of
                    lane l + 1 comes one lane down; the tile's left-edge parts
                    and right-edge face fluxes are computed for all km rows
                    before the march, one row per lane); writes
                    out = U + x terms (8 values: rho, m, E, cell-centred B)
  acc_march_kernel  y- (D = 1) and z-faces (D = 2), a thread walking km cells
                    of a line along D with the part of the cell just walked
                    carried to the next face (km + 1 faces for km cells);
                    out += D terms in place; the z pass ends the update: it
                    stores rho, m and the thermal pressure the first energy
                    refit recovers (update_vol_B before the face update)

Each pass stores the two field components of its faces that the
constrained-transport edge terms use (`emf`, 6 values per cell in all), and
ct_refit_kernel then updates the face field and refits the energy.  The
conserved variables are accumulated in the two-pass update's order (x, then
y, then z terms of each variable, each term ((0 + lo) - (0 + hi)) / len or
* (1 / len)), so the results are identical to it.  FP64 carries make_side's
record per cell; the mixed path the delta form's record (XPart).
*/
template <class T> struct APartOf;
template <> struct APartOf<double> { using type = SideState<double>; };
template <> struct APartOf<float> { using type = XPart<float>; };

template <class T> __device__ __forceinline__ void apart_make(const double U[8], const Phys& ph,
	typename APartOf<T>::type& p)
{
	if constexpr (std::is_same_v<T, double>) {
		p.cf = 0;
		p.stage = hlld_side<double>(U, ph, p.s, p.cf, p.code);
	} else {
		hlld_part<float>(U, ph, p);
	}
}

__device__ __forceinline__ SideState<double> apart_up(const SideState<double>& r, const SideState<double>& b)
{
	return lane_up(r, b);
}

__device__ __forceinline__ XPart<float> apart_up(const XPart<float>& r, const XPart<float>& b)
{
	return lane_up_words(r, b);
}

/* dt * F of the face between parts L and R in the global frame, zero on
   error; returns the error and sets the face's signal speed */
template <class T, int D> __device__ __forceinline__ int apart_face(const typename APartOf<T>::type& L,
	const typename APartOf<T>::type& R, const Phys& ph, double dt, double Fg[8], double& speed)
{
	double F[8];
	int e;
	if constexpr (std::is_same_v<T, double>) {
		double s = 0;
		if (L.stage | R.stage) e = first_side_error(L.stage, L.code, R.stage, R.code);
		else e = flux_hlld_face<double>(L.s, R.s, L.cf, R.cf, L.s.u, R.s.u, ph, F, s);
		speed = s;
	} else {
		double s = 0;
		const int sl = L.sc >> 4, sr = R.sc >> 4;
		if (sl | sr) e = first_side_error(sl, L.sc & 15, sr, R.sc & 15);
		else e = flux_hlld_mixed_face(L, R, ph, F, s);
		speed = s;
	}
	if (e) {
#pragma unroll
		for (int q = 0; q < 8; q++) Fg[q] = 0;
		speed = 0;
		return e;
	}
	Fg[RHO] = dt * F[RHO];
	Fg[NRJ] = dt * F[NRJ];
	Fg[M0 + D] = dt * F[M0];
	Fg[M0 + (D + 1) % 3] = dt * F[M1];
	Fg[M0 + (D + 2) % 3] = dt * F[M2];
	Fg[B0 + D] = dt * F[B0];
	Fg[B0 + (D + 1) % 3] = dt * F[B1];
	Fg[B0 + (D + 2) % 3] = dt * F[B2];
	return ERR_OK;
}

/* the two edge-term components of a D-face's flux: x (By, Bz), y (Bz, Bx),
   z (By, Bx), at emf[2 D] and emf[2 D + 1] */
template <int D> __device__ __forceinline__ void store_emf(double* __restrict__ emf, size_t n, int c,
	const double Fg[8])
{
	if constexpr (D == 0) {
		emf[c] = Fg[B1];
		emf[n + c] = Fg[B2];
	} else if constexpr (D == 1) {
		emf[2 * n + c] = Fg[B2];
		emf[3 * n + c] = Fg[B0];
	} else {
		emf[4 * n + c] = Fg[B1];
		emf[5 * n + c] = Fg[B0];
	}
}

constexpr int ACC_KMAX = 32;   /* cells per march of the accumulating x pass, at most */

template <class T, int MINB>
__global__ __launch_bounds__(BX * BY, MINB) void acc_x_kernel(GridDev g, StateArrays S,
	double* __restrict__ out, double* __restrict__ emf, const double* __restrict__ dtp,
	StepStatus* st, unsigned gx, int km, int remap)
{
	static_assert(BX == 64, "a wave is one x tile");
	using P = typename APartOf<T>::type;
	const unsigned lb = (remap & 1) ? logical_block(blockIdx.x, gridDim.x) : blockIdx.x;
	const unsigned gt = (unsigned)((g.nz + BY - 1) / BY);
	const unsigned bx = lb % gx, r = lb / gx;
	const int kk = (int)(r % gt) * BY + threadIdx.y;
	const int chunk = (int)(r / gt);
	const int lane = threadIdx.x;
	const int tid = lane + BX * threadIdx.y;
	const int i0 = (int)bx * BX;
	const int i = (i0 + lane) % g.nx;         /* lanes past the grid wrap */
	const bool own_x = i0 + lane < g.nx;
	const int a0 = chunk * km;
	const int a1 = min(a0 + km, g.ny);
	const size_t n = g.n;
	double sp = 0;
	unsigned zero = 0;
	/* per wave: the tile's left-edge parts and right-edge face fluxes of
	   the march's rows, and each lane's face flux (read by the lane below) */
	__shared__ P edge_l[BY][ACC_KMAX];
	__shared__ Flux8<double> edge_r[BY][ACC_KMAX];
	__shared__ double lo_s[8][BX * BY];
	/* kk, a0 and a1 are the same in every lane of the wave */
	if (kk < g.nz && a0 < a1) {
		const double dt = *dtp;
		const double il = 1 / g.dx;
		if (lane < a1 - a0) {
			const int row = a0 + lane;
			double U[8];
			load_face_state_d<0>(S, g, wrap_dec(i0, g.nx), row, kk, U);
			apart_make<T>(U, g.ph, edge_l[threadIdx.y][lane]);
			if (i0 + BX <= g.nx) {
				P pa, pb;
				load_face_state_d<0>(S, g, i0 + BX - 1, row, kk, U);
				apart_make<T>(U, g.ph, pa);
				load_face_state_d<0>(S, g, (i0 + BX) % g.nx, row, kk, U);
				apart_make<T>(U, g.ph, pb);
				double s_unused;
				apart_face<T, 0>(pa, pb, g.ph, dt, edge_r[threadIdx.y][lane].f, s_unused);
			}
		}
		__builtin_amdgcn_wave_barrier();
		for (int a = a0; a < a1; a++) {
			double s_face;
			int e;
			if constexpr (std::is_same_v<T, double>) {
				/* FP64: the face of the lane above comes down with a DPP
				   wavefront shift (lane 63 takes the tile's right-edge
				   face); every lane of the wave is active here */
				double lo[8], hi[8];
				{
					double U[8];
					load_face_state_d<0>(S, g, i, a, kk, U);
					P R;
					apart_make<T>(U, g.ph, R);
					const P L = apart_up(R, edge_l[threadIdx.y][a - a0]);
					e = apart_face<T, 0>(L, R, g.ph, dt, lo, s_face);
				}
#pragma unroll
				for (int q = 0; q < 8; q++) hi[q] = lane_down_dpp(lo[q], edge_r[threadIdx.y][a - a0].f[q]);
				if (own_x) {
					const int c = i + g.nx * (a + g.ny * kk);
					double U[8];
					if (e) record_error(st, e, (long long)c);
					sp = smax(sp, s_face);
					zero += s_face <= 0 ? 1u : 0u;
					store_emf<0>(emf, n, c, lo);
					load_face_state_d<0>(S, g, i, a, kk, U);
					add_terms<double>(U, lo, hi, g.dx, il);
#pragma unroll
					for (int q = 0; q < 8; q++) out[q * n + c] = U[q];
				}
				continue;
			}
			{
				double U[8];
				load_face_state_d<0>(S, g, i, a, kk, U);
				P R;
				apart_make<T>(U, g.ph, R);
				const P L = apart_up(R, edge_l[threadIdx.y][a - a0]);
				double lo[8];
				e = apart_face<T, 0>(L, R, g.ph, dt, lo, s_face);
#pragma unroll
				for (int q = 0; q < 8; q++) lo_s[q][tid] = lo[q];
			}
			/* the faces of this lane and of the lane above, from LDS (the
			   same wave wrote them: LDS operations of a wave complete in
			   order), and the cell's state again (a cache hit) rather than
			   held in registers across the face */
			__asm__ volatile("" ::: "memory");
			__builtin_amdgcn_wave_barrier();
			if (own_x) {
				const int c = i + g.nx * (a + g.ny * kk);
				double lo[8], hi[8], U[8];
				const bool last = lane == BX - 1;
				const int up = last ? tid : tid + 1;
#pragma unroll
				for (int q = 0; q < 8; q++) {
					lo[q] = lo_s[q][tid];
					hi[q] = last ? edge_r[threadIdx.y][a - a0].f[q] : lo_s[q][up];
				}
				if (e) record_error(st, e, (long long)c);
				sp = smax(sp, s_face);
				zero += s_face <= 0 ? 1u : 0u;
				store_emf<0>(emf, n, c, lo);
				load_face_state_d<0>(S, g, i, a, kk, U);
				add_terms<double>(U, lo, hi, g.dx, il);
#pragma unroll
				for (int q = 0; q < 8; q++) out[q * n + c] = U[q];
			}
			__builtin_amdgcn_wave_barrier();
			__asm__ volatile("" ::: "memory");
		}
	}
	reduce_speed(st, 0, sp, zero);
}

template <class T, int D, int MINB>
__global__ __launch_bounds__(BX * BY, MINB) void acc_march_kernel(GridDev g, StateArrays S,
	double* __restrict__ out, double* __restrict__ emf, const double* __restrict__ dtp,
	StepStatus* st, unsigned gx, int km, int remap, int refits)
{
	static_assert(D == 1 || D == 2, "march along y or z");
	using P = typename APartOf<T>::type;
	const unsigned lb = (remap & 1) ? logical_block(blockIdx.x, gridDim.x) : blockIdx.x;
	const int nt = D == 1 ? g.nz : g.ny;                 /* lines across */
	const unsigned gt = (unsigned)((nt + BY - 1) / BY);
	const unsigned bx = lb % gx, r = lb / gx;
	const int t = (int)(r % gt) * BY + threadIdx.y;
	const int chunk = (int)(r / gt);
	const int i = bx * BX + threadIdx.x;
	const int nd = D == 1 ? g.ny : g.nz;
	const int a0 = chunk * km;
	const int a1 = min(a0 + km, nd);
	const size_t n = g.n;
	double sp = 0;
	unsigned zero = 0;
	if (i < g.nx && t < nt && a0 < a1) {
		const int jj = D == 1 ? 0 : t;
		const int kk = D == 1 ? t : 0;
		const double dt = *dtp;
		const double len = D == 1 ? g.dy : g.dz;
		const double il = 1 / len;
		auto cell = [&](int a) { return D == 1 ? i + g.nx * (a + g.ny * kk) : i + g.nx * (jj + g.ny * a); };
		/* the flux of the cell's -D face, kept in LDS across the next face */
		__shared__ double flo_s[8][BX * BY];
		const int tid = threadIdx.x + BX * threadIdx.y;
		P pl, pc;
		{
			double U[8];
			const int am = wrap_dec(a0, nd);
			load_face_state_d<D>(S, g, i, D == 1 ? am : jj, D == 1 ? kk : am, U);
			apart_make<T>(U, g.ph, pl);
			load_face_state_d<D>(S, g, i, D == 1 ? a0 : jj, D == 1 ? kk : a0, U);
			apart_make<T>(U, g.ph, pc);
		}
		double flo[8];
		{
			double s_face;
			const int e = apart_face<T, D>(pl, pc, g.ph, dt, flo, s_face);
			const int c = cell(a0);
			if (e) record_error(st, e, (long long)D * g.n + c);
			sp = smax(sp, s_face);
			zero += s_face <= 0 ? 1u : 0u;
			store_emf<D>(emf, n, c, flo);
#pragma unroll
			for (int q = 0; q < 8; q++) flo_s[q][tid] = flo[q];
		}
		for (int a = a0; a < a1; a++) {
			const int an = wrap_inc(a, nd);
			P pn;
			{
				double U[8];
				load_face_state_d<D>(S, g, i, D == 1 ? an : jj, D == 1 ? kk : an, U);
				apart_make<T>(U, g.ph, pn);
			}
			double fhi[8];
			double s_face;
			const int e = apart_face<T, D>(pc, pn, g.ph, dt, fhi, s_face);
			if (a + 1 < a1) {
				const int cn = cell(a + 1);
				if (e) record_error(st, e, (long long)D * g.n + cn);
				sp = smax(sp, s_face);
				zero += s_face <= 0 ? 1u : 0u;
				store_emf<D>(emf, n, cn, fhi);
			}
			const int c = cell(a);
			double acc[8], lo[8];
			/* a real LDS round trip (the compiler would otherwise keep the
			   carried flux in registers across the face) */
			__asm__ volatile("" ::: "memory");
#pragma unroll
			for (int q = 0; q < 8; q++) {
				acc[q] = out[q * n + c];
				lo[q] = flo_s[q][tid];
			}
			add_terms<double>(acc, lo, fhi, len, il);
			if constexpr (D == 1) {
#pragma unroll
				for (int q = 0; q < 8; q++) out[q * n + c] = acc[q];
			} else {
#pragma unroll
				for (int q = 0; q < 4; q++) out[q * n + c] = acc[q];
				/* the thermal pressure the first energy refit recovers
				   (update_vol_B before the face update), or the energy
				   itself when no refit will use it */
				out[4 * n + c] = (refits > 0 && acc[RHO] > 0)
					? pressure64(acc, g.ph.gm1, g.ph.half_inv_mu0) : acc[NRJ];
			}
#pragma unroll
			for (int q = 0; q < 8; q++) flo_s[q][tid] = fhi[q];
			pc = pn;
		}
	}
	reduce_speed(st, D, sp, zero);
}

/* lagged CFL step of minimize_timestep + set_minmax_substepping_period with a
   single substep (source/mhd/solve.hpp:1961-2006, substepping.hpp:303-338) */
/* one wave: lane s reduces slot s of each dimension (max of the ordered bit
   patterns, i.e. of the non-negative speeds) and clears it for the next pass */
__global__ __launch_bounds__(NSLOT) void dt_kernel(const StepStatus* __restrict__ prev,
	StepStatus* __restrict__ next, double* __restrict__ dt_out, double cfl, double max_dt,
	double dx, double dy, double dz, int fixed)
{
	const int lane = threadIdx.x;
	unsigned long long m[3];
	for (int d = 0; d < 3; d++) {
		m[d] = prev->smax_bits[d][lane];
		for (int off = NSLOT / 2; off > 0; off >>= 1) {
			const unsigned long long o = __shfl_xor(m[d], off, NSLOT);
			m[d] = m[d] < o ? o : m[d];
		}
		next->smax_bits[d][lane] = 0;
	}
	if (lane != 0) return;
	const double len[3] = {dx, dy, dz};
	double cfl_dt = DBL_MAX;
	for (int d = 0; d < 3 && !fixed; d++) {
		if (prev->zero[d] > 0) {
			cfl_dt = 0;
			break;
		}
		cfl_dt = smin(cfl_dt, cfl * len[d] / __longlong_as_double((long long)m[d]));
	}
	*dt_out = (fixed || max_dt <= cfl_dt) ? max_dt : cfl_dt;
	for (int d = 0; d < 3; d++) next->zero[d] = 0;
	next->err = 0;
	next->err_face = -1;
}

/* energy refit of apply_boundaries for a periodic grid (update_vol_B with
   constant pressure over all cells, source/mhd/boundaries.hpp:1289-1292) */
__global__ __launch_bounds__(BX * BY) void refit_kernel(GridDev g, double* __restrict__ s)
{
	const int i = blockIdx.x * BX + threadIdx.x;
	const int j = blockIdx.y * BY + threadIdx.y;
	const int k = blockIdx.z;
	if (i >= g.nx || j >= g.ny) return;
	const long long n = g.n;
	const long long c = i + (long long)g.nx * (j + (long long)g.ny * k);
	const long long cx = wrap_inc(i, g.nx) + (long long)g.nx * (j + (long long)g.ny * k);
	const long long cy = i + (long long)g.nx * (wrap_inc(j, g.ny) + (long long)g.ny * k);
	const long long cz = i + (long long)g.nx * (j + (long long)g.ny * wrap_inc(k, g.nz));
	const double vb[3] = {
		0.5 * (s[5 * n + c] + s[5 * n + cx]),
		0.5 * (s[6 * n + c] + s[6 * n + cy]),
		0.5 * (s[7 * n + c] + s[7 * n + cz])};
	const double m[3] = {s[n + c], s[2 * n + c], s[3 * n + c]};
	double E = s[4 * n + c];
	refit_energy(s[c], m, E, vb, vb, g.ph);
	s[4 * n + c] = E;
}

/* sum of |div B| per block, source/math/nabla.hpp:64-88 */
__global__ __launch_bounds__(BX * BY) void div_kernel(GridDev g, const double* __restrict__ s,
	double* __restrict__ partial)
{
	const int i = blockIdx.x * BX + threadIdx.x;
	const int j = blockIdx.y * BY + threadIdx.y;
	const int k = blockIdx.z;
	double v = 0;
	if (i < g.nx && j < g.ny) {
		const long long n = g.n;
		const long long c = i + (long long)g.nx * (j + (long long)g.ny * k);
		const long long cx = wrap_inc(i, g.nx) + (long long)g.nx * (j + (long long)g.ny * k);
		const long long cy = i + (long long)g.nx * (wrap_inc(j, g.ny) + (long long)g.ny * k);
		const long long cz = i + (long long)g.nx * (j + (long long)g.ny * wrap_inc(k, g.nz));
		v = fabs((s[5 * n + cx] - s[5 * n + c]) / g.dx
			+ (s[6 * n + cy] - s[6 * n + c]) / g.dy
			+ (s[7 * n + cz] - s[7 * n + c]) / g.dz);
	}
	for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
	__shared__ double w[BX * BY / 64];
	const int tid = threadIdx.x + blockDim.x * threadIdx.y;
	if (tid % 64 == 0) w[tid / 64] = v;
	__syncthreads();
	if (tid == 0) {
		double t = 0;
		for (int q = 0; q < BX * BY / 64; q++) t += w[q];
		partial[blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * (size_t)blockIdx.z)] = t;
	}
}

/* right-hand side of the cleaning Poisson problem: the face divergence of
   each cell, and the squared roundoff scale sum_d 2|B_d|/len_d of its faces */
/* right-hand side of the cleaning's Poisson problem, the face divergence of
   each cell, into the multigrid level's layout; a block takes CLEAN_KZ
   planes of a BX x BY tile and writes the block sums of |div| and of the
   squared divergence scale (2 sum_d |B_d| / h_d)^2 of its cells as a pair */
constexpr int CLEAN_KZ = 8;

__global__ __launch_bounds__(BX * BY) void clean_rhs_kernel(GridDev g, const double* __restrict__ s,
	double* __restrict__ rhs, double* __restrict__ pairs, int packed)
{
	__shared__ double sh[2][BX * BY];
	const int i = blockIdx.x * BX + threadIdx.x;
	const int j = blockIdx.y * BY + threadIdx.y;
	const long long n = g.n;
	double sa = 0, ss = 0;
	if (i < g.nx && j < g.ny) {
		for (int m = 0; m < CLEAN_KZ; m++) {
			const int k = blockIdx.z * CLEAN_KZ + m;
			if (k >= g.nz) break;
			const long long c = i + (long long)g.nx * (j + (long long)g.ny * k);
			const long long cx = wrap_inc(i, g.nx) + (long long)g.nx * (j + (long long)g.ny * k);
			const long long cy = i + (long long)g.nx * (wrap_inc(j, g.ny) + (long long)g.ny * k);
			const long long cz = i + (long long)g.nx * (j + (long long)g.ny * wrap_inc(k, g.nz));
			const double d = (s[5 * n + cx] - s[5 * n + c]) / g.dx
				+ (s[6 * n + cy] - s[6 * n + c]) / g.dy
				+ (s[7 * n + cz] - s[7 * n + c]) / g.dz;
			rhs[mg_index(packed, g.nx, g.ny, (int)(n / 2), i, j, k)] = d;
			const double t = 2 * fabs(s[5 * n + c]) / g.dx + 2 * fabs(s[6 * n + c]) / g.dy
				+ 2 * fabs(s[7 * n + c]) / g.dz;
			sa += fabs(d);
			ss += t * t;
		}
	}
	const int tid = threadIdx.x + BX * threadIdx.y;
	sh[0][tid] = sa;
	sh[1][tid] = ss;
	__syncthreads();
	for (int off = BX * BY / 2; off > 0; off >>= 1) {
		if (tid < off) {
			sh[0][tid] += sh[0][tid + off];
			sh[1][tid] += sh[1][tid + off];
		}
		__syncthreads();
	}
	if (tid == 0) {
		const int b = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
		pairs[2 * b] = sh[0][0];
		pairs[2 * b + 1] = sh[1][0];
	}
}

/* projection B_face -= grad phi (a dimension of length 1 has no gradient)
   and constant-pressure energy refit of each cell, from the old state into
   the new one; the refit needs the new field of the cell's positive faces,
   which is recomputed here from phi */
__global__ __launch_bounds__(BX * BY) void project_kernel(GridDev g, const double* __restrict__ s,
	double* __restrict__ o, const double* __restrict__ phi_m, int packed)
{
	const int i = blockIdx.x * BX + threadIdx.x;
	const int j = blockIdx.y * BY + threadIdx.y;
	const int k = blockIdx.z;
	if (i >= g.nx || j >= g.ny) return;
	const long long n = g.n;
	auto id = [&](int a, int b, int e) { return a + (long long)g.nx * (b + (long long)g.ny * e); };
	/* phi in the multigrid level's layout */
	auto phi_at = [&](int a, int b, int e) { return phi_m[mg_index(packed, g.nx, g.ny, (int)(n / 2), a, b, e)]; };
	const long long c = id(i, j, k);
	const int ip = wrap_inc(i, g.nx), jp = wrap_inc(j, g.ny), kp = wrap_inc(k, g.nz);
	const int im = wrap_dec(i, g.nx), jm = wrap_dec(j, g.ny), km = wrap_dec(k, g.nz);
	const double pc = phi_at(i, j, k);
	double b[3] = {s[5 * n + c], s[6 * n + c], s[7 * n + c]};
	double bp[3] = {s[5 * n + id(ip, j, k)], s[6 * n + id(i, jp, k)], s[7 * n + id(i, j, kp)]};
	const double vb_old[3] = {0.5 * (b[0] + bp[0]), 0.5 * (b[1] + bp[1]), 0.5 * (b[2] + bp[2])};
	if (g.nx > 1) {
		b[0] -= (pc - phi_at(im, j, k)) / g.dx;
		bp[0] -= (phi_at(ip, j, k) - pc) / g.dx;
	}
	if (g.ny > 1) {
		b[1] -= (pc - phi_at(i, jm, k)) / g.dy;
		bp[1] -= (phi_at(i, jp, k) - pc) / g.dy;
	}
	if (g.nz > 1) {
		b[2] -= (pc - phi_at(i, j, km)) / g.dz;
		bp[2] -= (phi_at(i, j, kp) - pc) / g.dz;
	}
	const double vb_new[3] = {0.5 * (b[0] + bp[0]), 0.5 * (b[1] + bp[1]), 0.5 * (b[2] + bp[2])};
	const double rho = s[c];
	const double m[3] = {s[n + c], s[2 * n + c], s[3 * n + c]};
	double E = s[4 * n + c];
	refit_energy(rho, m, E, vb_old, vb_new, g.ph);
	o[c] = rho;
	o[n + c] = m[0];
	o[2 * n + c] = m[1];
	o[3 * n + c] = m[2];
	o[4 * n + c] = E;
	o[5 * n + c] = b[0];
	o[6 * n + c] = b[1];
	o[7 * n + c] = b[2];
}

/* the six transverse magnetic fluxes of the flux array: slot 2 d + q - 1
   (q = 1, 2) holds dt * F of global field component 5 + (d + q) % 3 on the
   d-faces */
__global__ __launch_bounds__(256) void magnetic_fluxes_kernel(const double* __restrict__ flux,
	double* __restrict__ out, long long n)
{
	const long long c = (long long)blockIdx.x * 256 + threadIdx.x;
	if (c >= n) return;
#pragma unroll
	for (int d = 0; d < 3; d++)
#pragma unroll
	for (int q = 1; q < 3; q++) out[(2 * d + q - 1) * n + c] = flux[(d * 8 + 5 + (d + q) % 3) * n + c];
}

/* batched Riemann solver for the pamhd_face_flux_batch unit-test surface */
template <int SOLVER, class T>
__global__ __launch_bounds__(256) void face_batch_kernel(size_t n, const double* __restrict__ Lp,
	const double* __restrict__ Rp, const double* __restrict__ bg, Phys ph,
	double* __restrict__ Fp, double* __restrict__ sp, int* __restrict__ status)
{
	const size_t f = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (f >= n) return;
	double L[8], R[8];
	for (int q = 0; q < 8; q++) {
		L[q] = Lp[q * n + f];
		R[q] = Rp[q * n + f];
	}
	T b[3] = {0, 0, 0};
	if (bg) for (int q = 0; q < 3; q++) b[q] = (T)bg[q * n + f];
	double F[8], s = 0;
	const int e = riemann_d<SOLVER, T>(L, R, b, ph, F, s);
	for (int q = 0; q < 8; q++) Fp[q * n + f] = e ? 0.0 : F[q];
	sp[f] = e ? 0.0 : s;
	if (status) status[f] = e;
}

/* ---------------------------------------------------------------------- */
/* host side                                                              */
/* ---------------------------------------------------------------------- */

std::mutex g_err_mutex;
std::string g_err;

void set_global_error(const std::string& m)
{
	std::lock_guard<std::mutex> lock(g_err_mutex);
	g_err = m;
}

const char* error_name(int e)
{
	switch (e) {
	case PAMHD_E_DENSITY: return "non-positive or invalid mass density";
	case PAMHD_E_PRESSURE: return "non-positive or invalid thermal pressure";
	case PAMHD_E_SPEED: return "invalid signal speed";
	case PAMHD_E_FLUX: return "non-finite momentum or flux";
	case PAMHD_E_FIELD: return "non-finite total magnetic field";
	case PAMHD_E_ROE_DENSITY: return "non-physical density in Roe intermediate state";
	case PAMHD_E_ROE_PRESSURE: return "non-physical pressure in Roe intermediate state";
	case PAMHD_E_SOLVER: return "invalid solver";
	case PAMHD_E_SUBSTEP: return "substep-min above the largest allowed substep level, or a level above 30";
	default: return "error";
	}
}

/*
Per-cell 2^N substepping of timestep(), source/mhd/solve.hpp:1837-1924 with
source/substepping.hpp, on a periodic grid of normal cells (FP64).  Working
layout: S 14 x n (rho, m, E, vol B, the cell's copies of its -x, +x, -y, +y,
-z, +z faces), per-face flux sums 48 x n ((d * 8 + q): the sum serving the
slower of the face's two cells, or both; ((3 + d) * 8 + q): the faster
cell's where the periods differ), face dB sums 9 x n (the same two sums
per face, and the +d faces at a closed end), per-face signal
speeds 3 x n, per-cell time step, level exponent, period and the smallest
period of the cell's face and edge neighbourhood.  The state the handle
keeps is the face-once layout; a substepped step expands it, runs the
substeps, syncs the face copies and writes it back.  A face copy's change
sums the same 12 edge terms in the same order as update_cell_kernel; the
sums a cell applies equal its flux slots and face dB of the reference.
*/
struct SubstepDev {
	double* S = nullptr;
	double* slot = nullptr;
	double* fdb = nullptr;
	double* maxv = nullptr;
	double* ts = nullptr;
	int* e[2] = {nullptr, nullptr};
	int* period = nullptr;
	int* minsub = nullptr;
	/* per-cell options: [0, n) substep-min shifted so that its smallest is 0,
	   [n, 2n) substep-max; null when every cell has substep-min 0 and
	   substep-max smax */
	int* opt = nullptr;
	int* opt_dev = nullptr;
	/* [0] smallest time step, [1] largest period, [2] substep length
	   (smallest time step / 2^substep-min), [3] largest level; the
	   non-negative doubles are kept as their bits */
	unsigned long long* scal = nullptr;
	double* sub_dt = nullptr;
	StepStatus* status = nullptr;
	unsigned long long* host = nullptr;   /* pinned: [0] largest period, [1] substep length (bits), [2] largest level */
	StepStatus* host_status = nullptr;
	int smax = 0;   /* uniform substep-max, or the largest per-cell one */
	int last_max = 1;
	/* grids with boundaries: S is the state itself; cell types (1 normal,
	   0 boundary, -1 not solved), face types of the six faces, the
	   variables' boundary kinds (1 normal, 2 not solved, 3 value, 4 copy),
	   periodicity, and the value boundaries' values (mask 4 n; number
	   density n, velocity 3 n, pressure n, face field 6 n) */
	bool bounded = false;
	const signed char* ct = nullptr;
	const signed char* fi = nullptr;
	const signed char* vt = nullptr;
	int per[3] = {1, 1, 1};
	unsigned char* bmask = nullptr;
	double* bval = nullptr;
	signed char* types = nullptr;   /* allocation behind ct, fi, vt */
	/* per BX x BY block of the cell grid (nblk of them): [0, 3 nblk) the
	   smallest owner period m of the block's x-, y-, z-faces (a block has a
	   face due at substep s exactly when s % m == 0), [3 nblk, 4 nblk) the
	   smallest period of its cells and their lower neighbours; and
	   whether a block's last flux pass of a direction left nonzero edge
	   fields (3 nblk) */
	int* blk = nullptr;
	int* dirty = nullptr;
	int nblk = 0;
	/* per substep, the blocks each pass works on (ss_lists_kernel): lists
	   SS_LISTS x nblk (the x-, y-, z-flux passes, the face-dB pass, the
	   update, the blocks of each direction whose edge fields only need
	   zeroing, and the face-dB pass's blocks whose dBs are all +0), and
	   their lengths, two sets of SS_LISTS used by odd and
	   even substeps */
	int* list = nullptr;
	int* count = nullptr;
	int walk = 0;   /* workgroups of the passes that walk a list */
};

/* lists of ss_lists_kernel; the grid of the passes that walk them */
enum { SS_LIST_DB = 3, SS_LIST_UPDATE = 4, SS_LIST_ZERO = 5, SS_LIST_DB0 = 8, SS_LISTS = 9 };
constexpr int SS_WALK_BLOCKS = 4096;

/* index of a neighbour coordinate: periodic dimensions wrap (within one
   length), closed ones give -1 beyond their ends */
__device__ __forceinline__ int ss_coord(int v, int len, int per)
{
	if (v < 0) return per ? v + len : -1;
	if (v >= len) return per ? v - len : -1;
	return v;
}

__device__ __forceinline__ bool ss_active(const SubstepDev& d, long long c)
{
	return !d.ct || d.ct[c] >= 0;
}

__device__ __forceinline__ int cidx(const GridDev& g, int i, int j, int k)
{
	return i + g.nx * (j + g.ny * k);
}

/* the step's five reductions (0: smallest cell time step, 1: largest
   period, 2: smallest time step / 2^substep-min, 3: largest level, 4: faces
   solved) are spread over SS_SLOTS slots each, a wave's atomic going to the
   slot of its block (a few addresses serialise the 2^18 waves' atomics of
   a 256^3 grid), and folded by ss_fold_kernel / ss_count_kernel before
   they are read */
constexpr int SS_SLOTS = 4096;
constexpr int SS_FOLD = 256;
__device__ __forceinline__ unsigned ss_slot()
{
	return (blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z)) % SS_SLOTS;
}

__global__ __launch_bounds__(SS_FOLD) void ss_clear_kernel(SubstepDev d)
{
	const unsigned long long big = (unsigned long long)__double_as_longlong(DBL_MAX);
	for (int s = threadIdx.x; s < SS_SLOTS; s += blockDim.x) {
		d.scal[0 * SS_SLOTS + s] = big;
		d.scal[1 * SS_SLOTS + s] = 0;
		d.scal[2 * SS_SLOTS + s] = big;
		d.scal[3 * SS_SLOTS + s] = 0;
	}
	for (int s = threadIdx.x; s < SS_SLOTS; s += blockDim.x) d.scal[4 * SS_SLOTS + s] = 0;
	if (threadIdx.x == 0) {
		d.status->err = 0;
		d.status->err_face = -1;
		d.status->flux_calcs = 0;
	}
}

/* the faces solved over the step: the fifth reduction's slots summed into
   the status record (one block of SS_FOLD threads) */
__global__ __launch_bounds__(SS_FOLD) void ss_count_kernel(SubstepDev d)
{
	__shared__ unsigned long long w[SS_FOLD / 64];
	unsigned long long v = 0;
	for (int s = threadIdx.x; s < SS_SLOTS; s += SS_FOLD) v += d.scal[4 * SS_SLOTS + s];
	for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
	if (threadIdx.x % 64 == 0) w[threadIdx.x / 64] = v;
	__syncthreads();
	if (threadIdx.x == 0) {
		for (int k = 1; k < SS_FOLD / 64; k++) v += w[k];
		d.status->flux_calcs = v;
	}
}

/* reduction q's SS_SLOTS slots into slot 0 (one block of SS_FOLD threads):
   min when is_min (the slots hold non-negative doubles' bit patterns, which
   order as the values do), max otherwise */
__global__ __launch_bounds__(SS_FOLD) void ss_fold_kernel(SubstepDev d, int q, int is_min)
{
	__shared__ unsigned long long w[SS_FOLD / 64];
	const unsigned long long* a = d.scal + (size_t)q * SS_SLOTS;
	unsigned long long v = a[threadIdx.x];
	for (int s = threadIdx.x + SS_FOLD; s < SS_SLOTS; s += SS_FOLD) {
		const unsigned long long o = a[s];
		v = is_min ? (o < v ? o : v) : (o > v ? o : v);
	}
	for (int off = 32; off > 0; off >>= 1) {
		const unsigned long long o = __shfl_xor(v, off, 64);
		v = is_min ? (o < v ? o : v) : (o > v ? o : v);
	}
	if (threadIdx.x % 64 == 0) w[threadIdx.x / 64] = v;
	__syncthreads();
	if (threadIdx.x == 0) {
		for (int k = 1; k < SS_FOLD / 64; k++) v = is_min ? (w[k] < v ? w[k] : v) : (w[k] > v ? w[k] : v);
		d.scal[(size_t)q * SS_SLOTS] = v;
	}
}

/* minimize_timestep, source/mhd/solve.hpp:1927-2014: a cell's limit over its
   six faces, the smallest over the grid and the smallest of time step /
   2^substep-min (set_minmax_substepping_period, substepping.hpp:333-342) */
__global__ __launch_bounds__(BX * BY) void ss_ts_kernel(GridDev g, SubstepDev d, double cfl)
{
	const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y, k = blockIdx.z;
	double t = DBL_MAX, r = DBL_MAX;
	if (i < g.nx && j < g.ny) {
		const int c = cidx(g, i, j, k);
		if (ss_active(d, c)) {
			const double len[3] = {g.dx, g.dy, g.dz};
			const int id[3] = {i, j, k};
			const int N[3] = {g.nx, g.ny, g.nz};
			for (int q = 0; q < 3; q++) {
				/* the faces on the cell's two sides, where the neighbour
				   exists and is of type >= 0 */
				int lo[3] = {i, j, k}, hi[3] = {i, j, k};
				lo[q] = ss_coord(id[q] - 1, N[q], d.per[q]);
				hi[q] = ss_coord(id[q] + 1, N[q], d.per[q]);
				const int bl = lo[q] < 0 ? -1 : cidx(g, lo[0], lo[1], lo[2]);
				const int bh = hi[q] < 0 ? -1 : cidx(g, hi[0], hi[1], hi[2]);
				if (bl >= 0 && ss_active(d, bl)) {
					const double v = d.maxv[(long long)q * g.n + c];
					t = v <= 0 ? 0.0 : smin(cfl * len[q] / v, t);
				}
				if (bh >= 0 && ss_active(d, bh)) {
					const double v = d.maxv[(long long)q * g.n + bh];
					t = v <= 0 ? 0.0 : smin(cfl * len[q] / v, t);
				}
			}
			r = d.opt ? t / (double)(1 << d.opt[c]) : t;
		}
		d.ts[c] = t;
	}
	for (int off = 32; off > 0; off >>= 1) {
		t = smin(t, __shfl_xor(t, off, 64));
		r = smin(r, __shfl_xor(r, off, 64));
	}
	if ((threadIdx.x + BX * threadIdx.y) % 64 == 0) {
		atomicMin(&d.scal[0 * SS_SLOTS + ss_slot()], (unsigned long long)__double_as_longlong(t));
		atomicMin(&d.scal[2 * SS_SLOTS + ss_slot()], (unsigned long long)__double_as_longlong(r));
	}
}

/* set_minmax_substepping_period, source/substepping.hpp:264-399: the
   substep length and each cell's largest level.  floor(log2 x) is exact
   (ilogb); a non-finite ratio gives level 0.  A cell whose substep-min
   exceeds its level (the reference aborts) or whose level exceeds 30 (its
   int period would overflow) reports PAMHD_E_SUBSTEP. */
__global__ __launch_bounds__(BX * BY) void ss_level_kernel(GridDev g, SubstepDev d, double max_dt)
{
	const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y, k = blockIdx.z;
	const double smallest = __longlong_as_double((long long)d.scal[0]);
	const double ret = __longlong_as_double((long long)d.scal[2 * SS_SLOTS]);
	const bool all_zero = max_dt <= smallest;
	if (i == 0 && j == 0 && k == 0) *d.sub_dt = all_zero ? max_dt : ret;
	int lev = 0;
	if (i < g.nx && j < g.ny) {
		const int c = cidx(g, i, j, k);
		const double t = d.ts[c];
		if (ss_active(d, c) && !all_zero && t > 0) {
			const double x = smin(t, max_dt) / ret;
			int step = (__builtin_isfinite(x) && x > 0) ? ilogb(x) : 0;
			if (step < 0) step = 0;
			const int cap = d.opt ? d.opt[g.n + c] : d.smax;
			lev = cap > step ? step : cap;
			if ((d.opt && d.opt[c] > lev) || lev > 30) {
				record_error(d.status, PAMHD_E_SUBSTEP, c);
				lev = 0;
			}
		}
		d.e[0][c] = lev;
	}
	for (int off = 32; off > 0; off >>= 1) {
		const int o = __shfl_xor(lev, off, 64);
		lev = o > lev ? o : lev;
	}
	if ((threadIdx.x + BX * threadIdx.y) % 64 == 0)
		atomicMax(&d.scal[3 * SS_SLOTS + ss_slot()], (unsigned long long)lev);
}

/* the levels of the cells around this BX x BY block of plane k (planes
   k - 1 .. k + 1, one cell of halo) in LDS, SS_NONE where there is no cell
   (beyond a closed end), the cell is not solved, or no thread of the block
   reads the entry */
constexpr int SS_NONE = INT_MAX / 2;
struct LevelTile {
	int v[3][BY + 2][BX + 2];
};

__device__ __forceinline__ void load_level_tile(const GridDev& g, const SubstepDev& d,
	const int* __restrict__ e, LevelTile& T)
{
	const int k = blockIdx.z;
	const int i0 = (int)blockIdx.x * BX - 1, j0 = (int)blockIdx.y * BY - 1;
	const int ihi = min(g.nx - 1, (int)blockIdx.x * BX + BX - 1) + 1;
	const int jhi = min(g.ny - 1, (int)blockIdx.y * BY + BY - 1) + 1;
	const int tid = threadIdx.x + BX * threadIdx.y;
	for (int t = tid; t < 3 * (BY + 2) * (BX + 2); t += BX * BY) {
		const int ti = t % (BX + 2), tj = (t / (BX + 2)) % (BY + 2), tk = t / ((BX + 2) * (BY + 2));
		const int vi = i0 + ti, vj = j0 + tj;
		int val = SS_NONE;
		if (vi <= ihi && vj <= jhi) {
			const int a = ss_coord(vi, g.nx, d.per[0]), b = ss_coord(vj, g.ny, d.per[1]);
			const int c = ss_coord(k + tk - 1, g.nz, d.per[2]);
			if (a >= 0 && b >= 0 && c >= 0) {
				const int id = cidx(g, a, b, c);
				if (ss_active(d, id)) val = e[id];
			}
		}
		T.v[tk][tj][ti] = val;
	}
	__syncthreads();
}

/* f(value) for each of the 18 face and edge neighbours of the thread's cell
   in the tile (offsets with one or two nonzero coordinates), SS_NONE for
   those that are no cell or not solved */
template <class F> __device__ __forceinline__ void for_tile_neighbours(const LevelTile& T, F&& f)
{
	const int ti = threadIdx.x + 1, tj = threadIdx.y + 1;
#pragma unroll
	for (int dk = -1; dk <= 1; dk++)
#pragma unroll
	for (int dj = -1; dj <= 1; dj++)
#pragma unroll
	for (int di = -1; di <= 1; di++) {
		const int nz_ = (di != 0) + (dj != 0) + (dk != 0);
		if (nz_ == 0 || nz_ == 3) continue;
		f(T.v[1 + dk][tj + dj][ti + di]);
	}
}

/* one Jacobi pass of restrict_substepping_period, substepping.hpp:99-191
   (a minimum over the neighbours, so their order does not matter) */
__global__ __launch_bounds__(BX * BY) void ss_restrict_kernel(GridDev g, SubstepDev d,
	const int* __restrict__ in, int* __restrict__ out)
{
	__shared__ LevelTile T;
	load_level_tile(g, d, in, T);
	const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y, k = blockIdx.z;
	if (i >= g.nx || j >= g.ny) return;
	const int c = cidx(g, i, j, k);
	int v = in[c];
	if (ss_active(d, c)) {
		for_tile_neighbours(T, [&](int w) { if (v > w + 1) v = w + 1; });
	}
	out[c] = v;
}

/* update_substeps (2^N periods), substepping.hpp:54-96, and the smallest
   period of each cell's face and edge neighbourhood (get_fluxes,
   source/mhd/solve.hpp:252-289) */
__global__ __launch_bounds__(BX * BY) void ss_period_kernel(GridDev g, SubstepDev d, const int* __restrict__ e)
{
	__shared__ LevelTile T;
	load_level_tile(g, d, e, T);
	const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y, k = blockIdx.z;
	unsigned long long p = 0;
	if (i < g.nx && j < g.ny) {
		const int c = cidx(g, i, j, k);
		int lo = e[c];
		const bool act = ss_active(d, c);
		if (act) for_tile_neighbours(T, [&](int w) { if (w < lo) lo = w; });
		d.period[c] = 1 << e[c];
		d.minsub[c] = 1 << lo;
		p = act ? 1ull << e[c] : 0ull;
	}
	for (int off = 32; off > 0; off >>= 1) {
		const unsigned long long o = __shfl_xor(p, off, 64);
		p = o > p ? o : p;
	}
	if ((threadIdx.x + BX * threadIdx.y) % 64 == 0) atomicMax(&d.scal[1 * SS_SLOTS + ss_slot()], p);
}

/* the working layout from the face-once state */
__global__ __launch_bounds__(BX * BY) void ss_expand_kernel(GridDev g, const double* __restrict__ s, SubstepDev d)
{
	const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y, k = blockIdx.z;
	if (i >= g.nx || j >= g.ny) return;
	const long long n = g.n;
	const int c = cidx(g, i, j, k);
	const int nb[3] = {cidx(g, wrap_inc(i, g.nx), j, k), cidx(g, i, wrap_inc(j, g.ny), k),
		cidx(g, i, j, wrap_inc(k, g.nz))};
	for (int q = 0; q < 5; q++) d.S[q * n + c] = s[q * n + c];
	for (int q = 0; q < 3; q++) {
		d.S[(5 + q) * n + c] = 0.5 * (s[(5 + q) * n + c] + s[(5 + q) * n + nb[q]]);
		d.S[(8 + 2 * q) * n + c] = s[(5 + q) * n + c];
		d.S[(9 + 2 * q) * n + c] = s[(5 + q) * n + nb[q]];
	}
}

/* per block of the cell grid: the smallest owner period of its x-, y- and
   z-faces and the smallest period of its cells and their lower neighbours
   (the tests of the block lists, ss_lists_kernel), and
   every block's edge fields marked as not yet known to be zero */
__global__ __launch_bounds__(BX * BY) void ss_blockmask_kernel(GridDev g, SubstepDev d)
{
	__shared__ int red[4][BX * BY];
	const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y, k = blockIdx.z;
	const int tid = threadIdx.x + BX * threadIdx.y;
	int m[3] = {INT_MAX, INT_MAX, INT_MAX}, p = INT_MAX;
	if (i < g.nx && j < g.ny) {
		const int c = cidx(g, i, j, k);
		p = d.period[c];
		const int o[3] = {
			ss_coord(i - 1, g.nx, d.per[0]) < 0 ? -1 : cidx(g, ss_coord(i - 1, g.nx, d.per[0]), j, k),
			ss_coord(j - 1, g.ny, d.per[1]) < 0 ? -1 : cidx(g, i, ss_coord(j - 1, g.ny, d.per[1]), k),
			ss_coord(k - 1, g.nz, d.per[2]) < 0 ? -1 : cidx(g, i, j, ss_coord(k - 1, g.nz, d.per[2]))};
		for (int e = 0; e < 3; e++) {
			if (o[e] < 0) continue;
			m[e] = d.minsub[o[e]];
			p = min(p, d.period[o[e]]);
		}
	}
	for (int e = 0; e < 3; e++) red[e][tid] = m[e];
	red[3][tid] = p;
	__syncthreads();
	for (int off = BX * BY / 2; off > 0; off >>= 1) {
		if (tid < off) {
			for (int e = 0; e < 4; e++) red[e][tid] = min(red[e][tid], red[e][tid + off]);
		}
		__syncthreads();
	}
	if (tid < 4) {
		const int b = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
		d.blk[tid * d.nblk + b] = red[tid][0];
		if (tid < 3) d.dirty[tid * d.nblk + b] = 1;
	}
	/* the list lengths of the first substep (odd) start at zero */
	if (blockIdx.x == 0 && blockIdx.y == 0 && blockIdx.z == 0 && tid < SS_LISTS) d.count[SS_LISTS + tid] = 0;
}

/*
The blocks the passes of substep `sub` work on, one thread per block of the
cell grid, appended in any order (the blocks are independent):
  flux pass D   a face of the block due at this substep; a block with
                none whose edge fields of the previous substep are not yet
                known to be zero goes to the zeroing list of D instead; the
                flag is updated here
  face-dB pass  a block around it (27, three directions) with a face due;
                else, a sum of the block that restarts at this substep goes
                to the list of blocks whose dBs are all +0 (else no sum
                changes)
  update        a cell of the block (or a lower neighbour) due: periods are
                powers of two, so none is when the smallest is not
Thread 0 also zeroes the other set of lengths, which the next substep uses
(the passes of the previous one, which read it, have finished).
*/
__global__ __launch_bounds__(256) void ss_lists_kernel(GridDev g, SubstepDev d, int sub)
{
	const int b = blockIdx.x * 256 + threadIdx.x;
	int* cnt = d.count + SS_LISTS * (sub & 1);
	if (b == 0) {
		for (int q = 0; q < SS_LISTS; q++) d.count[SS_LISTS * ((sub + 1) & 1) + q] = 0;
	}
	bool run[SS_LISTS] = {false, false, false, false, false, false, false, false, false};
	if (b < d.nblk) {
		for (int e = 0; e < 3; e++) {
			const bool due = (sub & (d.blk[e * d.nblk + b] - 1)) == 0;
			run[e] = due;
			run[SS_LIST_ZERO + e] = !due && d.dirty[e * d.nblk + b];
			d.dirty[e * d.nblk + b] = due ? 1 : 0;
		}
		const int gx = (g.nx + BX - 1) / BX, gy = (g.ny + BY - 1) / BY, gz = g.nz;
		const int bx0 = b % gx, by0 = (b / gx) % gy, bz0 = b / (gx * gy);
		const bool restart = ((sub - 1) & (d.blk[3 * d.nblk + b] - 1)) == 0;
		bool near = false;
		for (int dz = -1; dz <= 1 && !near; dz++)
		for (int dy = -1; dy <= 1 && !near; dy++)
		for (int dx = -1; dx <= 1 && !near; dx++) {
			const int bx = (bx0 + dx + gx) % gx, by = (by0 + dy + gy) % gy, bz = (bz0 + dz + gz) % gz;
			const int nbk = bx + gx * (by + gy * bz);
			for (int e = 0; e < 3; e++) {
				if ((sub & (d.blk[e * d.nblk + nbk] - 1)) == 0) near = true;
			}
		}
		run[SS_LIST_DB] = near;
		run[SS_LIST_DB0] = restart && !near;
		run[SS_LIST_UPDATE] = (sub & (d.blk[3 * d.nblk + b] - 1)) == 0;
	}
	const unsigned long long lt = (1ull << (threadIdx.x % 64)) - 1;
	for (int l = 0; l < SS_LISTS; l++) {
		const unsigned long long m = __ballot(run[l]);
		if (!m) continue;
		const int leader = __ffsll((long long)m) - 1;
		int base = 0;
		if ((int)(threadIdx.x % 64) == leader) base = atomicAdd(&cnt[l], __popcll(m));
		base = __shfl(base, leader, 64);
		if (run[l]) d.list[l * d.nblk + base + __popcll(m & lt)] = b;
	}
}

/* block b of the cell grid (cell_grid's order: x fastest, then y, then z) */
__device__ __forceinline__ void ss_block_coords(const GridDev& g, int b, int& bx, int& by, int& bz)
{
	const int gx = (g.nx + BX - 1) / BX, gy = (g.ny + BY - 1) / BY;
	bx = b % gx;
	by = (b / gx) % gy;
	bz = b / (gx * gy);
}

/*
Faces of direction D due at substep `sub` (source/mhd/solve.hpp:252-393 with
substepping): a face between two cells of type >= 0 is solved by its
negative-side owner o at the pace of o's smallest neighbourhood period m,
from the cells' stored volume field, and weighted by the substep length
times m.  Its contribution goes into per-face accumulators instead of the
two cells' flux slots: both cells receive the same contributions, so one
sum serves both when their periods are equal; where they differ, the
faster cell (period pmin) gets a second sum.  A sum restarts (0 + c) at
the first solve after the cell(s) it serves applied their slots: at
substep s with (s - m) % p == 0, m dividing every period around o.  The
slots of the reference hold exactly these sums (0 + c1 + c2 + ..., never
-0, so the zeros the reference's unsolved substeps would add change no
bit).  The face's two field components the edge terms use go to emf (0
when the face is not solved at this substep).
*/
template <int SOLVER, int D>
__device__ __forceinline__ void ss_flux_block(const GridDev& g, const SubstepDev& d,
	const double* __restrict__ bg, double* __restrict__ emf, int sub, int bx, int by, int k)
{
	const int i = bx * BX + threadIdx.x, j = by * BY + threadIdx.y;
	if (i >= g.nx || j >= g.ny) return;
	const long long n = g.n;
	const int c = cidx(g, i, j, k);
	const int oi = D == 0 ? ss_coord(i - 1, g.nx, d.per[0]) : i;
	const int oj = D == 1 ? ss_coord(j - 1, g.ny, d.per[1]) : j;
	const int ok = D == 2 ? ss_coord(k - 1, g.nz, d.per[2]) : k;
	double* A = d.slot + (long long)D * 8 * n;          /* sum of the slower cell */
	double* A2 = d.slot + (long long)(3 + D) * 8 * n;   /* sum of the faster cell */
	double* Em = emf + (long long)(2 * D) * n;
	const int o = (oi < 0 || oj < 0 || ok < 0) ? -1 : cidx(g, oi, oj, ok);
	const bool act = o >= 0 && ss_active(d, o) && ss_active(d, c);
	{
		/* faces solved (timestep()'s flux_calcs), counted per wave into the
		   block's slot of the fifth reduction */
		const bool solved = act && (sub & (d.minsub[o] - 1)) == 0;
		const unsigned long long b = __ballot(solved);
		const unsigned long long a = __ballot(1);
		if ((int)(threadIdx.x + BX * threadIdx.y) % 64 == __ffsll((long long)a) - 1 && b)
			atomicAdd(&d.scal[4 * SS_SLOTS + ss_slot()], (unsigned long long)__popcll(b));
	}
	if (!act) {
		/* no face, or a cell not solved: nothing is ever added */
		for (int q = 0; q < 8; q++) {
			A[q * n + c] = 0;
			A2[q * n + c] = 0;
		}
		Em[c] = 0;
		Em[n + c] = 0;
		return;
	}
	const int m = d.minsub[o];
	if (sub & (m - 1)) {
		Em[c] = 0;
		Em[n + c] = 0;
		return;
	}
	double L[8], R[8];
	const int cc[2] = {o, c};
	double* U[2] = {L, R};
	for (int sd = 0; sd < 2; sd++) {
		const int x = cc[sd];
		const double mm[3] = {d.S[1 * n + x], d.S[2 * n + x], d.S[3 * n + x]};
		const double B[3] = {d.S[5 * n + x], d.S[6 * n + x], d.S[7 * n + x]};
		U[sd][RHO] = d.S[x];
		U[sd][NRJ] = d.S[4 * n + x];
		U[sd][M0] = mm[D]; U[sd][M1] = mm[(D + 1) % 3]; U[sd][M2] = mm[(D + 2) % 3];
		U[sd][B0] = B[D]; U[sd][B1] = B[(D + 1) % 3]; U[sd][B2] = B[(D + 2) % 3];
	}
	double b[3] = {0, 0, 0};
	if (bg) {
		const double* bgd = bg + (size_t)(3 * D) * n;
		b[0] = bgd[(size_t)D * n + o];
		b[1] = bgd[(size_t)((D + 1) % 3) * n + o];
		b[2] = bgd[(size_t)((D + 2) % 3) * n + o];
	}
	double F[8], sp = 0;
	const int e = riemann<SOLVER, double>(L, R, b, g.ph, F, sp);
	double Fg[8];
	if (e) {
		/* the step fails; the face contributes nothing */
		record_error(d.status, e, (long long)D * n + c);
		for (int q = 0; q < 8; q++) Fg[q] = 0;
	} else {
		d.maxv[(long long)D * n + c] = sp;
		Fg[RHO] = F[RHO];
		Fg[NRJ] = F[NRJ];
		Fg[M0 + D] = F[M0]; Fg[M0 + (D + 1) % 3] = F[M1]; Fg[M0 + (D + 2) % 3] = F[M2];
		Fg[B0 + D] = F[B0]; Fg[B0 + (D + 1) % 3] = F[B1]; Fg[B0 + (D + 2) % 3] = F[B2];
	}
	const double min_dt = *d.sub_dt * m;
	double cq[8];
	for (int q = 0; q < 8; q++) cq[q] = min_dt * Fg[q];
	const int po = d.period[o], pc = d.period[c];
	const int pmax = po > pc ? po : pc, pmin = po > pc ? pc : po;
	const bool fresh = ((sub - m) & (pmax - 1)) == 0;
	for (int q = 0; q < 8; q++) A[q * n + c] = (fresh ? 0.0 : A[q * n + c]) + cq[q];
	if (pmin < pmax) {
		const bool fresh2 = ((sub - m) & (pmin - 1)) == 0;
		for (int q = 0; q < 8; q++) A2[q * n + c] = (fresh2 ? 0.0 : A2[q * n + c]) + cq[q];
	}
	/* the edge-term components: x-face (By, Bz), y (Bz, Bx), z (By, Bx) */
	Em[c] = cq[D == 1 ? B0 + 2 : B0 + 1];
	Em[n + c] = cq[D == 0 ? B0 + 2 : B0];
}

/* the flux pass of direction D: with WALK over the blocks of its list
   (those with a face due), else over every block of the cell grid (a
   substep at which every face is due) */
template <int SOLVER, int D, bool WALK>
__global__ __launch_bounds__(BX * BY) void ss_flux_kernel(GridDev g, SubstepDev d,
	const double* __restrict__ bg, double* __restrict__ emf, int sub)
{
	if constexpr (WALK) {
		const int cnt = d.count[SS_LISTS * (sub & 1) + D];
		const int* list = d.list + (size_t)D * d.nblk;
		for (int idx = blockIdx.x; idx < cnt; idx += gridDim.x) {
			int bx, by, bz;
			ss_block_coords(g, list[idx], bx, by, bz);
			ss_flux_block<SOLVER, D>(g, d, bg, emf, sub, bx, by, bz);
		}
	} else {
		ss_flux_block<SOLVER, D>(g, d, bg, emf, sub, blockIdx.x, blockIdx.y, blockIdx.z);
	}
}

/*
The flux pass of a substep at which every face is due, for HLLD's y- (D = 1)
and z-faces (D = 2) on a periodic grid without cell types: a thread walks
km faces of a line along D with HLLD's per-cell part (hlld_side) made once
per cell and carried to the next face, as flux_hlld_march_kernel does
(flux_hlld is the two parts, the side error order and flux_hlld_face, so
the fluxes are ss_flux_block's bit for bit); each face's flux then goes
into the substep's sums and edge fields as in ss_flux_block.  Blocks: x
tiles fastest, then BY-line groups, then chunks of km along D.
*/
template <int D>
__global__ __launch_bounds__(BX * BY) void ss_flux_march_kernel(GridDev g, SubstepDev d,
	double* __restrict__ emf, int sub, unsigned gx, int km)
{
	static_assert(D == 1 || D == 2, "march along y or z");
	const int nt = D == 1 ? g.nz : g.ny;                 /* lines across */
	const unsigned gt = (unsigned)((nt + BY - 1) / BY);
	const unsigned bx = blockIdx.x % gx, r = blockIdx.x / gx;
	const int t = (int)(r % gt) * BY + threadIdx.y;
	const int chunk = (int)(r / gt);
	const int i = bx * BX + threadIdx.x;
	const int nd = D == 1 ? g.ny : g.nz;
	const int a0 = chunk * km;
	const int a1 = min(a0 + km, nd);
	const long long n = g.n;
	unsigned solved = 0;
	if (i < g.nx && t < nt && a0 < a1) {
		const int jj = D == 1 ? 0 : t, kk = D == 1 ? t : 0;
		auto cell = [&](int a) { return D == 1 ? cidx(g, i, a, kk) : cidx(g, i, jj, a); };
		auto load = [&](int x, double U[8]) {
			const double mm[3] = {d.S[1 * n + x], d.S[2 * n + x], d.S[3 * n + x]};
			const double B[3] = {d.S[5 * n + x], d.S[6 * n + x], d.S[7 * n + x]};
			U[RHO] = d.S[x];
			U[NRJ] = d.S[4 * n + x];
			U[M0] = mm[D]; U[M1] = mm[(D + 1) % 3]; U[M2] = mm[(D + 2) % 3];
			U[B0] = B[D]; U[B1] = B[(D + 1) % 3]; U[B2] = B[(D + 2) % 3];
		};
		double* A = d.slot + (long long)D * 8 * n;
		double* A2 = d.slot + (long long)(3 + D) * 8 * n;
		double* Em = emf + (long long)(2 * D) * n;
		int o = cell(wrap_dec(a0, nd));
		Side<double> Ls;
		double cfl = 0;
		int code_l, stage_l;
		{
			double U[8];
			load(o, U);
			stage_l = hlld_side<double>(U, g.ph, Ls, cfl, code_l);
		}
		for (int a = a0; a < a1; a++) {
			const int c = cell(a);
			Side<double> Rs;
			double cfr = 0;
			int code_r, stage_r;
			{
				double U[8];
				load(c, U);
				stage_r = hlld_side<double>(U, g.ph, Rs, cfr, code_r);
			}
			double F[8], sT = 0;
			int e;
			if (stage_l | stage_r) e = first_side_error(stage_l, code_l, stage_r, code_r);
			else e = flux_hlld_face<double>(Ls, Rs, cfl, cfr, Ls.u, Rs.u, g.ph, F, sT);
			double Fg[8];
			if (e) {
				record_error(d.status, e, (long long)D * n + c);
				for (int q = 0; q < 8; q++) Fg[q] = 0;
			} else {
				d.maxv[(long long)D * n + c] = sT;
				Fg[RHO] = F[RHO];
				Fg[NRJ] = F[NRJ];
				Fg[M0 + D] = F[M0]; Fg[M0 + (D + 1) % 3] = F[M1]; Fg[M0 + (D + 2) % 3] = F[M2];
				Fg[B0 + D] = F[B0]; Fg[B0 + (D + 1) % 3] = F[B1]; Fg[B0 + (D + 2) % 3] = F[B2];
			}
			const int m = d.minsub[o];
			const double min_dt = *d.sub_dt * m;
			double cq[8];
			for (int q = 0; q < 8; q++) cq[q] = min_dt * Fg[q];
			const int po = d.period[o], pc = d.period[c];
			const int pmax = po > pc ? po : pc, pmin = po > pc ? pc : po;
			const bool fresh = ((sub - m) & (pmax - 1)) == 0;
			for (int q = 0; q < 8; q++) A[q * n + c] = (fresh ? 0.0 : A[q * n + c]) + cq[q];
			if (pmin < pmax) {
				const bool fresh2 = ((sub - m) & (pmin - 1)) == 0;
				for (int q = 0; q < 8; q++) A2[q * n + c] = (fresh2 ? 0.0 : A2[q * n + c]) + cq[q];
			}
			Em[c] = cq[D == 1 ? B0 + 2 : B0 + 1];
			Em[n + c] = cq[B0];
			solved++;
			Ls = Rs;
			cfl = cfr;
			stage_l = stage_r;
			code_l = code_r;
			o = c;
		}
	}
	/* faces solved (timestep()'s flux_calcs), one atomic per wave */
	for (int off = 32; off > 0; off >>= 1) solved += __shfl_xor(solved, off, 64);
	if ((threadIdx.x + BX * threadIdx.y) % 64 == 0 && solved)
		atomicAdd(&d.scal[4 * SS_SLOTS + ss_slot()], (unsigned long long)solved);
}

/* the edge fields of direction D set to zero in the blocks of its zeroing
   list (no face due, the previous substep's values still there), what the
   flux pass would write there */
template <int D>
__global__ __launch_bounds__(BX * BY) void ss_zero_emf_kernel(GridDev g, SubstepDev d,
	double* __restrict__ emf, int sub)
{
	const int cnt = d.count[SS_LISTS * (sub & 1) + SS_LIST_ZERO + D];
	const int* list = d.list + (size_t)(SS_LIST_ZERO + D) * d.nblk;
	const long long n = g.n;
	double* Em = emf + (long long)(2 * D) * n;
	for (int idx = blockIdx.x; idx < cnt; idx += gridDim.x) {
		int bx, by, bz;
		ss_block_coords(g, list[idx], bx, by, bz);
		const int i = bx * BX + threadIdx.x, j = by * BY + threadIdx.y;
		if (i >= g.nx || j >= g.ny) continue;
		const int c = cidx(g, i, j, bz);
		Em[c] = 0;
		Em[n + c] = 0;
	}
}

/* constrained-transport edge terms from the substep's 6 field components
   per cell (emf), 0 beyond a closed end */
struct EdgeTerms6B {
	const double* P;
	long long n;
	double dx, dy, dz;
	__device__ __forceinline__ double at(int q, long long c) const { return c < 0 ? 0.0 : P[q * n + c]; }
	__device__ __forceinline__ double ezx(long long c) const { return +(at(0, c) * dz / 4); }
	__device__ __forceinline__ double eyx(long long c) const { return -(at(1, c) * dy / 4); }
	__device__ __forceinline__ double exy(long long c) const { return +(at(2, c) * dx / 4); }
	__device__ __forceinline__ double ezy(long long c) const { return -(at(3, c) * dz / 4); }
	__device__ __forceinline__ double exz(long long c) const { return -(at(4, c) * dx / 4); }
	__device__ __forceinline__ double eyz(long long c) const { return +(at(5, c) * dy / 4); }
};

/*
Face dB sums of substep `sub`: each face's dB (the same 12 edge terms, in
the same order, for both of its copies) is summed per face like the
fluxes: one sum where the two cells' periods are equal, a second for the
faster cell where they differ, each restarting at the first substep after
the cells it serves applied (x % p is x & (p - 1): periods are powers of
two).  A cell computes the sums of the faces it owns (its -x, -y, -z
faces) and of its +d faces at a closed end (kept in its end slots).
*/
/* ZERO_DB: no face around the block is due, so every edge field its faces
   use is zero and every dB is +0 (a sum of +-0 terms starting at +0); the
   sums are written as they would be, without the edge terms */
template <bool ZERO_DB>
__device__ __forceinline__ void ss_dbsum_block(const GridDev& g, const SubstepDev& d,
	const double* __restrict__ emf, int sub, int bx, int by, int k)
{
	const int i = bx * BX + threadIdx.x, j = by * BY + threadIdx.y;
	if (i >= g.nx || j >= g.ny) return;
	const long long n = g.n;
	const int c = cidx(g, i, j, k);
	const int im = ss_coord(i - 1, g.nx, d.per[0]), ip = ss_coord(i + 1, g.nx, d.per[0]);
	const int jm = ss_coord(j - 1, g.ny, d.per[1]), jp = ss_coord(j + 1, g.ny, d.per[1]);
	const int km = ss_coord(k - 1, g.nz, d.per[2]), kp = ss_coord(k + 1, g.nz, d.per[2]);
	const NbB L{g.nx, g.ny};
	const EdgeTerms6B E{emf, n, g.dx, g.dy, g.dz};
	const int pc = d.period[c];
	double* fdb = d.fdb;
	auto accumulate = [&](long long slow_i, long long fast_i, long long other, double db) {
		const int po = other >= 0 ? d.period[other] : pc;
		const int pmax = po > pc ? po : pc, pmin = po > pc ? pc : po;
		fdb[slow_i] = ((((sub - 1) & (pmax - 1)) == 0) ? 0.0 : fdb[slow_i]) + db;
		if (pmin < pmax) fdb[fast_i] = ((((sub - 1) & (pmin - 1)) == 0) ? 0.0 : fdb[fast_i]) + db;
	};
	if constexpr (ZERO_DB) {
		(void)E;
		accumulate(c, 3 * n + c, L(im, j, k), 0.0);
		accumulate(n + c, 4 * n + c, L(i, jm, k), 0.0);
		accumulate(2 * n + c, 5 * n + c, L(i, j, km), 0.0);
		if (ip < 0) accumulate(6 * n + c, 6 * n + c, -1, 0.0);
		if (jp < 0) accumulate(7 * n + c, 7 * n + c, -1, 0.0);
		if (kp < 0) accumulate(8 * n + c, 8 * n + c, -1, 0.0);
	} else {
		accumulate(c, 3 * n + c, L(im, j, k), dBx_at(E, L, i, im, j, jm, jp, k, km, kp));
		accumulate(n + c, 4 * n + c, L(i, jm, k), dBy_at(E, L, i, im, ip, j, jm, k, km, kp));
		accumulate(2 * n + c, 5 * n + c, L(i, j, km), dBz_at(E, L, i, im, ip, j, jm, jp, k, km));
		if (ip < 0) accumulate(6 * n + c, 6 * n + c, -1, dBx_at(E, L, ip, i, j, jm, jp, k, km, kp));
		if (jp < 0) accumulate(7 * n + c, 7 * n + c, -1, dBy_at(E, L, i, im, ip, jp, j, k, km, kp));
		if (kp < 0) accumulate(8 * n + c, 8 * n + c, -1, dBz_at(E, L, i, im, ip, j, jm, jp, kp, k));
	}
}

/* the face-dB pass, with WALK over the blocks of its list (those around
   which a face is due) and, with ZERO_DB, of its list of blocks with a sum
   that restarts and no face due around them (a block with neither is in
   no list: every dB is +0 and no sum changes); else over every block */
template <bool WALK, bool ZERO_DB>
__global__ __launch_bounds__(BX * BY) void ss_dbsum_kernel(GridDev g, SubstepDev d,
	const double* __restrict__ emf, int sub)
{
	if constexpr (WALK) {
		const int l = ZERO_DB ? SS_LIST_DB0 : SS_LIST_DB;
		const int cnt = d.count[SS_LISTS * (sub & 1) + l];
		const int* list = d.list + (size_t)l * d.nblk;
		for (int idx = blockIdx.x; idx < cnt; idx += gridDim.x) {
			int bx, by, bz;
			ss_block_coords(g, list[idx], bx, by, bz);
			ss_dbsum_block<ZERO_DB>(g, d, emf, sub, bx, by, bz);
		}
	} else {
		ss_dbsum_block<false>(g, d, emf, sub, blockIdx.x, blockIdx.y, blockIdx.z);
	}
}

/*
The update of cells due at substep `sub` (update_mhd_state,
source/mhd/solve.hpp:1503-1603, and update_vol_B, :1618-1683, from their
own face copies): a due cell applies the flux and face dB sums of its six
faces that serve it (the faster cell's where the two periods differ).
*/
__device__ __forceinline__ void ss_update_block(const GridDev& g, const SubstepDev& d, int sub,
	int bx, int by, int k)
{
	const int i = bx * BX + threadIdx.x, j = by * BY + threadIdx.y;
	if (i >= g.nx || j >= g.ny) return;
	const long long n = g.n;
	const int c = cidx(g, i, j, k);
	const int pc = d.period[c];
	if (sub & (pc - 1)) return;
	const int ct = d.ct ? d.ct[c] : 1;
	if (ct <= 0) {
		/* boundary cells: update_vol_B at their period (constant
		   pressure, own face copies); cells not solved: nothing */
		if (ct == 0) {
			const double rho = d.S[c], m[3] = {d.S[n + c], d.S[2 * n + c], d.S[3 * n + c]};
			if (rho > 0) {
				double En = d.S[4 * n + c];
				const double vb[3] = {d.S[5 * n + c], d.S[6 * n + c], d.S[7 * n + c]};
				const double vb_new[3] = {
					0.5 * (d.S[8 * n + c] + d.S[9 * n + c]),
					0.5 * (d.S[10 * n + c] + d.S[11 * n + c]),
					0.5 * (d.S[12 * n + c] + d.S[13 * n + c])};
				refit_energy(rho, m, En, vb, vb_new, g.ph);
				for (int q = 0; q < 3; q++) d.S[(5 + q) * n + c] = vb_new[q];
				d.S[4 * n + c] = En;
			}
		}
		return;
	}
	const int im = ss_coord(i - 1, g.nx, d.per[0]), ip = ss_coord(i + 1, g.nx, d.per[0]);
	const int jm = ss_coord(j - 1, g.ny, d.per[1]), jp = ss_coord(j + 1, g.ny, d.per[1]);
	const int km = ss_coord(k - 1, g.nz, d.per[2]), kp = ss_coord(k + 1, g.nz, d.per[2]);
	const NbB L{g.nx, g.ny};
	/* the neighbour across each face (-1: none) */
	const long long nb[6] = {L(im, j, k), L(ip, j, k), L(i, jm, k), L(i, jp, k), L(i, j, km), L(i, j, kp)};
	double sl[6][8], fd[6];
	for (int q3 = 0; q3 < 3; q3++)
	for (int side = 0; side < 2; side++) {
		const int f = 2 * q3 + side;
		const long long other = nb[f];
		const int po = other >= 0 ? d.period[other] : pc;
		const bool fast = pc < po;    /* this cell is the faster of the two */
		/* the face's owner: c for -d, the +d neighbour for +d (none: the
		   cell's end slot, and a flux sum of +0, the value of a slot
		   nothing was added to) */
		const long long own = side == 0 ? c : other;
		fd[f] = own >= 0 ? d.fdb[(long long)((fast ? 3 : 0) + q3) * n + own] : d.fdb[(long long)(6 + q3) * n + c];
		const double* A = d.slot + (long long)((fast ? 3 : 0) + q3) * 8 * n;
		for (int q = 0; q < 8; q++) sl[f][q] = own >= 0 ? A[q * n + own] : 0.0;
	}
	const double len[3] = {g.dx, g.dy, g.dz};
	const double area[3] = {g.dy * g.dz, g.dx * g.dz, g.dx * g.dy};
	double rho = d.S[c], m[3] = {d.S[n + c], d.S[2 * n + c], d.S[3 * n + c]}, En = d.S[4 * n + c];
	double vb[3] = {d.S[5 * n + c], d.S[6 * n + c], d.S[7 * n + c]};
	for (int q3 = 0; q3 < 3; q3++) rho += (sl[2 * q3][RHO] - sl[2 * q3 + 1][RHO]) / len[q3];
	for (int q3 = 0; q3 < 3; q3++) {
		const double il = 1 / len[q3];
		for (int q = 0; q < 3; q++) m[q] = m[q] + (sl[2 * q3][M0 + q] - sl[2 * q3 + 1][M0 + q]) * il;
	}
	for (int q3 = 0; q3 < 3; q3++) En += (sl[2 * q3][NRJ] - sl[2 * q3 + 1][NRJ]) / len[q3];
	for (int q3 = 0; q3 < 3; q3++) {
		const double il = 1 / len[q3];
		for (int q = 0; q < 3; q++) vb[q] = vb[q] + (sl[2 * q3][B0 + q] - sl[2 * q3 + 1][B0 + q]) * il;
	}
	double fc[6];
	for (int q3 = 0; q3 < 3; q3++) {
		fc[2 * q3] = d.S[(8 + 2 * q3) * n + c] + fd[2 * q3] / area[q3];
		fc[2 * q3 + 1] = d.S[(9 + 2 * q3) * n + c] + fd[2 * q3 + 1] / area[q3];
	}
	for (int f = 0; f < 6; f++) d.S[(8 + f) * n + c] = fc[f];
	d.S[c] = rho;
	for (int q = 0; q < 3; q++) d.S[(1 + q) * n + c] = m[q];
	if (rho > 0) {
		const double vb_new[3] = {0.5 * (fc[0] + fc[1]), 0.5 * (fc[2] + fc[3]), 0.5 * (fc[4] + fc[5])};
		refit_energy(rho, m, En, vb, vb_new, g.ph);
		for (int q = 0; q < 3; q++) d.S[(5 + q) * n + c] = vb_new[q];
	} else {
		for (int q = 0; q < 3; q++) d.S[(5 + q) * n + c] = vb[q];
	}
	d.S[4 * n + c] = En;
}

/* the update, with WALK over the blocks of its list (those with a cell
   that may be due), else over every block */
template <bool WALK>
__global__ __launch_bounds__(BX * BY) void ss_update_kernel(GridDev g, SubstepDev d, int sub)
{
	if constexpr (WALK) {
		const int cnt = d.count[SS_LISTS * (sub & 1) + SS_LIST_UPDATE];
		const int* list = d.list + (size_t)SS_LIST_UPDATE * d.nblk;
		for (int idx = blockIdx.x; idx < cnt; idx += gridDim.x) {
			int bx, by, bz;
			ss_block_coords(g, list[idx], bx, by, bz);
			ss_update_block(g, d, sub, bx, by, bz);
		}
	} else {
		ss_update_block(g, d, sub, blockIdx.x, blockIdx.y, blockIdx.z);
	}
}

/* sync_magnetic_field, common_functions.hpp:184-269: the two visits to a
   face, the cell with the smaller index first */
template <int D>
__global__ __launch_bounds__(BX * BY) void ss_sync_kernel(GridDev g, SubstepDev d)
{
	const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y, k = blockIdx.z;
	if (i >= g.nx || j >= g.ny) return;
	const long long n = g.n;
	const int b = cidx(g, i, j, k);
	const int ai = D == 0 ? ss_coord(i - 1, g.nx, d.per[0]) : i;
	const int aj = D == 1 ? ss_coord(j - 1, g.ny, d.per[1]) : j;
	const int ak = D == 2 ? ss_coord(k - 1, g.nz, d.per[2]) : k;
	if (ai < 0 || aj < 0 || ak < 0) return;
	const int a = cidx(g, ai, aj, ak);
	if (!ss_active(d, a) || !ss_active(d, b)) return;
	/* a's + copy and b's - copy; the visit from the smaller index first,
	   from the -D side when a cell is its own neighbour */
	double pa = d.S[(9 + 2 * D) * n + a], pb = d.S[(8 + 2 * D) * n + b];
	if (a < b) {
		double diff = pa - pb; pa -= diff / 2; pb += diff / 2;
		diff = pb - pa; pb -= diff / 2; pa += diff / 2;
	} else {
		double diff = pb - pa; pb -= diff / 2; pa += diff / 2;
		diff = pa - pb; pa -= diff / 2; pb += diff / 2;
	}
	d.S[(9 + 2 * D) * n + a] = pa;
	d.S[(8 + 2 * D) * n + b] = pb;
}

/* update_vol_B(0) of all cells and the face-once state; with `refit`
   also the constant-pressure refit of the main loop's apply_boundaries
   (refit_kernel's, with the field of the face-once state: the cell's and
   its +d neighbour's own -d copies, which this kernel does not write).
   The working layout is left as ss_expand_kernel would make it from the
   state written here (energy, face-average field, +d copies from the
   neighbours' -d copies), so the next step may skip that expansion */
__global__ __launch_bounds__(BX * BY) void ss_final_kernel(GridDev g, SubstepDev d, double* __restrict__ s,
	int refit)
{
	const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y, k = blockIdx.z;
	if (i >= g.nx || j >= g.ny) return;
	const long long n = g.n;
	const int c = cidx(g, i, j, k);
	const double rho = d.S[c], m[3] = {d.S[n + c], d.S[2 * n + c], d.S[3 * n + c]};
	double En = d.S[4 * n + c];
	const double vb[3] = {d.S[5 * n + c], d.S[6 * n + c], d.S[7 * n + c]};
	const double vb_new[3] = {
		0.5 * (d.S[8 * n + c] + d.S[9 * n + c]),
		0.5 * (d.S[10 * n + c] + d.S[11 * n + c]),
		0.5 * (d.S[12 * n + c] + d.S[13 * n + c])};
	if (d.bounded) {
		/* the state stays in the working layout: update_vol_B(0) of the
		   cells of type >= 0 in place */
		if (ss_active(d, c) && rho > 0) {
			refit_energy(rho, m, En, vb, vb_new, g.ph);
			for (int q = 0; q < 3; q++) d.S[(5 + q) * n + c] = vb_new[q];
			d.S[4 * n + c] = En;
		}
		return;
	}
	if (rho > 0) refit_energy(rho, m, En, vb, vb_new, g.ph);
	const int nb[3] = {cidx(g, wrap_inc(i, g.nx), j, k), cidx(g, i, wrap_inc(j, g.ny), k),
		cidx(g, i, j, wrap_inc(k, g.nz))};
	double lo[3], up[3], vf[3];
	for (int q = 0; q < 3; q++) {
		lo[q] = d.S[(8 + 2 * q) * n + c];
		up[q] = d.S[(8 + 2 * q) * n + nb[q]];
		vf[q] = 0.5 * (lo[q] + up[q]);
	}
	if (refit) refit_energy(rho, m, En, vf, vf, g.ph);
	s[c] = rho;
	for (int q = 0; q < 3; q++) s[(1 + q) * n + c] = m[q];
	s[4 * n + c] = En;
	for (int q = 0; q < 3; q++) s[(5 + q) * n + c] = lo[q];
	d.S[4 * n + c] = En;
	for (int q = 0; q < 3; q++) {
		d.S[(5 + q) * n + c] = vf[q];
		d.S[(9 + 2 * q) * n + c] = up[q];
	}
}

__global__ __launch_bounds__(BX * BY) void ss_fill_kernel(double* __restrict__ a, long long n, double v)
{
	const long long t = (long long)blockIdx.x * blockDim.x * blockDim.y + threadIdx.x + blockDim.x * threadIdx.y;
	if (t < n) a[t] = v;
}

} // namespace

struct pamhd_handle {
	GridDev g;
	pamhd_grid_desc desc;
	pamhd_physics phys;
	int precision = PAMHD_FP64;
	double* state[2] = {nullptr, nullptr};
	double* flux = nullptr;
	double* bg = nullptr;
	double* dt = nullptr;
	double* scratch = nullptr;
	size_t scratch_len = 0;
	StepStatus* status = nullptr; /* 2 entries */
	double* host_dt = nullptr;    /* pinned */
	StepStatus* host_status = nullptr;
	/* launch choices, set by pamhd_autotune (defaults: per-cell update) */
	/* until pamhd_autotune or pamhd_set_launch: the variants the autotuner
	   picks for the 256^3 HLLD step (every variant gives the same bits) */
	struct Tune {
		int flux_remap = 1;
		int update_kz = 8;
		int update_remap = 0;
		int slabs = 1;   /* > 1: flux pass and update pipelined over z-slabs */
		int flux_occ = 0; /* 1: flux kernels compiled for 4 blocks per CU */
		/* HLLD's per-cell part: 0 per face, 1 shared through LDS, 2 carried
		   along marches of flux_km cells (y- and z-faces), 3 as 2 and the
		   x-faces marched along y with the part moved across lanes */
		int flux_shared = 3;
		int flux_km = 16;
		int update_tile = 1;   /* column update tile: 64x4, 128x2, 32x8, 128x4, 64x8 */
		/* HLLD: 1 = the fused step (every face flux and the conserved
		   update in one pass, then the face-field update and refits),
		   0 = flux passes and update */
		/* HLLD step structure: 0 = flux passes and update, 1 = the fused
		   single pass, 2 = the accumulating marches (each followed by the
		   face-field update and refits), -1 = the default (2) */
		int fused = -1;
		int fused_kc = 16;     /* planes per chunk of the fused pass (<= FKC_MAX) */
		int acc_km = 32;       /* cells per march of the accumulating step */
	} tune;
	double* emf = nullptr;     /* 6 x n field components of the fused step, allocated on first use */
	/* second stream and events of the slab pipeline, created on first use */
	hipStream_t aux = nullptr;
	std::vector<hipEvent_t> ev;
	/* optional per-kernel device time accounting with HIP events */
	bool timing = false;
	struct Timed { int kernel; hipEvent_t start, stop; };
	std::vector<Timed> timed;
	double kernel_ms[4] = {0, 0, 0, 0};
	long kernel_launches[4] = {0, 0, 0, 0};
	BoxSet* box = nullptr;    /* solar-wind box (non-periodic) handles only */
	double* adj = nullptr;    /* adjoint scratch, allocated on first use */
	PoissonMG* mg = nullptr;  /* multigrid hierarchy, allocated on first cleaning */
	SubstepDev* sub = nullptr; /* per-cell substepping, allocated when enabled */
	int mg_fused = 0;   /* per-colour passes (faster; see pamhd_set_poisson_smoother) */
	int adj_reverse = 1;   /* flux Jacobian of the adjoint: 1 reverse accumulation, 0 dual numbers */
	int flux_step = 0;     /* the last step left its FP64 dt * F in flux (pamhd_get_magnetic_fluxes) */
	int ss_walk = 0;       /* workgroups of the substep passes that walk block lists (0: default) */
	double* pe_out = nullptr;   /* where the update pass also writes the step's magnetic fluxes */
	int cur = 0;    /* state buffer holding the current state */
	int last = 0;   /* status entry of the last flux pass */
	unsigned long long flux_calcs = 0;   /* faces solved by the last step */
	/* the substepped step leaves its working layout as the expansion of the
	   state it returns, and the next one skips the expansion when nothing
	   else reached the handle in between: `calls` counts the entry points
	   that take the handle, ss_at is its value after the last substepped
	   step (ss_valid), and a state pointer handed out (pamhd_state_device_ptr)
	   turns the skip off for good */
	unsigned long long calls = 0, ss_at = 0;
	bool ss_valid = false;
	bool state_escaped = false;
	std::string err;
};

namespace {

/* an entry point that takes the handle was called */
inline void touch(pamhd_handle* h)
{
	if (h) h->calls++;
}

enum { K_DT = 0, K_FLUX = 1, K_UPDATE = 2, K_REFIT = 3 };
const char* const kernel_names[4] = {"dt", "flux_pass", "update", "refit"};

/* brackets a launch with events when timing is enabled */
struct TimedLaunch {
	pamhd_handle* h;
	hipStream_t st;
	int k;
	hipEvent_t a = nullptr, b = nullptr;
	TimedLaunch(pamhd_handle* h_, hipStream_t st_, int k_) : h(h_), st(st_), k(k_)
	{
		if (h->timing && hipEventCreate(&a) == hipSuccess && hipEventCreate(&b) == hipSuccess) {
			(void)hipEventRecord(a, st);
		} else {
			a = b = nullptr;
		}
	}
	~TimedLaunch()
	{
		if (a && b) {
			(void)hipEventRecord(b, st);
			h->timed.push_back({k, a, b});
		}
	}
};

void collect_times(pamhd_handle* h)
{
	for (auto& t : h->timed) {
		(void)hipEventSynchronize(t.stop);
		float ms = 0;
		if (hipEventElapsedTime(&ms, t.start, t.stop) == hipSuccess) {
			h->kernel_ms[t.kernel] += ms;
			h->kernel_launches[t.kernel]++;
		}
		(void)hipEventDestroy(t.start);
		(void)hipEventDestroy(t.stop);
	}
	h->timed.clear();
}

int fail(pamhd_handle* h, int code, const std::string& m)
{
	if (h) h->err = m;
	else set_global_error(m);
	return code;
}

int hip_fail(pamhd_handle* h, hipError_t e, const char* what)
{
	return fail(h, PAMHD_E_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIP_TRY(h, expr) do { hipError_t e_ = (expr); if (e_ != hipSuccess) return hip_fail((h), e_, #expr); } while (0)

dim3 cell_grid(const GridDev& g, int slabs = 1)
{
	return dim3((g.nx + BX - 1) / BX, (g.ny + BY - 1) / BY, g.nz * slabs);
}

template <int SOLVER, class T, class PT>
void launch_flux_dirs(pamhd_handle* h, hipStream_t st, StepStatus* stat, int k0, int nk)
{
	const double* s = h->state[h->cur];
	StateArrays S;
	for (int q = 0; q < 8; q++) S.q[q] = s + (size_t)q * h->g.n;
	PT* P = reinterpret_cast<PT*>(h->flux);
	const unsigned gx = (h->g.nx + BX - 1) / BX, gy = (h->g.ny + BY - 1) / BY;
	const dim3 grid(gx * gy * (unsigned)nk), block(BX, BY, 1);
	const int r = h->tune.flux_remap;
	constexpr bool hlld64 = SOLVER == HLLD_ATHENA && std::is_same_v<T, double>;
	if constexpr (hlld64) if (h->tune.flux_shared >= 2) {
		const int km = h->tune.flux_km;
		const unsigned g1 = gx * (unsigned)((nk + BY - 1) / BY) * (unsigned)((h->g.ny + km - 1) / km);
		const unsigned g2 = gx * (unsigned)((h->g.ny + BY - 1) / BY) * (unsigned)((nk + km - 1) / km);
		const int ke = k0 + nk;
		/* flux_shared 3: the x-faces marched too (km <= 64: one row per lane
		   for the tile's left edge) */
		const bool xm = h->tune.flux_shared == 3 && km <= 64;
		if (h->tune.flux_occ) {
			if (xm) flux_hlld_xmarch_kernel<T, PT, 4><<<g1, block, 0, st>>>(h->g, S, P, h->dt, stat, gx, km, r, k0, ke);
			else flux_pass_kernel<SOLVER, T, PT, 0, 4><<<grid, block, 0, st>>>(h->g, S, h->bg, P, h->dt, stat, gx, gy, r, k0);
			flux_hlld_march_kernel<T, PT, 1, 4><<<g1, block, 0, st>>>(h->g, S, P, h->dt, stat, gx, km, r, k0, ke);
			flux_hlld_march_kernel<T, PT, 2, 4><<<g2, block, 0, st>>>(h->g, S, P, h->dt, stat, gx, km, r, k0, ke);
		} else {
			if (xm) flux_hlld_xmarch_kernel<T, PT, 1><<<g1, block, 0, st>>>(h->g, S, P, h->dt, stat, gx, km, r, k0, ke);
			else flux_pass_kernel<SOLVER, T, PT, 0, 1><<<grid, block, 0, st>>>(h->g, S, h->bg, P, h->dt, stat, gx, gy, r, k0);
			flux_hlld_march_kernel<T, PT, 1, 1><<<g1, block, 0, st>>>(h->g, S, P, h->dt, stat, gx, km, r, k0, ke);
			flux_hlld_march_kernel<T, PT, 2, 1><<<g2, block, 0, st>>>(h->g, S, P, h->dt, stat, gx, km, r, k0, ke);
		}
		return;
	}
	if ((SOLVER == HLL_ATHENA || SOLVER == RUSANOV) && h->tune.flux_shared >= 2 && !h->bg) {
		const int km = h->tune.flux_km;
		const unsigned gyz = gx * (unsigned)((nk + BY - 1) / BY) * (unsigned)((h->g.ny + km - 1) / km);
		const unsigned gz = gx * (unsigned)((h->g.ny + BY - 1) / BY) * (unsigned)((nk + km - 1) / km);
		const int ke = k0 + nk;
		constexpr int SV = SOLVER == HLL_ATHENA ? HLL_ATHENA : RUSANOV;
		if (h->tune.flux_shared == 3 && km <= 32)
			flux_cell_march_kernel<SV, T, PT, 0, 1><<<gyz, block, 0, st>>>(h->g, S, P, h->dt, stat, gx, km, r, k0, ke);
		else
			flux_pass_kernel<SOLVER, T, PT, 0, 1><<<grid, block, 0, st>>>(h->g, S, h->bg, P, h->dt, stat, gx, gy, r, k0);
		flux_cell_march_kernel<SV, T, PT, 1, 1><<<gyz, block, 0, st>>>(h->g, S, P, h->dt, stat, gx, km, r, k0, ke);
		flux_cell_march_kernel<SV, T, PT, 2, 1><<<gz, block, 0, st>>>(h->g, S, P, h->dt, stat, gx, km, r, k0, ke);
		return;
	}
	if constexpr (hlld64) if (h->tune.flux_shared == 1) {
		const dim3 grid2(gx * (unsigned)h->g.ny * (unsigned)((nk + BY - 1) / BY));
		const int ke = k0 + nk;
		if (h->tune.flux_occ) {
			flux_hlld_shared_kernel<T, PT, 0, 4><<<grid, block, 0, st>>>(h->g, S, P, h->dt, stat, gx, gy, r, k0, ke);
			flux_hlld_shared_kernel<T, PT, 1, 4><<<grid, block, 0, st>>>(h->g, S, P, h->dt, stat, gx, gy, r, k0, ke);
			flux_hlld_shared_kernel<T, PT, 2, 4><<<grid2, block, 0, st>>>(h->g, S, P, h->dt, stat, gx, gy, r, k0, ke);
		} else {
			flux_hlld_shared_kernel<T, PT, 0, 1><<<grid, block, 0, st>>>(h->g, S, P, h->dt, stat, gx, gy, r, k0, ke);
			flux_hlld_shared_kernel<T, PT, 1, 1><<<grid, block, 0, st>>>(h->g, S, P, h->dt, stat, gx, gy, r, k0, ke);
			flux_hlld_shared_kernel<T, PT, 2, 1><<<grid2, block, 0, st>>>(h->g, S, P, h->dt, stat, gx, gy, r, k0, ke);
		}
		return;
	}
	if (h->tune.flux_occ) {
		flux_pass_kernel<SOLVER, T, PT, 0, 4><<<grid, block, 0, st>>>(h->g, S, h->bg, P, h->dt, stat, gx, gy, r, k0);
		flux_pass_kernel<SOLVER, T, PT, 1, 4><<<grid, block, 0, st>>>(h->g, S, h->bg, P, h->dt, stat, gx, gy, r, k0);
		flux_pass_kernel<SOLVER, T, PT, 2, 4><<<grid, block, 0, st>>>(h->g, S, h->bg, P, h->dt, stat, gx, gy, r, k0);
		return;
	}
	flux_pass_kernel<SOLVER, T, PT, 0, 1><<<grid, block, 0, st>>>(h->g, S, h->bg, P, h->dt, stat, gx, gy, r, k0);
	flux_pass_kernel<SOLVER, T, PT, 1, 1><<<grid, block, 0, st>>>(h->g, S, h->bg, P, h->dt, stat, gx, gy, r, k0);
	flux_pass_kernel<SOLVER, T, PT, 2, 1><<<grid, block, 0, st>>>(h->g, S, h->bg, P, h->dt, stat, gx, gy, r, k0);
}

/* faces owned by the cells of planes [k0, k0 + nk), all three directions */
template <class T, class PT>
int launch_flux_pass(pamhd_handle* h, int solver, hipStream_t st, StepStatus* stat, int k0 = 0,
	int nk = -1)
{
	if (nk < 0) nk = h->g.nz;
	switch (solver) {
	case RUSANOV: launch_flux_dirs<RUSANOV, T, PT>(h, st, stat, k0, nk); break;
	case HLL_ATHENA: launch_flux_dirs<HLL_ATHENA, T, PT>(h, st, stat, k0, nk); break;
	case HLLD_ATHENA: launch_flux_dirs<HLLD_ATHENA, T, PT>(h, st, stat, k0, nk); break;
	case ROE_ATHENA: launch_flux_dirs<ROE_ATHENA, T, PT>(h, st, stat, k0, nk); break;
	default:
		return fail(h, PAMHD_E_SOLVER, "invalid solver " + std::to_string(solver));
	}
	HIP_TRY(h, hipGetLastError());
	return PAMHD_OK;
}

/* update_kz = 0 selects the per-cell update kernel, > 0 the z-marching one
   with that many planes per block */
template <class PT>
void launch_update(pamhd_handle* h, hipStream_t st, int refits, int k0 = 0, int nk = -1)
{
	const PT* P = reinterpret_cast<const PT*>(h->flux);
	/* the step's magnetic fluxes, when a caller asked for them (FP64 only) */
	double* pe = std::is_same<PT, double>::value ? h->pe_out : nullptr;
	const int kz = h->tune.update_kz < h->g.nz ? h->tune.update_kz : h->g.nz;
	if (kz <= 0 || nk >= 0) {
		if (nk < 0) nk = h->g.nz;
		const unsigned gx = (h->g.nx + BX - 1) / BX, gy = (h->g.ny + BY - 1) / BY;
		update_cell_kernel<PT><<<dim3(gx * gy * nk), dim3(BX, BY, 1), 0, st>>>(
			h->g, h->state[h->cur], h->state[1 - h->cur], P, refits, gx, gy, h->tune.update_remap, k0,
			pe);
		return;
	}
	const unsigned gz = (h->g.nz + kz - 1) / kz;
	auto column = [&](auto tx, auto ty) {
		constexpr int TX = decltype(tx)::value, TY = decltype(ty)::value;
		const unsigned gx = (h->g.nx + TX - 1) / TX, gy = (h->g.ny + TY - 1) / TY;
		update_column_kernel<PT, TX, TY><<<dim3(gx * gy * gz), dim3(TX, TY, 1), 0, st>>>(
			h->g, h->state[h->cur], h->state[1 - h->cur], P, refits, kz, gx, gy, h->tune.update_remap,
			pe);
	};
	switch (h->tune.update_tile) {
	case 1: column(std::integral_constant<int, 128>(), std::integral_constant<int, 2>()); break;
	case 2: column(std::integral_constant<int, 32>(), std::integral_constant<int, 8>()); break;
	case 3: column(std::integral_constant<int, 128>(), std::integral_constant<int, 4>()); break;
	case 4: column(std::integral_constant<int, 64>(), std::integral_constant<int, 8>()); break;
	default: column(std::integral_constant<int, 64>(), std::integral_constant<int, 4>()); break;
	}
}

int report_status(pamhd_handle* h, const StepStatus& st)
{
	if (st.err == 0) return PAMHD_OK;
	const long long n = h->g.n;
	const long long face = st.err_face;
	if (st.err == PAMHD_E_SUBSTEP) {
		const int i = (int)(face % h->g.nx), j = (int)((face / h->g.nx) % h->g.ny),
			k = (int)(face / ((long long)h->g.nx * h->g.ny));
		char buf[256];
		snprintf(buf, sizeof(buf), "cell (%d,%d,%d): %s", i, j, k, error_name(st.err));
		return fail(h, st.err, buf);
	}
	const int d = (int)(face / n);
	const long long c = face - (long long)d * n;
	const int i = (int)(c % h->g.nx), j = (int)((c / h->g.nx) % h->g.ny), k = (int)(c / ((long long)h->g.nx * h->g.ny));
	char buf[256];
	snprintf(buf, sizeof(buf),
		"solution failed between cells (%d,%d,%d) and its -%c neighbour in direction %d: %s",
		i, j, k, "xyz"[d], d + 1, error_name(st.err));
	return fail(h, st.err, buf);
}

/* the fused step applies to HLLD (which ignores the background field) */
/* the step structure used for HLLD (which ignores the background field):
   0 flux passes and update, 1 fused single pass, 2 accumulating marches */
int hlld_mode(const pamhd_handle* h)
{
	return h->tune.fused >= 0 ? h->tune.fused : 2;
}

bool fused_step(const pamhd_handle* h, int solver)
{
	return solver == HLLD_ATHENA && hlld_mode(h) != 0;
}

/* the face-field update and energy refits after the fused or accumulating
   passes: reads the old face field from s, rho, m and the recovered
   pressure from out, the field components from emf; completes out */
int launch_ct_refit(pamhd_handle* h, hipStream_t st, const double* s, double* out, int refits)
{
	TimedLaunch t(h, st, K_UPDATE);
	const int kz = h->tune.update_kz > 0 ? min(h->tune.update_kz, h->g.nz) : min(8, h->g.nz);
	const unsigned gzc = (h->g.nz + kz - 1) / kz;
	auto column = [&](auto tx, auto ty) {
		constexpr int TX = decltype(tx)::value, TY = decltype(ty)::value;
		const unsigned cx = (h->g.nx + TX - 1) / TX, cy = (h->g.ny + TY - 1) / TY;
		ct_refit_kernel<TX, TY><<<dim3(cx * cy * gzc), dim3(TX, TY, 1), 0, st>>>(
			h->g, s, out, h->emf, refits, kz, cx, cy, h->tune.update_remap & 1);
	};
	switch (h->tune.update_tile) {
	case 1: column(std::integral_constant<int, 128>(), std::integral_constant<int, 2>()); break;
	case 2: column(std::integral_constant<int, 32>(), std::integral_constant<int, 8>()); break;
	case 3: column(std::integral_constant<int, 128>(), std::integral_constant<int, 4>()); break;
	case 4: column(std::integral_constant<int, 64>(), std::integral_constant<int, 8>()); break;
	default: column(std::integral_constant<int, 64>(), std::integral_constant<int, 4>()); break;
	}
	HIP_TRY(h, hipGetLastError());
	return PAMHD_OK;
}

constexpr int FNW = 8;   /* waves per block of the fused pass: 6 rows and their two neighbours */

/* the two passes of the fused step: pass 1 reads state[cur] and writes rho,
   m and the recovered pressure into state[1 - cur] and the field components
   into emf; pass 2 completes state[1 - cur] */
template <class T>
int launch_fused(pamhd_handle* h, hipStream_t st, StepStatus* next, int refits)
{
	const long long n = h->g.n;
	if (!h->emf) HIP_TRY(h, hipMalloc(&h->emf, sizeof(double) * 6 * n));
	const double* s = h->state[h->cur];
	double* out = h->state[1 - h->cur];
	StateArrays S;
	for (int q = 0; q < 8; q++) S.q[q] = s + (size_t)q * n;
	int kc = h->tune.fused_kc;
	if (kc < 1) kc = 1;
	if (kc > FKC_MAX) kc = FKC_MAX;
	const unsigned gx = (h->g.nx + FLANES - 1) / FLANES, gy = (h->g.ny + FNW - 3) / (FNW - 2);
	const unsigned gz = (h->g.nz + kc - 1) / kc;
	{
		TimedLaunch t(h, st, K_FLUX);
		fused_flux_kernel<T, FNW><<<gx * gy * gz, dim3(FLANES, FNW, 1), 0, st>>>(h->g, S, out, h->emf,
			h->dt, next, gx, gy, kc, refits, h->tune.flux_remap);
	}
	HIP_TRY(h, hipGetLastError());
	return launch_ct_refit(h, st, s, out, refits);
}

/* the accumulating step: x, y and z marches into state[1 - cur], then the
   face-field update and refits */
template <class T>
int launch_accumulate(pamhd_handle* h, hipStream_t st, StepStatus* next, int refits)
{
	const long long n = h->g.n;
	if (!h->emf) HIP_TRY(h, hipMalloc(&h->emf, sizeof(double) * 6 * n));
	const double* s = h->state[h->cur];
	double* out = h->state[1 - h->cur];
	StateArrays S;
	for (int q = 0; q < 8; q++) S.q[q] = s + (size_t)q * n;
	const int km = std::min(std::max(1, h->tune.acc_km), ACC_KMAX);
	const unsigned gx = (h->g.nx + BX - 1) / BX;
	const unsigned tz = (h->g.nz + BY - 1) / BY, ty = (h->g.ny + BY - 1) / BY;
	const unsigned cy = (h->g.ny + km - 1) / km, cz = (h->g.nz + km - 1) / km;
	const int r = h->tune.flux_remap;
	const dim3 block(BX, BY, 1);
	{
		TimedLaunch t(h, st, K_FLUX);
		acc_x_kernel<T, 1><<<gx * tz * cy, block, 0, st>>>(h->g, S, out, h->emf, h->dt, next, gx, km, r);
		acc_march_kernel<T, 1, 3><<<gx * tz * cy, block, 0, st>>>(h->g, S, out, h->emf, h->dt, next,
			gx, km, r, refits);
		acc_march_kernel<T, 2, 3><<<gx * ty * cz, block, 0, st>>>(h->g, S, out, h->emf, h->dt, next,
			gx, km, r, refits);
	}
	HIP_TRY(h, hipGetLastError());
	return launch_ct_refit(h, st, s, out, refits);
}

/*
Flux pass and update of one step.  With tune.slabs = S > 1 the grid is cut
into S z-slabs: the flux pass of slab s runs on the caller's stream, and the
update of slab s (which reads the fluxes of slabs s-1, s and s+1) runs on a
second stream as soon as the flux pass of slab s+1 is done, so the
memory-bound update overlaps the compute-bound flux pass of later slabs.
Per-cell arithmetic is unchanged, so results are identical for every S.
*/
template <class T, class PT>
int launch_step_body(pamhd_handle* h, int solver, hipStream_t st, StepStatus* next, int refits)
{
	if (fused_step(h, solver)) {
		return hlld_mode(h) == 2 ? launch_accumulate<T>(h, st, next, refits)
			: launch_fused<T>(h, st, next, refits);
	}
	int S = h->tune.slabs;
	if (S > h->g.nz / 2) S = h->g.nz / 2;
	if (h->timing || h->tune.update_kz > 0) S = 1;
	if (S <= 1) {
		int e;
		{
			TimedLaunch t(h, st, K_FLUX);
			e = launch_flux_pass<T, PT>(h, solver, st, next);
		}
		if (e) return e;
		TimedLaunch t(h, st, K_UPDATE);
		launch_update<PT>(h, st, refits);
		return PAMHD_OK;
	}
	if (!h->aux) HIP_TRY(h, hipStreamCreateWithFlags(&h->aux, hipStreamNonBlocking));
	while ((int)h->ev.size() < S + 2) {
		hipEvent_t e;
		HIP_TRY(h, hipEventCreateWithFlags(&e, hipEventDisableTiming));
		h->ev.push_back(e);
	}
	hipEvent_t start = h->ev[S], join = h->ev[S + 1];
	HIP_TRY(h, hipEventRecord(start, st));
	HIP_TRY(h, hipStreamWaitEvent(h->aux, start, 0));
	auto k0 = [&](int q) { return (int)((long long)q * h->g.nz / S); };
	for (int q = 0; q < S; q++) {
		const int e = launch_flux_pass<T, PT>(h, solver, st, next, k0(q), k0(q + 1) - k0(q));
		if (e) return e;
		HIP_TRY(h, hipEventRecord(h->ev[q], st));
		if (q >= 2) {
			HIP_TRY(h, hipStreamWaitEvent(h->aux, h->ev[q], 0));
			launch_update<PT>(h, h->aux, refits, k0(q - 1), k0(q) - k0(q - 1));
		}
	}
	if (S == 2) HIP_TRY(h, hipStreamWaitEvent(h->aux, h->ev[1], 0));
	launch_update<PT>(h, h->aux, refits, k0(S - 1), k0(S) - k0(S - 1));
	launch_update<PT>(h, h->aux, refits, k0(0), k0(1) - k0(0));
	HIP_TRY(h, hipGetLastError());
	HIP_TRY(h, hipEventRecord(join, h->aux));
	HIP_TRY(h, hipStreamWaitEvent(st, join, 0));
	return PAMHD_OK;
}

/*
apply_boundaries on a grid with boundaries (source/mhd/boundaries.hpp:
1261-1297): value and copy boundaries of number density, velocity and
pressure (apply_fluid_boundaries, :738-1026), of the face field
(apply_magnetic_field_boundaries_staggered, :250-640, uniform grids), then
update_vol_B(0) at constant pressure of every cell of type >= 0.  One
launch per stage; a stage writes only boundary cells and reads only its
own cell and normal cells, which no stage writes.  Neighbours are visited
in the order of the neighbourhood (offsets -1..1, z slowest).
*/
struct BndNb { int id; int fn; int en0, en1, en2; };

template <class F> __device__ __forceinline__ void for_neighbours(const GridDev& g, const SubstepDev& d,
	int i, int j, int k, F&& f)
{
	for (int oz = -1; oz <= 1; oz++)
	for (int oy = -1; oy <= 1; oy++)
	for (int ox = -1; ox <= 1; ox++) {
		const int nzc = (ox != 0) + (oy != 0) + (oz != 0);
		if (nzc == 0) continue;
		const int a = ss_coord(i + ox, g.nx, d.per[0]);
		const int b = ss_coord(j + oy, g.ny, d.per[1]);
		const int c = ss_coord(k + oz, g.nz, d.per[2]);
		if (a < 0 || b < 0 || c < 0) continue;
		BndNb nb{cidx(g, a, b, c), 0, -1, 0, 0};
		if (nzc == 1) {
			nb.fn = ox ? ox : (oy ? 2 * oy : 3 * oz);
		} else if (nzc == 2) {
			/* edge: its dimension, then the offsets in the lower and the
			   higher other dimension */
			if (!ox) { nb.en0 = 0; nb.en1 = oy; nb.en2 = oz; }
			else if (!oy) { nb.en0 = 1; nb.en1 = ox; nb.en2 = oz; }
			else { nb.en0 = 2; nb.en1 = ox; nb.en2 = oy; }
		}
		f(nb);
	}
}

__device__ __forceinline__ int face_of(int dir) { return 2 * ((dir < 0 ? -dir : dir) - 1) + (dir > 0); }

__device__ __forceinline__ void bnd_velocity(const SubstepDev& d, long long n, int c, double v[3])
{
	if (d.S[c] > 0) {
		const double inv = 1 / d.S[c];
		for (int q = 0; q < 3; q++) v[q] = d.S[(1 + q) * n + c] * inv;
	} else {
		v[0] = v[1] = v[2] = 0;
	}
}

__device__ __forceinline__ double bnd_pressure(const SubstepDev& d, long long n, int c, const Phys& ph)
{
	double U[8];
	for (int q = 0; q < 8; q++) U[q] = d.S[q * n + c];
	return pressure64(U, ph.gm1, ph.half_inv_mu0);
}

/* STAGE 0 number density, 1 velocity, 2 pressure */
template <int STAGE>
__global__ __launch_bounds__(BX * BY) void bnd_fluid_kernel(GridDev g, SubstepDev d, double proton_mass)
{
	const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y, k = blockIdx.z;
	if (i >= g.nx || j >= g.ny) return;
	const long long n = g.n;
	const int c = cidx(g, i, j, k);
	if (d.bmask[STAGE * n + c]) {
		if (STAGE == 0) {
			d.S[c] = proton_mass * d.bval[c];
		} else if (STAGE == 1) {
			for (int q = 0; q < 3; q++) d.S[(1 + q) * n + c] = d.bval[(1 + q) * n + c] * d.S[c];
		} else {
			if (d.S[c] <= 0) { record_error(d.status, PAMHD_E_DENSITY, 2 * n + c); return; }
			double v[3];
			bnd_velocity(d, n, c, v);
			const double B[3] = {d.S[5 * n + c], d.S[6 * n + c], d.S[7 * n + c]};
			d.S[4 * n + c] = total_energy64(d.S[c], v, d.bval[4 * n + c], B, g.ph.gm1, g.ph.half_inv_mu0);
		}
	}
	if (d.vt[STAGE * n + c] != 4) return;
	/* copy: the average over normal face neighbours, else over normal
	   edge neighbours */
	double fs[3] = {0, 0, 0}, es[3] = {0, 0, 0};
	unsigned long long nf = 0, ne = 0;
	bool bad = false;
	for_neighbours(g, d, i, j, k, [&](const BndNb& b) {
		if (d.ct[b.id] < 1) return;
		double v[3] = {0, 0, 0};
		if (b.fn != 0 || b.en0 >= 0) {
			if (STAGE == 0) v[0] = d.S[b.id];
			else if (STAGE == 1) bnd_velocity(d, n, b.id, v);
			else {
				if (d.S[b.id] <= 0) bad = true;
				else v[0] = bnd_pressure(d, n, b.id, g.ph);
			}
		}
		if (b.fn != 0) {
			for (int q = 0; q < 3; q++) fs[q] = fs[q] + v[q];
			nf++;
		}
		if (b.en0 >= 0) {
			for (int q = 0; q < 3; q++) es[q] = es[q] + v[q];
			ne++;
		}
	});
	if (bad) { record_error(d.status, PAMHD_E_DENSITY, 2 * n + c); return; }
	if (nf == 0 && ne == 0) return;
	const double* src = nf > 0 ? fs : es;
	const unsigned long long cnt = nf > 0 ? nf : ne;
	if (STAGE == 0) {
		d.S[c] = src[0] / cnt;
	} else if (STAGE == 1) {
		/* mul(mul(Mas, sum), 1 / count): an integer quotient, 1 for one
		   source and 0 for more */
		const double f = (double)(1ull / cnt);
		for (int q = 0; q < 3; q++) d.S[(1 + q) * n + c] = src[q] * d.S[c] * f;
	} else {
		if (d.S[c] <= 0) { record_error(d.status, PAMHD_E_DENSITY, 2 * n + c); return; }
		double v[3];
		bnd_velocity(d, n, c, v);
		const double B[3] = {d.S[5 * n + c], d.S[6 * n + c], d.S[7 * n + c]};
		d.S[4 * n + c] = total_energy64(d.S[c], v, src[0] / cnt, B, g.ph.gm1, g.ph.half_inv_mu0);
	}
}

__global__ __launch_bounds__(BX * BY) void bnd_field_value_kernel(GridDev g, SubstepDev d)
{
	const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y, k = blockIdx.z;
	if (i >= g.nx || j >= g.ny) return;
	const long long n = g.n;
	const int c = cidx(g, i, j, k);
	if (!d.bmask[3 * n + c]) return;
	for (int f = 0; f < 6; f++) {
		if (d.fi[f * n + c] == 0) d.S[(8 + f) * n + c] = d.bval[(5 + f) * n + c];
	}
}

__global__ __launch_bounds__(BX * BY) void bnd_field_copy_kernel(GridDev g, SubstepDev d)
{
	const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y, k = blockIdx.z;
	if (i >= g.nx || j >= g.ny) return;
	const long long n = g.n;
	const int c = cidx(g, i, j, k);
	if (d.vt[3 * n + c] != 4) return;
	double F[6];
	int cnt[6] = {0, 0, 0, 0, 0, 0};
	signed char fi[6];
	for (int f = 0; f < 6; f++) {
		fi[f] = d.fi[f * n + c];
		F[f] = d.S[(8 + f) * n + c];
	}
	for (int q3 = 0; q3 < 3; q3++) {
		for (int s = 0; s < 2; s++) if (fi[2 * q3 + s] < 1) F[2 * q3 + s] = 0;
		if (fi[2 * q3] == 0 && fi[2 * q3 + 1] == 1) { F[2 * q3] += F[2 * q3 + 1]; cnt[2 * q3]++; }
		if (fi[2 * q3] == 1 && fi[2 * q3 + 1] == 0) { F[2 * q3 + 1] += F[2 * q3]; cnt[2 * q3 + 1]++; }
	}
	for_neighbours(g, d, i, j, k, [&](const BndNb& b) {
		if (b.fn == 0 && b.en0 < 0) return;
		if (b.fn != 0) {
			for (int dir = -3; dir <= 3; dir++) {
				if (dir == 0) continue;
				const int f = face_of(dir);
				if (fi[f] == 0 && d.fi[f * n + b.id] == 1) { F[f] += d.S[(8 + f) * n + b.id]; cnt[f]++; }
			}
		}
		if (b.en0 >= 0) {
			const int d1 = b.en0 == 0 ? 1 : 0, d2 = b.en0 == 2 ? 1 : 2;
			const int dirs[2] = {b.en1 * (d1 + 1), b.en2 * (d2 + 1)};
			for (int r = 0; r < 2; r++) {
				const int f = face_of(dirs[r]), fo = face_of(-dirs[r]);
				if (fi[f] == 0 && d.fi[fo * n + b.id] == 1) { F[f] += d.S[(8 + fo) * n + b.id]; cnt[f]++; }
			}
		}
	});
	for (int f = 0; f < 6; f++) {
		if (fi[f] == 0) {
			if (cnt[f] == 0) { record_error(d.status, PAMHD_E_SOLVER, (long long)f * n + c); return; }
			F[f] /= cnt[f];
		}
		d.S[(8 + f) * n + c] = F[f];
	}
}

/* update_vol_B(0) with constant pressure of the cells of type >= 0 */
__global__ __launch_bounds__(BX * BY) void bnd_refit_kernel(GridDev g, SubstepDev d)
{
	const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y, k = blockIdx.z;
	if (i >= g.nx || j >= g.ny) return;
	const long long n = g.n;
	const int c = cidx(g, i, j, k);
	if (!ss_active(d, c)) return;
	const double rho = d.S[c];
	if (rho <= 0) return;
	const double m[3] = {d.S[n + c], d.S[2 * n + c], d.S[3 * n + c]};
	double En = d.S[4 * n + c];
	const double vb[3] = {d.S[5 * n + c], d.S[6 * n + c], d.S[7 * n + c]};
	const double vb_new[3] = {
		0.5 * (d.S[8 * n + c] + d.S[9 * n + c]),
		0.5 * (d.S[10 * n + c] + d.S[11 * n + c]),
		0.5 * (d.S[12 * n + c] + d.S[13 * n + c])};
	refit_energy(rho, m, En, vb, vb_new, g.ph);
	for (int q = 0; q < 3; q++) d.S[(5 + q) * n + c] = vb_new[q];
	d.S[4 * n + c] = En;
}

void substep_free(SubstepDev* d)
{
	if (!d) return;
	(void)hipFree(d->S);
	(void)hipFree(d->slot);
	(void)hipFree(d->fdb);
	(void)hipFree(d->maxv);
	(void)hipFree(d->ts);
	(void)hipFree(d->e[0]);
	(void)hipFree(d->e[1]);
	(void)hipFree(d->period);
	(void)hipFree(d->minsub);
	(void)hipFree(d->blk);
	(void)hipFree(d->dirty);
	(void)hipFree(d->list);
	(void)hipFree(d->count);
	(void)hipFree(d->opt_dev);
	(void)hipFree(d->types);
	(void)hipFree(d->bmask);
	(void)hipFree(d->bval);
	(void)hipFree(d->scal);
	(void)hipFree(d->sub_dt);
	(void)hipFree(d->status);
	(void)hipHostFree(d->host);
	(void)hipHostFree(d->host_status);
	delete d;
}

int substep_alloc(pamhd_handle* h)
{
	const size_t n = (size_t)h->g.n;
	SubstepDev* d = new SubstepDev();
	hipError_t e = hipSuccess;
	auto al = [&](auto** p, size_t bytes) {
		if (e == hipSuccess) e = hipMalloc((void**)p, bytes);
	};
	al(&d->S, 14 * n * sizeof(double));
	al(&d->slot, 48 * n * sizeof(double));
	al(&d->fdb, 9 * n * sizeof(double));
	al(&d->maxv, 3 * n * sizeof(double));
	al(&d->ts, n * sizeof(double));
	al(&d->e[0], n * sizeof(int));
	al(&d->e[1], n * sizeof(int));
	al(&d->period, n * sizeof(int));
	al(&d->minsub, n * sizeof(int));
	al(&d->opt_dev, 2 * n * sizeof(int));
	{
		const dim3 cg = cell_grid(h->g);
		d->nblk = (int)(cg.x * cg.y * cg.z);
	}
	d->walk = h->ss_walk;
	al(&d->blk, 4 * (size_t)d->nblk * sizeof(int));
	al(&d->dirty, 3 * (size_t)d->nblk * sizeof(int));
	al(&d->list, SS_LISTS * (size_t)d->nblk * sizeof(int));
	al(&d->count, 2 * SS_LISTS * sizeof(int));
	al(&d->scal, 5 * SS_SLOTS * sizeof(unsigned long long));
	al(&d->sub_dt, sizeof(double));
	al(&d->status, sizeof(StepStatus));
	if (e == hipSuccess) e = hipHostMalloc((void**)&d->host, 3 * sizeof(unsigned long long));
	if (e == hipSuccess) e = hipHostMalloc((void**)&d->host_status, sizeof(StepStatus));
	if (e != hipSuccess) {
		substep_free(d);
		return hip_fail(h, e, "substepping buffers");
	}
	/* slots and face dB start at zero and are zero again after every step;
	   speeds are unknown until the next flux pass */
	const unsigned blocks = (unsigned)((48 * n + 255) / 256);
	ss_fill_kernel<<<blocks, dim3(BX, BY, 1)>>>(d->slot, (long long)(48 * n), 0.0);
	ss_fill_kernel<<<(unsigned)((9 * n + 255) / 256), dim3(BX, BY, 1)>>>(d->fdb, (long long)(9 * n), 0.0);
	ss_fill_kernel<<<(unsigned)((3 * n + 255) / 256), dim3(BX, BY, 1)>>>(d->maxv, (long long)(3 * n), -1.0);
	HIP_TRY(h, hipGetLastError());
	HIP_TRY(h, hipDeviceSynchronize());
	h->sub = d;
	return PAMHD_OK;
}

/* grid of the passes that walk a block list */
dim3 ss_walk_grid(const SubstepDev& d)
{
	const int w = d.walk > 0 ? d.walk : SS_WALK_BLOCKS;
	return dim3((unsigned)(d.nblk < w ? d.nblk : w), 1, 1);
}

/* the flux passes of substep `sub`; at a substep where every face is due
   (full) they cover the whole cell grid, and no edge field needs zeroing */
template <int SOLVER>
void ss_launch_flux(pamhd_handle* h, double* P, int sub, bool full, hipStream_t st)
{
	const SubstepDev& d = *h->sub;
	const dim3 block(BX, BY, 1);
	if (full) {
		const dim3 grid = cell_grid(h->g);
		ss_flux_kernel<SOLVER, 0, false><<<grid, block, 0, st>>>(h->g, d, h->bg, P, sub);
		if constexpr (SOLVER == HLLD_ATHENA) {
			if (!d.bounded && !d.ct) {
				/* HLLD's per-cell part carried along marches of 16 cells */
				const int km = 16;
				const unsigned gx = (h->g.nx + BX - 1) / BX;
				const unsigned g1 = gx * ((h->g.nz + BY - 1) / BY) * ((h->g.ny + km - 1) / km);
				const unsigned g2 = gx * ((h->g.ny + BY - 1) / BY) * ((h->g.nz + km - 1) / km);
				ss_flux_march_kernel<1><<<g1, block, 0, st>>>(h->g, d, P, sub, gx, km);
				ss_flux_march_kernel<2><<<g2, block, 0, st>>>(h->g, d, P, sub, gx, km);
				return;
			}
		}
		ss_flux_kernel<SOLVER, 1, false><<<grid, block, 0, st>>>(h->g, d, h->bg, P, sub);
		ss_flux_kernel<SOLVER, 2, false><<<grid, block, 0, st>>>(h->g, d, h->bg, P, sub);
		return;
	}
	const dim3 grid = ss_walk_grid(d);
	ss_zero_emf_kernel<0><<<grid, block, 0, st>>>(h->g, d, P, sub);
	ss_zero_emf_kernel<1><<<grid, block, 0, st>>>(h->g, d, P, sub);
	ss_zero_emf_kernel<2><<<grid, block, 0, st>>>(h->g, d, P, sub);
	ss_flux_kernel<SOLVER, 0, true><<<grid, block, 0, st>>>(h->g, d, h->bg, P, sub);
	ss_flux_kernel<SOLVER, 1, true><<<grid, block, 0, st>>>(h->g, d, h->bg, P, sub);
	ss_flux_kernel<SOLVER, 2, true><<<grid, block, 0, st>>>(h->g, d, h->bg, P, sub);
}

/* apply_boundaries and the main loop's sync_magnetic_field on a grid with
   boundaries */
void launch_boundaries(pamhd_handle* h, hipStream_t st)
{
	SubstepDev& d = *h->sub;
	const dim3 grid = cell_grid(h->g), block(BX, BY, 1);
	const double mp = h->phys.proton_mass;
	bnd_fluid_kernel<0><<<grid, block, 0, st>>>(h->g, d, mp);
	bnd_fluid_kernel<1><<<grid, block, 0, st>>>(h->g, d, mp);
	bnd_fluid_kernel<2><<<grid, block, 0, st>>>(h->g, d, mp);
	bnd_field_value_kernel<<<grid, block, 0, st>>>(h->g, d);
	bnd_field_copy_kernel<<<grid, block, 0, st>>>(h->g, d);
	bnd_refit_kernel<<<grid, block, 0, st>>>(h->g, d);
	ss_sync_kernel<0><<<grid, block, 0, st>>>(h->g, d);
	ss_sync_kernel<1><<<grid, block, 0, st>>>(h->g, d);
	ss_sync_kernel<2><<<grid, block, 0, st>>>(h->g, d);
}

/* timestep() with per-cell substepping (and apply_boundaries' refit when
   refits == 3), in place on the current state */
int substep_step(pamhd_handle* h, int solver, double max_dt, double cfl, int refits,
	double* dt_taken, hipStream_t st)
{
	SubstepDev& d = *h->sub;
	const dim3 grid = cell_grid(h->g), block(BX, BY, 1);
	double* P = h->flux;   /* the substep's field components per face, 6 x n (of 24 x n) */
	ss_clear_kernel<<<1, SS_FOLD, 0, st>>>(d);
	ss_ts_kernel<<<grid, block, 0, st>>>(h->g, d, cfl);
	ss_fold_kernel<<<1, SS_FOLD, 0, st>>>(d, 0, 1);
	ss_fold_kernel<<<1, SS_FOLD, 0, st>>>(d, 2, 1);
	ss_level_kernel<<<grid, block, 0, st>>>(h->g, d, max_dt);
	ss_fold_kernel<<<1, SS_FOLD, 0, st>>>(d, 3, 0);
	HIP_TRY(h, hipGetLastError());
	HIP_TRY(h, hipMemcpyAsync(d.host + 2, d.scal + 3 * SS_SLOTS, sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
	HIP_TRY(h, hipMemcpyAsync(d.host_status, d.status, sizeof(StepStatus), hipMemcpyDeviceToHost, st));
	HIP_TRY(h, hipStreamSynchronize(st));
	if (d.host_status->err) return report_status(h, *d.host_status);
	/* Jacobi passes of the restriction: a cell's restricted level is the
	   smallest over cells m of level(m) + distance, reached within as many
	   passes as the largest level */
	const int passes = (int)d.host[2];
	int cur = 0;
	for (int r = 0; r < passes; r++) {
		ss_restrict_kernel<<<grid, block, 0, st>>>(h->g, d, d.e[cur], d.e[1 - cur]);
		cur = 1 - cur;
	}
	ss_period_kernel<<<grid, block, 0, st>>>(h->g, d, d.e[cur]);
	ss_fold_kernel<<<1, SS_FOLD, 0, st>>>(d, 1, 0);
	ss_blockmask_kernel<<<grid, block, 0, st>>>(h->g, d);
	HIP_TRY(h, hipGetLastError());
	HIP_TRY(h, hipMemcpyAsync(d.host, d.scal + 1 * SS_SLOTS, sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
	HIP_TRY(h, hipMemcpyAsync(d.host + 1, d.sub_dt, sizeof(double), hipMemcpyDeviceToHost, st));
	HIP_TRY(h, hipStreamSynchronize(st));
	const int maxp = (int)d.host[0];
	double sub_dt;
	memcpy(&sub_dt, d.host + 1, sizeof(double));
	if (maxp < 1) return fail(h, PAMHD_E_DEVICE, "substep periods not computed");

	/* the working layout is still the expansion of the state when the last
	   substepped step made it and nothing else reached the handle since
	   (this call and, in pamhd_run's loop, none) */
	const bool live = !d.bounded && h->ss_valid && !h->state_escaped && h->calls - h->ss_at <= 1;
	h->ss_valid = false;
	if (!d.bounded && !live) ss_expand_kernel<<<grid, block, 0, st>>>(h->g, h->state[h->cur], d);
	double total = 0;
	for (int sub = 1; sub <= maxp; sub++) {
		total += sub_dt;
		/* the last substep has every face and cell due (periods are powers
		   of two up to maxp): its passes cover the whole grid; the others
		   walk their lists */
		const bool full = sub == maxp;
		ss_lists_kernel<<<(unsigned)((d.nblk + 255) / 256), 256, 0, st>>>(h->g, d, sub);
		switch (solver) {
		case RUSANOV: ss_launch_flux<RUSANOV>(h, P, sub, full, st); break;
		case HLL_ATHENA: ss_launch_flux<HLL_ATHENA>(h, P, sub, full, st); break;
		case HLLD_ATHENA: ss_launch_flux<HLLD_ATHENA>(h, P, sub, full, st); break;
		case ROE_ATHENA: ss_launch_flux<ROE_ATHENA>(h, P, sub, full, st); break;
		default: return fail(h, PAMHD_E_SOLVER, "invalid solver " + std::to_string(solver));
		}
		if (full) {
			ss_dbsum_kernel<false, false><<<grid, block, 0, st>>>(h->g, d, P, sub);
		} else {
			ss_dbsum_kernel<true, true><<<ss_walk_grid(d), block, 0, st>>>(h->g, d, P, sub);
			ss_dbsum_kernel<true, false><<<ss_walk_grid(d), block, 0, st>>>(h->g, d, P, sub);
		}
		if (full) ss_update_kernel<false><<<grid, block, 0, st>>>(h->g, d, sub);
		else ss_update_kernel<true><<<ss_walk_grid(d), block, 0, st>>>(h->g, d, sub);
	}
	ss_sync_kernel<0><<<grid, block, 0, st>>>(h->g, d);
	ss_sync_kernel<1><<<grid, block, 0, st>>>(h->g, d);
	ss_sync_kernel<2><<<grid, block, 0, st>>>(h->g, d);
	/* on a periodic grid the final kernel also does apply_boundaries'
	   refit (refits == 3) */
	ss_final_kernel<<<grid, block, 0, st>>>(h->g, d, h->state[h->cur], !d.bounded && refits == 3);
	if (!d.bounded) {
		h->ss_valid = true;
		h->ss_at = h->calls;
	}
	ss_count_kernel<<<1, SS_FOLD, 0, st>>>(d);
	if (refits == 3 && d.bounded) launch_boundaries(h, st);
	HIP_TRY(h, hipGetLastError());
	HIP_TRY(h, hipMemcpyAsync(d.host_status, d.status, sizeof(StepStatus), hipMemcpyDeviceToHost, st));
	HIP_TRY(h, hipStreamSynchronize(st));
	d.last_max = maxp;
	h->flux_calcs = d.host_status->flux_calcs;
	if (dt_taken) *dt_taken = total;
	return report_status(h, *d.host_status);
}

int do_step(pamhd_handle* h, int solver, double max_dt, double cfl, int refits,
	double* dt_taken, void* stream_, int fixed_dt = 0)
{
	if (!h) return fail(nullptr, PAMHD_E_ARG, "null handle");
	if (solver == HYBRID) return fail(h, PAMHD_E_UNSUPPORTED, "hybrid solver: MHD fluxes are not computed");
	if (solver != RUSANOV && solver != HLL_ATHENA && solver != HLLD_ATHENA && solver != ROE_ATHENA) {
		return fail(h, PAMHD_E_SOLVER, "invalid solver " + std::to_string(solver));
	}
	if (!(cfl >= 0)) return fail(h, PAMHD_E_ARG, "time step factor must be >= 0");
	hipStream_t st = (hipStream_t)stream_;
	h->flux_step = 0;
	if (h->box) {
		/* timestep() alone, or with the main loop's apply_boundaries_sw_box
		   and sync_magnetic_field (tests/mhd/solar_wind_box.cpp:522-553) */
		std::string err;
		int e = boxset_launch_step(h->box, solver, max_dt, cfl, refits == 3, st, err);
		if (e) return fail(h, e, err);
		h->flux_calcs = (unsigned long long)boxset_faces(h->box, 0);
		if (dt_taken) {
			e = boxset_readback(h->box, st, dt_taken, err);
			if (e) return fail(h, e, err);
		}
		return PAMHD_OK;
	}
	if (h->sub && h->sub->bounded) {
		if (fixed_dt) return fail(h, PAMHD_E_UNSUPPORTED, "fixed steps on a grid with boundaries");
		return substep_step(h, solver, max_dt, cfl, refits, dt_taken, st);
	}
	/* once substepping is set up, its own per-face speeds carry the lagged
	   time step from step to step (periods of 1 give the global step) */
	if (h->sub && !fixed_dt) {
		return substep_step(h, solver, max_dt, cfl, refits, dt_taken, st);
	}
	StepStatus* prev = h->status + h->last;
	StepStatus* next = h->status + (1 - h->last);
	{
		TimedLaunch t(h, st, K_DT);
		dt_kernel<<<1, NSLOT, 0, st>>>(prev, next, h->dt, cfl, max_dt, h->g.dx, h->g.dy, h->g.dz, fixed_dt);
	}
	HIP_TRY(h, hipGetLastError());
	const int e = h->precision == PAMHD_MIXED
		? launch_step_body<float, double>(h, solver, st, next, refits)
		: launch_step_body<double, double>(h, solver, st, next, refits);
	if (e) return e;
	HIP_TRY(h, hipGetLastError());
	h->flux_step = (h->precision == PAMHD_FP64 && !fused_step(h, solver)) ? 1 : 0;
	h->cur = 1 - h->cur;
	h->last = 1 - h->last;
	h->flux_calcs = 3ull * (unsigned long long)h->g.n;
	if (dt_taken) {
		HIP_TRY(h, hipMemcpyAsync(h->host_dt, h->dt, sizeof(double), hipMemcpyDeviceToHost, st));
		HIP_TRY(h, hipMemcpyAsync(h->host_status, next, sizeof(StepStatus), hipMemcpyDeviceToHost, st));
		HIP_TRY(h, hipStreamSynchronize(st));
		*dt_taken = *h->host_dt;
		return report_status(h, *h->host_status);
	}
	return PAMHD_OK;
}

} // namespace

extern "C" {

int pamhd_abi_version(void) { return 1; }

int pamhd_create(const pamhd_grid_desc* grid, const pamhd_physics* physics, int precision,
	pamhd_handle** out)
{
	if (!grid || !physics || !out) return fail(nullptr, PAMHD_E_ARG, "null argument");
	*out = nullptr;
	for (int d = 0; d < 3; d++) {
		if (grid->cells[d] < 1) return fail(nullptr, PAMHD_E_ARG, "cells must be >= 1");
		if (!(grid->cell_length[d] > 0)) return fail(nullptr, PAMHD_E_ARG, "cell length must be > 0");
		if (!grid->periodic[d]) {
			return fail(nullptr, PAMHD_E_UNSUPPORTED, "non-periodic grids need boundary cells; not supported");
		}
	}
	/* get_total_energy_density's checks, source/mhd/common.hpp:275-289 */
	if (!(physics->adiabatic_index > 1)) return fail(nullptr, PAMHD_E_ARG, "adiabatic index must be > 1");
	if (!(physics->vacuum_permeability > 0)) return fail(nullptr, PAMHD_E_ARG, "vacuum permeability must be > 0");
	if (precision != PAMHD_FP64 && precision != PAMHD_MIXED) return fail(nullptr, PAMHD_E_ARG, "invalid precision");
	const long long n = (long long)grid->cells[0] * grid->cells[1] * grid->cells[2];
	if (n >= (1ll << 28)) return fail(nullptr, PAMHD_E_UNSUPPORTED, "more than 2^28 cells per grid");

	pamhd_handle* h = new pamhd_handle();
	h->desc = *grid;
	h->phys = *physics;
	h->precision = precision;
	h->g = GridDev{grid->cells[0], grid->cells[1], grid->cells[2], n,
		grid->cell_length[0], grid->cell_length[1], grid->cell_length[2],
		make_phys(physics->adiabatic_index, physics->vacuum_permeability)};
	/* thin grids: the x/y marches put one wave per block on each of fewer
	   than BY planes and the column update walks fewer planes than its
	   setup loads, so their defaults are per face and per cell */
	if (grid->cells[2] < 4 * BY) {
		h->tune.flux_remap = 0;
		h->tune.flux_shared = 0;
		h->tune.update_kz = 0;
		h->tune.update_tile = 0;
	}
	auto cleanup = [&](int code, const char* what, hipError_t e) {
		std::string m = std::string(what) + ": " + hipGetErrorString(e);
		pamhd_destroy(h);
		return fail(nullptr, code, m);
	};
	hipError_t e;
	if ((e = hipMalloc(&h->state[0], sizeof(double) * 8 * n)) != hipSuccess) return cleanup(PAMHD_E_DEVICE, "state", e);
	if ((e = hipMalloc(&h->state[1], sizeof(double) * 8 * n)) != hipSuccess) return cleanup(PAMHD_E_DEVICE, "state", e);
	if ((e = hipMalloc(&h->flux, sizeof(double) * 24 * n)) != hipSuccess) return cleanup(PAMHD_E_DEVICE, "flux", e);
	if ((e = hipMalloc(&h->dt, sizeof(double))) != hipSuccess) return cleanup(PAMHD_E_DEVICE, "dt", e);
	if ((e = hipMalloc(&h->status, 2 * sizeof(StepStatus))) != hipSuccess) return cleanup(PAMHD_E_DEVICE, "status", e);
	if ((e = hipHostMalloc(&h->host_dt, sizeof(double))) != hipSuccess) return cleanup(PAMHD_E_DEVICE, "pinned", e);
	if ((e = hipHostMalloc(&h->host_status, sizeof(StepStatus))) != hipSuccess) return cleanup(PAMHD_E_DEVICE, "pinned", e);
	/* before the first flux pass every face's Max_v is -1
	   (tests/mhd/test.cpp:383), so the first lagged step is 0 */
	StepStatus init[2];
	for (auto& s : init) {
		for (int d = 0; d < 3; d++) {
			for (int q = 0; q < NSLOT; q++) s.smax_bits[d][q] = 0;
			s.zero[d] = 1;
		}
		s.err = 0;
		s.err_face = -1;
	}
	if ((e = hipMemcpy(h->status, init, sizeof(init), hipMemcpyHostToDevice)) != hipSuccess) return cleanup(PAMHD_E_DEVICE, "status", e);
	if ((e = hipMemset(h->state[0], 0, sizeof(double) * 8 * n)) != hipSuccess) return cleanup(PAMHD_E_DEVICE, "state", e);
	*out = h;
	return PAMHD_OK;
}

int pamhd_create_solar_wind_box(const pamhd_grid_desc* grid, const pamhd_physics* physics,
	const pamhd_solar_wind_desc* sw, pamhd_handle** out)
{
	if (!grid || !physics || !sw || !out) return fail(nullptr, PAMHD_E_ARG, "null argument");
	*out = nullptr;
	BoxSet* set = nullptr;
	std::string err;
	const int e = boxset_create(&set, 1, grid, physics, sw, err);
	if (e) return fail(nullptr, e, err);
	pamhd_handle* h = new pamhd_handle();
	h->desc = *grid;
	h->phys = *physics;
	h->box = set;
	h->g = GridDev{grid->cells[0], grid->cells[1], grid->cells[2],
		(long long)grid->cells[0] * grid->cells[1] * grid->cells[2],
		grid->cell_length[0], grid->cell_length[1], grid->cell_length[2],
		make_phys(physics->adiabatic_index, physics->vacuum_permeability)};
	*out = h;
	return PAMHD_OK;
}

int pamhd_create_bounded(const pamhd_grid_desc* grid, const pamhd_physics* physics,
	const signed char* var_type, const signed char* cell_type, const signed char* face_type,
	pamhd_handle** out)
{
	if (!grid || !physics || !var_type || !cell_type || !face_type || !out) {
		return fail(nullptr, PAMHD_E_ARG, "null argument");
	}
	*out = nullptr;
	pamhd_grid_desc periodic = *grid;
	for (int d = 0; d < 3; d++) periodic.periodic[d] = 1;
	pamhd_handle* h = nullptr;
	int e = pamhd_create(&periodic, physics, PAMHD_FP64, &h);
	if (e) return e;
	h->desc = *grid;
	e = substep_alloc(h);
	if (e) { pamhd_destroy(h); return e; }
	SubstepDev& d = *h->sub;
	const long long n = h->g.n;
	hipError_t he = hipMalloc(&d.types, (size_t)11 * n);
	if (he == hipSuccess) he = hipMalloc(&d.bmask, (size_t)4 * n);
	if (he == hipSuccess) he = hipMalloc(&d.bval, sizeof(double) * 11 * n);
	if (he == hipSuccess) he = hipMemcpy(d.types, cell_type, (size_t)n, hipMemcpyHostToDevice);
	if (he == hipSuccess) he = hipMemcpy(d.types + n, face_type, (size_t)6 * n, hipMemcpyHostToDevice);
	if (he == hipSuccess) he = hipMemcpy(d.types + 7 * n, var_type, (size_t)4 * n, hipMemcpyHostToDevice);
	if (he == hipSuccess) he = hipMemset(d.bmask, 0, (size_t)4 * n);
	if (he == hipSuccess) he = hipMemset(d.S, 0, sizeof(double) * 14 * n);
	if (he != hipSuccess) {
		const int r = fail(nullptr, PAMHD_E_DEVICE, std::string("boundary arrays: ") + hipGetErrorString(he));
		pamhd_destroy(h);
		return r;
	}
	d.bounded = true;
	d.ct = d.types;
	d.fi = d.types + n;
	d.vt = d.types + 7 * n;
	for (int q = 0; q < 3; q++) d.per[q] = grid->periodic[q] ? 1 : 0;
	d.smax = 0;
	d.opt = nullptr;
	*out = h;
	return PAMHD_OK;
}

int pamhd_set_boundary_values(pamhd_handle* h, const unsigned char* mask, const double* number_density,
	const double* velocity, const double* pressure, const double* face_field)
{
	touch(h);
	if (!h || !mask || !number_density || !velocity || !pressure || !face_field) {
		return fail(h, PAMHD_E_ARG, "null argument");
	}
	if (!h->sub || !h->sub->bounded) return fail(h, PAMHD_E_ARG, "not a grid with boundaries");
	SubstepDev& d = *h->sub;
	const size_t n = (size_t)h->g.n;
	HIP_TRY(h, hipMemcpy(d.bmask, mask, 4 * n, hipMemcpyHostToDevice));
	HIP_TRY(h, hipMemcpy(d.bval, number_density, sizeof(double) * n, hipMemcpyHostToDevice));
	HIP_TRY(h, hipMemcpy(d.bval + n, velocity, sizeof(double) * 3 * n, hipMemcpyHostToDevice));
	HIP_TRY(h, hipMemcpy(d.bval + 4 * n, pressure, sizeof(double) * n, hipMemcpyHostToDevice));
	HIP_TRY(h, hipMemcpy(d.bval + 5 * n, face_field, sizeof(double) * 6 * n, hipMemcpyHostToDevice));
	return PAMHD_OK;
}

int pamhd_box_cell_types(pamhd_handle* h, signed char* cell_type, unsigned char* rule)
{
	touch(h);
	if (!h) return fail(nullptr, PAMHD_E_ARG, "null handle");
	if (!h->box) return fail(h, PAMHD_E_ARG, "not a solar-wind box");
	HIP_TRY(h, boxset_cell_types(h->box, 0, cell_type, rule));
	return PAMHD_OK;
}

int pamhd_box_background(pamhd_handle* h, double* bg)
{
	touch(h);
	if (!h || !bg) return fail(h, PAMHD_E_ARG, "null argument");
	if (!h->box) return fail(h, PAMHD_E_ARG, "not a solar-wind box");
	HIP_TRY(h, boxset_background(h->box, 0, bg));
	return PAMHD_OK;
}

int pamhd_destroy(pamhd_handle* h)
{
	touch(h);
	if (!h) return PAMHD_OK;
	collect_times(h);
	(void)hipFree(h->state[0]);
	(void)hipFree(h->state[1]);
	(void)hipFree(h->flux);
	(void)hipFree(h->emf);
	(void)hipFree(h->bg);
	(void)hipFree(h->dt);
	(void)hipFree(h->status);
	(void)hipFree(h->scratch);
	if (h->mg) {
		mg_destroy(*h->mg);
		delete h->mg;
	}
	boxset_destroy(h->box);
	substep_free(h->sub);
	(void)hipFree(h->adj);
	for (auto e : h->ev) (void)hipEventDestroy(e);
	if (h->aux) (void)hipStreamDestroy(h->aux);
	(void)hipHostFree(h->host_dt);
	(void)hipHostFree(h->host_status);
	delete h;
	return PAMHD_OK;
}

int pamhd_upload_state(pamhd_handle* h, const double* state)
{
	touch(h);
	if (!h || !state) return fail(h, PAMHD_E_ARG, "null argument");
	if (h->box) {
		HIP_TRY(h, boxset_upload(h->box, 0, state));
		return PAMHD_OK;
	}
	if (h->sub && h->sub->bounded) {
		HIP_TRY(h, hipMemcpy(h->sub->S, state, sizeof(double) * 14 * h->g.n, hipMemcpyHostToDevice));
		return PAMHD_OK;
	}
	HIP_TRY(h, hipMemcpy(h->state[h->cur], state, sizeof(double) * 8 * h->g.n, hipMemcpyHostToDevice));
	return PAMHD_OK;
}

int pamhd_download_state(pamhd_handle* h, double* state)
{
	touch(h);
	if (!h || !state) return fail(h, PAMHD_E_ARG, "null argument");
	if (h->box) {
		HIP_TRY(h, boxset_download(h->box, 0, state));
		return PAMHD_OK;
	}
	if (h->sub && h->sub->bounded) {
		HIP_TRY(h, hipDeviceSynchronize());
		HIP_TRY(h, hipMemcpy(state, h->sub->S, sizeof(double) * 14 * h->g.n, hipMemcpyDeviceToHost));
		return PAMHD_OK;
	}
	HIP_TRY(h, hipMemcpy(state, h->state[h->cur], sizeof(double) * 8 * h->g.n, hipMemcpyDeviceToHost));
	return PAMHD_OK;
}

int pamhd_state_components(pamhd_handle* h, int* components)
{
	touch(h);
	if (!h || !components) return fail(h, PAMHD_E_ARG, "null argument");
	*components = (h->box || (h->sub && h->sub->bounded)) ? 14 : 8;
	return PAMHD_OK;
}

int pamhd_grid_kind(pamhd_handle* h, int* kind)
{
	touch(h);
	if (!h || !kind) return fail(h, PAMHD_E_ARG, "null argument");
	*kind = h->box ? PAMHD_GRID_SOLAR_WIND_BOX : ((h->sub && h->sub->bounded) ? PAMHD_GRID_BOUNDED : PAMHD_GRID_PERIODIC);
	return PAMHD_OK;
}

int pamhd_state_device_ptr(pamhd_handle* h, double** ptr)
{
	touch(h);
	if (h) h->state_escaped = true;
	if (!h || !ptr) return fail(h, PAMHD_E_ARG, "null argument");
	if (h->box) return fail(h, PAMHD_E_UNSUPPORTED, "solar-wind box state is not a flat array");
	if (h->sub && h->sub->bounded) {
		*ptr = h->sub->S;
		return PAMHD_OK;
	}
	*ptr = h->state[h->cur];
	return PAMHD_OK;
}

int pamhd_set_background_faces(pamhd_handle* h, const double* bg)
{
	touch(h);
	if (!h) return fail(nullptr, PAMHD_E_ARG, "null handle");
	if (h->box) return fail(h, PAMHD_E_UNSUPPORTED, "solar-wind boxes compute their background field");
	if (!bg) {
		(void)hipFree(h->bg);
		h->bg = nullptr;
		return PAMHD_OK;
	}
	if (!h->bg) HIP_TRY(h, hipMalloc(&h->bg, sizeof(double) * 9 * h->g.n));
	HIP_TRY(h, hipMemcpy(h->bg, bg, sizeof(double) * 9 * h->g.n, hipMemcpyHostToDevice));
	return PAMHD_OK;
}

int pamhd_timestep(pamhd_handle* h, int solver, double max_time_step, double time_step_factor,
	double* dt_taken, void* stream)
{
	touch(h);
	return do_step(h, solver, max_time_step, time_step_factor, 2, dt_taken, stream);
}

int pamhd_step(pamhd_handle* h, int solver, double max_time_step, double time_step_factor,
	double* dt_taken, void* stream)
{
	touch(h);
	return do_step(h, solver, max_time_step, time_step_factor, 3, dt_taken, stream);
}

int pamhd_apply_boundaries(pamhd_handle* h, double /*simulation_time*/, void* stream)
{
	touch(h);
	if (!h) return fail(nullptr, PAMHD_E_ARG, "null handle");
	if (h->box) {
		if (boxset_launch_boundaries(h->box, (hipStream_t)stream)) return fail(h, PAMHD_E_DEVICE, "boundary launch");
		return PAMHD_OK;
	}
	if (h->sub && h->sub->bounded) {
		/* value and copy boundaries, refit, sync; synchronous so that a
		   boundary error is reported here */
		const hipStream_t st = (hipStream_t)stream;
		ss_clear_kernel<<<1, SS_FOLD, 0, st>>>(*h->sub);
		launch_boundaries(h, st);
		HIP_TRY(h, hipGetLastError());
		HIP_TRY(h, hipMemcpyAsync(h->sub->host_status, h->sub->status, sizeof(StepStatus), hipMemcpyDeviceToHost, st));
		HIP_TRY(h, hipStreamSynchronize(st));
		return report_status(h, *h->sub->host_status);
	}
	{
		TimedLaunch t(h, (hipStream_t)stream, K_REFIT);
		refit_kernel<<<cell_grid(h->g), dim3(BX, BY, 1), 0, (hipStream_t)stream>>>(h->g, h->state[h->cur]);
	}
	HIP_TRY(h, hipGetLastError());
	return PAMHD_OK;
}

int pamhd_check(pamhd_handle* h)
{
	touch(h);
	if (!h) return fail(nullptr, PAMHD_E_ARG, "null handle");
	if (h->box) {
		std::string err;
		const int e = boxset_readback(h->box, nullptr, nullptr, err);
		return e ? fail(h, e, err) : PAMHD_OK;
	}
	HIP_TRY(h, hipDeviceSynchronize());
	if (h->sub && h->sub->bounded) {
		HIP_TRY(h, hipMemcpy(h->sub->host_status, h->sub->status, sizeof(StepStatus), hipMemcpyDeviceToHost));
		return report_status(h, *h->sub->host_status);
	}
	HIP_TRY(h, hipMemcpy(h->host_status, h->status + h->last, sizeof(StepStatus), hipMemcpyDeviceToHost));
	return report_status(h, *h->host_status);
}

int pamhd_max_signal_speeds(pamhd_handle* h, double smax_out[3])
{
	touch(h);
	if (h && h->sub && h->sub->bounded) return fail(h, PAMHD_E_UNSUPPORTED, "not available on a grid with boundaries");
	if (!h || !smax_out) return fail(h, PAMHD_E_ARG, "null argument");
	if (h->box) return fail(h, PAMHD_E_UNSUPPORTED, "per-box speeds: use the ensemble interface");
	HIP_TRY(h, hipDeviceSynchronize());
	StepStatus st;
	HIP_TRY(h, hipMemcpy(&st, h->status + h->last, sizeof(StepStatus), hipMemcpyDeviceToHost));
	for (int d = 0; d < 3; d++) {
		unsigned long long m = 0;
		for (int q = 0; q < NSLOT; q++) m = m < st.smax_bits[d][q] ? st.smax_bits[d][q] : m;
		double v;
		memcpy(&v, &m, sizeof(v));
		smax_out[d] = v;
	}
	return PAMHD_OK;
}

int pamhd_div_face2volume(pamhd_handle* h, double* avg)
{
	touch(h);
	if (h && h->sub && h->sub->bounded) return fail(h, PAMHD_E_UNSUPPORTED, "not available on a grid with boundaries");
	if (!h || !avg) return fail(h, PAMHD_E_ARG, "null argument");
	if (h->box) {
		HIP_TRY(h, boxset_div_face2volume(h->box, 0, avg));
		return PAMHD_OK;
	}
	const dim3 grid = cell_grid(h->g);
	const size_t nb = (size_t)grid.x * grid.y * grid.z;
	if (h->scratch_len < nb) {
		(void)hipFree(h->scratch);
		h->scratch = nullptr;
		HIP_TRY(h, hipMalloc(&h->scratch, sizeof(double) * nb));
		h->scratch_len = nb;
	}
	div_kernel<<<grid, dim3(BX, BY, 1)>>>(h->g, h->state[h->cur], h->scratch);
	HIP_TRY(h, hipGetLastError());
	std::string tmp;
	double* host = new double[nb];
	const hipError_t e = hipMemcpy(host, h->scratch, sizeof(double) * nb, hipMemcpyDeviceToHost);
	if (e != hipSuccess) {
		delete[] host;
		return hip_fail(h, e, "div");
	}
	double total = 0;
	for (size_t b = 0; b < nb; b++) total += host[b];
	delete[] host;
	*avg = total / (double)h->g.n;
	return PAMHD_OK;
}

int pamhd_clean_divergence(pamhd_handle* h, double tol, int max_cycles, double* div_before,
	double* rel_residual, int* cycles, void* stream_)
{
	touch(h);
	if (h && h->sub && h->sub->bounded) return fail(h, PAMHD_E_UNSUPPORTED, "not available on a grid with boundaries");
	if (!h) return fail(nullptr, PAMHD_E_ARG, "null handle");
	if (h->box) return fail(h, PAMHD_E_UNSUPPORTED, "divergence cleaning needs a periodic grid");
	if (!(tol >= 0) || max_cycles < 0) return fail(h, PAMHD_E_ARG, "tol and max_cycles must be >= 0");
	hipStream_t st = (hipStream_t)stream_;
	if (!h->mg) {
		h->mg = new PoissonMG();
		const int dims[3] = {h->g.nx, h->g.ny, h->g.nz};
		const double len[3] = {h->g.dx, h->g.dy, h->g.dz};
		const hipError_t e = mg_create(*h->mg, dims, len);
		if (e != hipSuccess) {
			mg_destroy(*h->mg);
			delete h->mg;
			h->mg = nullptr;
			return hip_fail(h, e, "multigrid levels");
		}
	}
	PoissonMG& mg = *h->mg;
	mg.fused_smoother = h->mg_fused;
	MgLevel& F = mg.lv[0];
	const int n = (int)h->g.n;
	const dim3 rgrid((h->g.nx + BX - 1) / BX, (h->g.ny + BY - 1) / BY, (h->g.nz + CLEAN_KZ - 1) / CLEAN_KZ);
	const int nb = (int)(rgrid.x * rgrid.y * rgrid.z);
	if ((size_t)2 * nb > mg.norm_part_len) return fail(h, PAMHD_E_DEVICE, "cleaning: partial-sum buffer too small");
	clean_rhs_kernel<<<rgrid, dim3(BX, BY, 1), 0, st>>>(h->g, h->state[h->cur], F.f, mg.norm_part,
		F.packed);
	HIP_TRY(h, hipGetLastError());
	double sums[2];
	HIP_TRY(h, mg_sum_pairs(mg, st, mg.norm_part, nb, sums));
	const double dsum = sums[0], fl = sums[1];
	const double atol = 8 * DBL_EPSILON * std::sqrt(fl);
	double rel = 0;
	int it = 0;
	bool ok = true;
	HIP_TRY(h, mg_solve(mg, st, tol, atol, max_cycles, &rel, &it, &ok));
	project_kernel<<<cell_grid(h->g), dim3(BX, BY, 1), 0, st>>>(h->g, h->state[h->cur],
		h->state[1 - h->cur], F.phi, F.packed);
	HIP_TRY(h, hipGetLastError());
	HIP_TRY(h, hipStreamSynchronize(st));
	h->cur = 1 - h->cur;
	if (div_before) *div_before = dsum / (double)n;
	if (rel_residual) *rel_residual = rel;
	if (cycles) *cycles = it;
	if (!ok) {
		return fail(h, PAMHD_E_POISSON, "divergence cleaning: relative residual " + std::to_string(rel) +
			" after " + std::to_string(it) + " V-cycles");
	}
	return PAMHD_OK;
}

int pamhd_step_fixed(pamhd_handle* h, int solver, double dt, void* stream)
{
	touch(h);
	if (!h) return fail(nullptr, PAMHD_E_ARG, "null handle");
	if (h->box) return fail(h, PAMHD_E_UNSUPPORTED, "fixed steps need a periodic grid");
	if (!(dt >= 0)) return fail(h, PAMHD_E_ARG, "dt must be >= 0");
	return do_step(h, solver, dt, 0.0, 3, nullptr, stream, 1);
}

int pamhd_step_fixed_from(pamhd_handle* h, int solver, double dt, const double* state_in,
	double* state_out, double* fluxes, void* stream)
{
	touch(h);
	if (!h) return fail(nullptr, PAMHD_E_ARG, "null handle");
	if (!state_in || !state_out) return fail(h, PAMHD_E_ARG, "null argument");
	/* fixed steps never substep; grids with boundaries keep their state in
	   the substepping layout */
	if (h->box || (h->sub && h->sub->bounded))
		return fail(h, PAMHD_E_UNSUPPORTED, "steps between caller states need a periodic grid");
	if (!(dt >= 0)) return fail(h, PAMHD_E_ARG, "dt must be >= 0");
	const size_t bytes = sizeof(double) * 8 * (size_t)h->g.n;
	const char* a = reinterpret_cast<const char*>(state_in);
	const char* b = reinterpret_cast<const char*>(state_out);
	if (a < b + bytes && b < a + bytes) return fail(h, PAMHD_E_ARG, "input and output states overlap");
	if (fluxes && (h->precision != PAMHD_FP64 || fused_step(h, solver))) {
		return fail(h, PAMHD_E_UNSUPPORTED, "the step keeps no FP64 flux array (HLLD's marches and "
			"mixed precision do not)");
	}
	/* the step reads the current and writes the other state buffer: lend it
	   the caller's two for this call, then restore the grid's own */
	double* own[2] = {h->state[0], h->state[1]};
	const int cur = h->cur;
	h->state[cur] = const_cast<double*>(state_in);
	h->state[1 - cur] = state_out;
	h->pe_out = fluxes;
	const int e = do_step(h, solver, dt, 0.0, 3, nullptr, stream, 1);
	h->pe_out = nullptr;
	h->state[0] = own[0];
	h->state[1] = own[1];
	h->cur = cur;
	return e;
}

int pamhd_set_state_device(pamhd_handle* h, const double* state, void* stream)
{
	touch(h);
	if (h && h->sub && h->sub->bounded) return fail(h, PAMHD_E_UNSUPPORTED, "not available on a grid with boundaries");
	if (!h || !state) return fail(h, PAMHD_E_ARG, "null argument");
	if (h->box) return fail(h, PAMHD_E_UNSUPPORTED, "solar-wind box state is not a flat array");
	HIP_TRY(h, hipMemcpyAsync(h->state[h->cur], state, sizeof(double) * 8 * h->g.n,
		hipMemcpyDeviceToDevice, (hipStream_t)stream));
	return PAMHD_OK;
}

int pamhd_get_state_device(pamhd_handle* h, double* state, void* stream)
{
	touch(h);
	if (h && h->sub && h->sub->bounded) return fail(h, PAMHD_E_UNSUPPORTED, "not available on a grid with boundaries");
	if (!h || !state) return fail(h, PAMHD_E_ARG, "null argument");
	if (h->box) return fail(h, PAMHD_E_UNSUPPORTED, "solar-wind box state is not a flat array");
	HIP_TRY(h, hipMemcpyAsync(state, h->state[h->cur], sizeof(double) * 8 * h->g.n,
		hipMemcpyDeviceToDevice, (hipStream_t)stream));
	return PAMHD_OK;
}

int pamhd_adjoint_step(pamhd_handle* h, int solver, double dt, const double* state, double* lambda,
	void* stream)
{
	touch(h);
	if (h && h->sub && h->sub->bounded) return fail(h, PAMHD_E_UNSUPPORTED, "not available on a grid with boundaries");
	if (!h || !state || !lambda) return fail(h, PAMHD_E_ARG, "null argument");
	if (h->box) return fail(h, PAMHD_E_UNSUPPORTED, "the adjoint needs a periodic grid");
	if (solver != HLL_ATHENA) return fail(h, PAMHD_E_UNSUPPORTED, "the adjoint is implemented for hll-athena");
	if (h->precision != PAMHD_FP64) return fail(h, PAMHD_E_UNSUPPORTED, "the adjoint is FP64");
	if (h->bg) return fail(h, PAMHD_E_UNSUPPORTED, "the adjoint has no background field");
	if (!h->adj) HIP_TRY(h, hipMalloc(&h->adj, sizeof(double) * adjoint_scratch_doubles(h->g.nx, h->g.ny, h->g.nz)));
	const double len[3] = {h->g.dx, h->g.dy, h->g.dz};
	adjoint_launch(h->g.nx, h->g.ny, h->g.nz, len, dt, h->g.ph, state, lambda, h->adj,
		(hipStream_t)stream, h->adj_reverse, nullptr);
	HIP_TRY(h, hipGetLastError());
	return PAMHD_OK;
}

int pamhd_get_magnetic_fluxes(pamhd_handle* h, double* fluxes, void* stream)
{
	touch(h);
	if (!h || !fluxes) return fail(h, PAMHD_E_ARG, "null argument");
	if (!h->flux_step) {
		return fail(h, PAMHD_E_UNSUPPORTED, "the last step kept no FP64 flux array (HLLD's marches, "
			"mixed precision, substepping and boundaries do not)");
	}
	const long long n = h->g.n;
	magnetic_fluxes_kernel<<<(unsigned)((n + 255) / 256), 256, 0, (hipStream_t)stream>>>(h->flux, fluxes, n);
	HIP_TRY(h, hipGetLastError());
	return PAMHD_OK;
}

int pamhd_adjoint_steps(pamhd_handle* h, int solver, int nsteps, const double* dts,
	const double* const* states, const double* const* fluxes, double* lambda, void* stream)
{
	touch(h);
	if (h && h->sub && h->sub->bounded) return fail(h, PAMHD_E_UNSUPPORTED, "not available on a grid with boundaries");
	if (!h || !dts || !states || !lambda) return fail(h, PAMHD_E_ARG, "null argument");
	if (nsteps < 0) return fail(h, PAMHD_E_ARG, "nsteps must be >= 0");
	for (int q = 0; q < nsteps; q++) {
		if (!states[q]) return fail(h, PAMHD_E_ARG, "null state");
	}
	if (h->box) return fail(h, PAMHD_E_UNSUPPORTED, "the adjoint needs a periodic grid");
	if (solver != HLL_ATHENA) return fail(h, PAMHD_E_UNSUPPORTED, "the adjoint is implemented for hll-athena");
	if (h->precision != PAMHD_FP64) return fail(h, PAMHD_E_UNSUPPORTED, "the adjoint is FP64");
	if (h->bg) return fail(h, PAMHD_E_UNSUPPORTED, "the adjoint has no background field");
	if (nsteps == 0) return PAMHD_OK;
	if (!h->adj) HIP_TRY(h, hipMalloc(&h->adj, sizeof(double) * adjoint_scratch_doubles(h->g.nx, h->g.ny, h->g.nz)));
	const double len[3] = {h->g.dx, h->g.dy, h->g.dz};
	adjoint_steps_launch(h->g.nx, h->g.ny, h->g.nz, len, nsteps, dts, states, fluxes, h->g.ph, lambda,
		h->adj, (hipStream_t)stream, h->adj_reverse);
	HIP_TRY(h, hipGetLastError());
	return PAMHD_OK;
}

int pamhd_adjoint_step_fluxes(pamhd_handle* h, int solver, double dt, const double* state,
	const double* fluxes, double* lambda, void* stream)
{
	touch(h);
	if (!fluxes) return fail(h, PAMHD_E_ARG, "null argument");
	if (h && h->sub && h->sub->bounded) return fail(h, PAMHD_E_UNSUPPORTED, "not available on a grid with boundaries");
	if (!h || !state || !lambda) return fail(h, PAMHD_E_ARG, "null argument");
	if (h->box) return fail(h, PAMHD_E_UNSUPPORTED, "the adjoint needs a periodic grid");
	if (solver != HLL_ATHENA) return fail(h, PAMHD_E_UNSUPPORTED, "the adjoint is implemented for hll-athena");
	if (h->precision != PAMHD_FP64) return fail(h, PAMHD_E_UNSUPPORTED, "the adjoint is FP64");
	if (h->bg) return fail(h, PAMHD_E_UNSUPPORTED, "the adjoint has no background field");
	if (!h->adj) HIP_TRY(h, hipMalloc(&h->adj, sizeof(double) * adjoint_scratch_doubles(h->g.nx, h->g.ny, h->g.nz)));
	const double len[3] = {h->g.dx, h->g.dy, h->g.dz};
	adjoint_launch(h->g.nx, h->g.ny, h->g.nz, len, dt, h->g.ph, state, lambda, h->adj,
		(hipStream_t)stream, h->adj_reverse, fluxes);
	HIP_TRY(h, hipGetLastError());
	return PAMHD_OK;
}

int pamhd_set_adjoint_jacobian(pamhd_handle* h, int reverse)
{
	touch(h);
	if (!h) return fail(nullptr, PAMHD_E_ARG, "null handle");
	h->adj_reverse = reverse != 0;
	return PAMHD_OK;
}

int pamhd_set_poisson_smoother(pamhd_handle* h, int fused)
{
	touch(h);
	if (!h) return fail(nullptr, PAMHD_E_ARG, "null handle");
	h->mg_fused = fused != 0;
	return PAMHD_OK;
}

int pamhd_set_substepping(pamhd_handle* h, int substep_min, int substep_max)
{
	touch(h);
	if (!h) return fail(nullptr, PAMHD_E_ARG, "null handle");
	if (h->box) return fail(h, PAMHD_E_UNSUPPORTED, "substepping needs a periodic grid");
	if (substep_max < 0) return fail(h, PAMHD_E_ARG, "substep_max must be >= 0");
	if (substep_min < 0) return fail(h, PAMHD_E_ARG, "substep_min must be >= 0");
	if (substep_max > 0 && h->precision != PAMHD_FP64) {
		return fail(h, PAMHD_E_UNSUPPORTED, "substepping runs in FP64 only");
	}
	if (substep_max > 0 && !h->sub) {
		const int e = substep_alloc(h);
		if (e) return e;
	}
	if (h->sub) {
		/* the same minimum in every cell is shifted to 0 */
		h->sub->opt = nullptr;
		h->sub->smax = substep_max;
	}
	return PAMHD_OK;
}

int pamhd_set_substep_walk(pamhd_handle* h, int workgroups)
{
	touch(h);
	if (!h) return fail(nullptr, PAMHD_E_ARG, "null handle");
	if (workgroups < 0) return fail(h, PAMHD_E_ARG, "workgroups must be >= 0");
	h->ss_walk = workgroups;
	if (h->sub) h->sub->walk = workgroups;
	return PAMHD_OK;
}

int pamhd_set_substepping_cells(pamhd_handle* h, const int* substep_min, const int* substep_max)
{
	touch(h);
	if (!h) return fail(nullptr, PAMHD_E_ARG, "null handle");
	if (!substep_min || !substep_max) return fail(h, PAMHD_E_ARG, "null argument");
	if (h->box) return fail(h, PAMHD_E_UNSUPPORTED, "substepping needs a periodic grid");
	const long long n = h->g.n;
	int lo = INT_MAX, hi = 0;
	for (long long c = 0; c < n; c++) {
		if (substep_min[c] < 0 || substep_max[c] < 0) {
			return fail(h, PAMHD_E_ARG, "substep options must be >= 0 (cell " + std::to_string(c) + ")");
		}
		lo = substep_min[c] < lo ? substep_min[c] : lo;
		hi = substep_max[c] > hi ? substep_max[c] : hi;
	}
	if (h->precision != PAMHD_FP64) {
		if (hi > 0) return fail(h, PAMHD_E_UNSUPPORTED, "substepping runs in FP64 only");
		return PAMHD_OK;   /* every period is 1: the global step */
	}
	if (!h->sub) {
		const int e = substep_alloc(h);
		if (e) return e;
	}
	std::vector<int> opt((size_t)(2 * n));
	for (long long c = 0; c < n; c++) {
		opt[(size_t)c] = substep_min[c] - lo;
		opt[(size_t)(n + c)] = substep_max[c];
	}
	HIP_TRY(h, hipMemcpy(h->sub->opt_dev, opt.data(), opt.size() * sizeof(int), hipMemcpyHostToDevice));
	h->sub->opt = h->sub->opt_dev;
	h->sub->smax = hi;
	return PAMHD_OK;
}

int pamhd_last_flux_calcs(pamhd_handle* h, unsigned long long* flux_calcs)
{
	touch(h);
	if (!h || !flux_calcs) return fail(h, PAMHD_E_ARG, "null argument");
	*flux_calcs = h->flux_calcs;
	return PAMHD_OK;
}

int pamhd_last_max_substep(pamhd_handle* h, int* max_substep)
{
	touch(h);
	if (!h || !max_substep) return fail(h, PAMHD_E_ARG, "null argument");
	*max_substep = h->sub ? h->sub->last_max : 1;
	return PAMHD_OK;
}

int pamhd_last_substep_periods(pamhd_handle* h, int* periods)
{
	touch(h);
	if (!h || !periods) return fail(h, PAMHD_E_ARG, "null argument");
	if (!h->sub || !h->sub->period) {
		for (long long c = 0; c < h->g.n; c++) periods[c] = 1;
		return PAMHD_OK;
	}
	HIP_TRY(h, hipDeviceSynchronize());
	HIP_TRY(h, hipMemcpy(periods, h->sub->period, sizeof(int) * (size_t)h->g.n, hipMemcpyDeviceToHost));
	return PAMHD_OK;
}

int pamhd_plain1d_run(size_t ncells, double cell_size, double adiabatic_index,
	double vacuum_permeability, double cfl, int solver, double t_end, long max_steps,
	double* state, long* steps_taken, void* stream)
{
	if (!state) return fail(nullptr, PAMHD_E_ARG, "null state");
	if (!(cell_size > 0) || !(adiabatic_index > 1) || !(vacuum_permeability > 0) || !(cfl > 0) ||
		max_steps < 0)
		return fail(nullptr, PAMHD_E_ARG, "invalid cell size, physics constants, cfl or step limit");
	std::string err;
	const int e = pamhd_dev::plain1d_run((long long)ncells, cell_size, adiabatic_index,
		vacuum_permeability, cfl, solver, t_end, max_steps, state, steps_taken,
		(hipStream_t)stream, err);
	if (e) {
		set_global_error(e == PAMHD_E_ARG || e == PAMHD_E_SOLVER || e == PAMHD_E_DEVICE ? err :
			std::string(error_name(e)) + " (" + err + ")");
		return e;
	}
	return PAMHD_OK;
}

int pamhd_plain1d_run_noalloc(size_t ncells, double cell_size, double adiabatic_index,
	double vacuum_permeability, double cfl, int solver, double t_end, long max_steps,
	double* state_device, long* steps_taken, void* stream)
{
	if (!state_device) return fail(nullptr, PAMHD_E_ARG, "null state");
	if (!(cell_size > 0) || !(adiabatic_index > 1) || !(vacuum_permeability > 0) || !(cfl > 0) ||
		max_steps < 0)
		return fail(nullptr, PAMHD_E_ARG, "invalid cell size, physics constants, cfl or step limit");
	std::string err;
	const int e = pamhd_dev::plain1d_run_device((long long)ncells, cell_size, adiabatic_index,
		vacuum_permeability, cfl, solver, t_end, max_steps, state_device, steps_taken,
		(hipStream_t)stream, err);
	if (e) {
		set_global_error(e == PAMHD_E_ARG || e == PAMHD_E_SOLVER || e == PAMHD_E_DEVICE ? err :
			std::string(error_name(e)) + " (" + err + ")");
		return e;
	}
	return PAMHD_OK;
}

int pamhd_face_flux_batch(int solver, int precision, size_t n, const double* L, const double* R,
	const double* bg, double gamma, double mu0, double* F, double* smax_out, int* status,
	void* stream)
{
	if (n == 0) return PAMHD_OK;
	if (!L || !R || !F || !smax_out) return fail(nullptr, PAMHD_E_ARG, "null argument");
	if (precision != PAMHD_FP64 && precision != PAMHD_MIXED) return fail(nullptr, PAMHD_E_ARG, "invalid precision");
	hipStream_t st = (hipStream_t)stream;
	const dim3 grid((unsigned)((n + 255) / 256)), block(256);
	const Phys ph = make_phys(gamma, mu0);
#define BATCH(S, T) face_batch_kernel<S, T><<<grid, block, 0, st>>>(n, L, R, bg, ph, F, smax_out, status)
	if (precision == PAMHD_FP64) {
		switch (solver) {
		case RUSANOV: BATCH(RUSANOV, double); break;
		case HLL_ATHENA: BATCH(HLL_ATHENA, double); break;
		case HLLD_ATHENA: BATCH(HLLD_ATHENA, double); break;
		case ROE_ATHENA: BATCH(ROE_ATHENA, double); break;
		default: return fail(nullptr, PAMHD_E_SOLVER, "invalid solver");
		}
	} else {
		switch (solver) {
		case RUSANOV: BATCH(RUSANOV, float); break;
		case HLL_ATHENA: BATCH(HLL_ATHENA, float); break;
		case HLLD_ATHENA: BATCH(HLLD_ATHENA, float); break;
		case ROE_ATHENA: BATCH(ROE_ATHENA, float); break;
		default: return fail(nullptr, PAMHD_E_SOLVER, "invalid solver");
		}
	}
#undef BATCH
	const hipError_t e = hipGetLastError();
	if (e != hipSuccess) return hip_fail(nullptr, e, "face_batch_kernel");
	return PAMHD_OK;
}

int pamhd_autotune(pamhd_handle* h, int solver, char* buf, size_t buflen, void* stream_)
{
	touch(h);
	if (h && h->sub && h->sub->bounded) return fail(h, PAMHD_E_UNSUPPORTED, "not available on a grid with boundaries");
	if (!h) return fail(nullptr, PAMHD_E_ARG, "null handle");
	if (h->box) return fail(h, PAMHD_E_UNSUPPORTED, "solar-wind box: use the ensemble autotuner");
	if (solver != RUSANOV && solver != HLL_ATHENA && solver != HLLD_ATHENA && solver != ROE_ATHENA) {
		return fail(h, PAMHD_E_SOLVER, "invalid solver " + std::to_string(solver));
	}
	hipStream_t st = (hipStream_t)stream_;
	const bool timing = h->timing;
	h->timing = false;
	StepStatus* prev = h->status + h->last;
	StepStatus* next = h->status + (1 - h->last);
	hipEvent_t a, b;
	HIP_TRY(h, hipEventCreate(&a));
	HIP_TRY(h, hipEventCreate(&b));
	/* average of 3 timed launches after one warm-up; neither the current
	   state nor the status of the last flux pass is modified */
	auto time_it = [&](auto&& launch) -> float {
		launch();
		float best = 0;
		(void)hipEventRecord(a, st);
		for (int r = 0; r < 3; r++) launch();
		(void)hipEventRecord(b, st);
		(void)hipEventSynchronize(b);
		(void)hipEventElapsedTime(&best, a, b);
		return best / 3;
	};
	auto flux = [&]() {
		dt_kernel<<<1, NSLOT, 0, st>>>(prev, next, h->dt, 0.0, 0.0, h->g.dx, h->g.dy, h->g.dz, 0);
		if (h->precision == PAMHD_MIXED) launch_flux_pass<float, double>(h, solver, st, next);
		else launch_flux_pass<double, double>(h, solver, st, next);
	};
	std::string report = "flux_pass:";
	float best_flux = 1e30f;
	int best_remap = 0, best_occ = 0, best_shared = 0, best_km = 16;
	/* (shared mode, cells per march) */
	const int modes[][2] = {{0, 16}, {1, 16}, {2, 8}, {2, 16}, {2, 32}, {3, 8}, {3, 16}, {3, 32}};
	for (const auto& md : modes)
	for (int occ = 0; occ < 2; occ++)
	for (int remap = 0; remap < (md[0] ? 2 : 4); remap++) {
		/* the per-cell part: HLLD, and HLL / Rusanov without a background
		   field (marches only, one occupancy build) */
		if (md[0] && solver != HLLD_ATHENA &&
			(md[0] == 1 || occ || h->bg || (solver != HLL_ATHENA && solver != RUSANOV)))
			continue;
		h->tune.flux_remap = remap;
		h->tune.flux_occ = occ;
		h->tune.flux_shared = md[0];
		h->tune.flux_km = md[1];
		const float ms = time_it(flux);
		report += std::string(md[0] == 1 ? " shared/" : md[0] == 2 ? " march" + std::to_string(md[1]) + "/"
			: md[0] == 3 ? " xmarch" + std::to_string(md[1]) + "/" : " ") +
			"remap" + std::to_string(remap) + (occ ? "/occ4" : "") + "=" + std::to_string(ms) + "ms";
		if (ms < best_flux) {
			best_flux = ms;
			best_remap = remap;
			best_occ = occ;
			best_shared = md[0];
			best_km = md[1];
		}
	}
	h->tune.flux_remap = best_remap;
	h->tune.flux_occ = best_occ;
	h->tune.flux_shared = best_shared;
	h->tune.flux_km = best_km;
	flux();
	report += "; update:";
	const int kzs[] = {0, 8, 16, 32, 64};
	float best_upd = 1e30f;
	pamhd_handle::Tune best = h->tune;
	for (int kz : kzs)
	for (int tile = 0; tile < (kz ? 5 : 1); tile++) {
		if (kz > 0 && kz > h->g.nz) continue;
		for (int remap = 0; remap < (kz ? 2 : 4); remap++) {
			h->tune.update_kz = kz;
			h->tune.update_remap = remap;
			h->tune.update_tile = tile;
			const float ms = time_it([&]() {
				launch_update<double>(h, st, 3);
			});
			static const char* tiles[] = {"64x4", "128x2", "32x8", "128x4", "64x8"};
			report += " " + std::string(kz ? "column" + std::to_string(kz) + "/" + tiles[tile] : "cell") +
				"/remap" + std::to_string(remap) + "=" + std::to_string(ms) + "ms";
			if (ms < best_upd) {
				best_upd = ms;
				best = h->tune;
			}
		}
	}
	h->tune = best;
	/* z-slab pipelining of flux pass and per-cell update (whole step) */
	if (h->tune.update_kz == 0) {
		report += "; step:";
		float best_step = 1e30f;
		int best_slabs = 1;
		const int slabs[] = {1, 4, 8, 16};
		for (int S : slabs) {
			if (S > 1 && S > h->g.nz / 2) continue;
			h->tune.slabs = S;
			const float ms = time_it([&]() {
				dt_kernel<<<1, NSLOT, 0, st>>>(prev, next, h->dt, 0.0, 0.0, h->g.dx, h->g.dy, h->g.dz, 0);
				if (h->precision == PAMHD_MIXED) launch_step_body<float, double>(h, solver, st, next, 3);
				else launch_step_body<double, double>(h, solver, st, next, 3);
			});
			report += " slabs" + std::to_string(S) + "=" + std::to_string(ms) + "ms";
			if (ms < best_step) {
				best_step = ms;
				best_slabs = S;
			}
		}
		h->tune.slabs = best_slabs;
	}
	/* HLLD: the step structure (flux passes and update, fused single pass,
	   accumulating marches) and its chunk length, then the tile of the
	   face-field update that follows the latter two */
	if (solver == HLLD_ATHENA) {
		auto step = [&]() {
			dt_kernel<<<1, NSLOT, 0, st>>>(prev, next, h->dt, 0.0, 0.0, h->g.dx, h->g.dy, h->g.dz, 0);
			if (h->precision == PAMHD_MIXED) launch_step_body<float, double>(h, solver, st, next, 3);
			else launch_step_body<double, double>(h, solver, st, next, 3);
		};
		report += "; hlld step:";
		pamhd_handle::Tune best_t = h->tune;
		best_t.fused = 0;
		h->tune.fused = 0;
		float best_ms = time_it(step);
		report += " passes+update=" + std::to_string(best_ms) + "ms";
		const int chunks[] = {8, 16, 32};
		for (int mode = 1; mode <= 2; mode++)
		for (int c : chunks) {
			h->tune = best_t;
			h->tune.fused = mode;
			if (mode == 1) h->tune.fused_kc = c;
			else h->tune.acc_km = c;
			h->tune.update_kz = 8;
			h->tune.update_tile = 1;
			const float ms = time_it(step);
			report += std::string(mode == 1 ? " fused/kc" : " accumulate/km") + std::to_string(c) + "=" +
				std::to_string(ms) + "ms";
			if (ms < best_ms) {
				best_ms = ms;
				best_t = h->tune;
			}
		}
		h->tune = best_t;
		if (h->tune.fused) {
			const int ct_kz[] = {8, 16, 32};
			for (int kz : ct_kz)
			for (int tile = 0; tile < 5; tile++) {
				if (kz > h->g.nz) continue;
				h->tune = best_t;
				h->tune.update_kz = kz;
				h->tune.update_tile = tile;
				const float ms = time_it(step);
				report += " ct" + std::to_string(kz) + "/tile" + std::to_string(tile) + "=" +
					std::to_string(ms) + "ms";
				if (ms < best_ms) {
					best_ms = ms;
					best_t = h->tune;
				}
			}
			h->tune = best_t;
		}
	}
	(void)hipEventDestroy(a);
	(void)hipEventDestroy(b);
	HIP_TRY(h, hipGetLastError());
	HIP_TRY(h, hipStreamSynchronize(st));
	h->timing = timing;
	report += "; chosen fused=" + std::to_string(h->tune.fused) +
		(h->tune.fused == 1 ? "/kc" + std::to_string(h->tune.fused_kc) :
		 h->tune.fused == 2 ? "/km" + std::to_string(h->tune.acc_km) : std::string()) +
		" flux_remap=" + std::to_string(h->tune.flux_remap) +
		" flux_occ4=" + std::to_string(h->tune.flux_occ) +
		" flux_shared=" + std::to_string(h->tune.flux_shared) +
		(h->tune.flux_shared >= 2 ? "/km" + std::to_string(h->tune.flux_km) : std::string()) +
		" update=" + (h->tune.update_kz ? "column" + std::to_string(h->tune.update_kz) + "/tile" +
			std::to_string(h->tune.update_tile) : std::string("cell")) +
		" update_remap=" + std::to_string(h->tune.update_remap) +
		" slabs=" + std::to_string(h->tune.slabs);
	if (buf && buflen) snprintf(buf, buflen, "%s", report.c_str());
	return PAMHD_OK;
}

int pamhd_set_launch(pamhd_handle* h, int flux_remap, int update_kz, int update_remap)
{
	touch(h);
	if (!h) return fail(nullptr, PAMHD_E_ARG, "null handle");
	if (update_kz < 0) return fail(h, PAMHD_E_ARG, "update_kz must be >= 0");
	h->tune.flux_remap = flux_remap & 3;
	h->tune.flux_occ = (flux_remap >> 2) & 1;
	h->tune.flux_shared = (flux_remap & 64) ? 3 : (flux_remap & 16) ? 2 : (flux_remap >> 3) & 1;
	h->tune.flux_km = (flux_remap & 32) ? 4 : 16;
	h->tune.update_kz = update_kz;
	h->tune.update_remap = update_kz ? (update_remap & 1) : (update_remap & 3);
	h->tune.update_tile = update_kz ? (update_remap >> 2) % 5 : 0;
	return PAMHD_OK;
}

int pamhd_set_fused_step(pamhd_handle* h, int fused, int planes_per_chunk)
{
	touch(h);
	if (!h) return fail(nullptr, PAMHD_E_ARG, "null handle");
	if (fused < -1 || fused > 2) return fail(h, PAMHD_E_ARG, "fused must be -1, 0, 1 or 2");
	const int most = fused == 2 ? ACC_KMAX : FKC_MAX;
	if (planes_per_chunk < 0 || planes_per_chunk > most)
		return fail(h, PAMHD_E_ARG, "planes per chunk must be in 0.." + std::to_string(most));
	h->tune.fused = fused;
	if (planes_per_chunk > 0) {
		if (fused == 2) h->tune.acc_km = planes_per_chunk;
		else h->tune.fused_kc = planes_per_chunk;
	}
	return PAMHD_OK;
}

int pamhd_set_slabs(pamhd_handle* h, int slabs)
{
	touch(h);
	if (!h) return fail(nullptr, PAMHD_E_ARG, "null handle");
	if (slabs < 1) return fail(h, PAMHD_E_ARG, "slabs must be >= 1");
	h->tune.slabs = slabs;
	return PAMHD_OK;
}

int pamhd_enable_timing(pamhd_handle* h, int enable)
{
	touch(h);
	if (!h) return fail(nullptr, PAMHD_E_ARG, "null handle");
	collect_times(h);
	h->timing = enable != 0;
	for (int k = 0; k < 4; k++) {
		h->kernel_ms[k] = 0;
		h->kernel_launches[k] = 0;
	}
	return PAMHD_OK;
}

int pamhd_kernel_time(pamhd_handle* h, const char* kernel, double* total_ms, long* launches)
{
	touch(h);
	if (!h || !kernel || !total_ms || !launches) return fail(h, PAMHD_E_ARG, "null argument");
	collect_times(h);
	for (int k = 0; k < 4; k++) {
		if (strcmp(kernel, kernel_names[k]) == 0) {
			*total_ms = h->kernel_ms[k];
			*launches = h->kernel_launches[k];
			return PAMHD_OK;
		}
	}
	return fail(h, PAMHD_E_ARG, std::string("unknown kernel ") + kernel);
}

int pamhd_last_error(pamhd_handle* h, char* buf, size_t n)
{
	touch(h);
	if (!buf || n == 0) return PAMHD_E_ARG;
	std::string m;
	if (h) {
		m = h->err;
	} else {
		std::lock_guard<std::mutex> lock(g_err_mutex);
		m = g_err;
	}
	snprintf(buf, n, "%s", m.c_str());
	return PAMHD_OK;
}

} // extern "C"
