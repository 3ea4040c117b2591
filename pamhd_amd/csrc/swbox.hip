/*
Solar-wind box path on MI355X (gfx950): the program of
tests/mhd/solar_wind_box.cpp on uniform grids, for one box or an ensemble of
boxes of different sizes advanced together.

Data layout (all boxes concatenated, cell g = box offset + local index
i + nx (j + ny k), component-major over the N cells of the set):
  S      20 * N doubles: rho, m (3), E, Vol_B (3), Face_B copies of the
         cell's -x, +x, -y, +y, -z, +z faces, and a second set of the six
         face copies (scratch within a step).  Each cell keeps its own copy
         of each face, like the reference's Face_Magnetic_Field, because the
         boundary rules and sync_magnetic_field act on copies.
  P      24 * N doubles: dt * F of the d-face on each cell's negative side
         (0 where no flux is computed)
  ctype  N: 1 normal, 0 boundary, -1 not solved (Cell_Type)
  rule   N: boundary rule (1..27 copy from a neighbour offset, 28 planet,
         29 inflow)
  bg     9 * N: Bg_B(cell)(+d) at the cell's positive face centres

A main-loop step's launches, each covering every box:
  dt            lagged CFL step per box (one thread per box)
  flux          the three face directions of each cell in one launch
                (or one launch per direction); per-box max-speed reductions
  update        normal cells: constrained-transport change of their six
                face copies, conserved update; cells of type >= 0: refit
  sync + refit  each cell's six face copies synced with the other copies
                into the second set, cells of type >= 0 refit from them
  boundary      copy / planet / inflow rules on the second set (reads only
                normal cells; a thread per cell of the set's list of cells
                with a rule)
  inflow face   interior neighbours of the inflow cells (a thread per cell
                of the list of inflow cells)
  sync          from the second set back into the first
A step without boundaries runs update, sync (one thread per face writes
both copies in place) and refit instead.  No kernel reads what another
thread of the same launch writes.  A persistent run (box_run_kernel) does
the main-loop stages for all its steps in one cooperative launch, with a
grid-wide barrier between stages in place of the launch boundaries.
*/
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstring>
#include <vector>

#include "riemann.hpp"
#include "swbox.hpp"

namespace pamhd_dev {

namespace {

enum { R_NONE = 0, R_PLANET = 28, R_SW = 29 };
/* NQ values per cell; Q_F2: the second set of face copies that a step with
   boundaries syncs into and back from (scratch between steps) */
enum { Q_RHO = 0, Q_M = 1, Q_E = 4, Q_VB = 5, Q_F = 8, NQ = 14, Q_F2 = NQ, NQ_ALLOC = NQ + 6 };

struct BoxDev {
	int nx, ny, nz, n;
	long long off;
	double dx, dy, dz;
	int gx, gy, blk0;
	int sw_dim, sw_sg;
	double sw_rho, sw_m[3], sw_E, sw_B[3];
	double in_rho, in_m[3], in_E;
};

struct BoxStatus {
	unsigned long long smax_bits[3];
	unsigned int zero[3];
	int err;
	long long err_face;
};

struct Tile { int tx, ty; };

/* a cell of a set's list: global index and box */
struct CellRef {
	long long g;
	int b, pad;
};
constexpr Tile TILES[3] = {{64, 4}, {32, 8}, {16, 16}};

} // namespace

struct BoxSet {
	int nbox = 0;
	long long N = 0;
	Phys ph;
	double mp = 0;
	std::vector<BoxDev> host;
	std::vector<pamhd_solar_wind_desc> sw;
	BoxDev* boxes = nullptr;
	int* blockbox = nullptr;
	int nblocks = 0;
	int tile = 0;
	int flux3 = 1;   /* the three flux directions in one launch (box_flux3_kernel) */
	double* S = nullptr;
	double* P = nullptr;
	double* bg = nullptr;
	signed char* ctype = nullptr;
	unsigned char* rule = nullptr;
	BoxStatus* status = nullptr;   /* 2 * nbox */
	double* dt = nullptr;          /* nbox */
	double* partial = nullptr;
	unsigned* bar = nullptr;       /* persistent run: arrival count and error flag */
	CellRef* bnd = nullptr;        /* cells with a boundary rule, in index order */
	long long nbnd = 0;
	CellRef* inflow = nullptr;     /* inflow cells, in index order */
	long long ninflow = 0;
	unsigned bar_polls = 0;        /* polls before a barrier gives up (0: RUN_SPIN_LIMIT) */
	BoxStatus* host_status = nullptr;
	double* host_dt = nullptr;
	int last = 0;
	std::vector<long long> faces;  /* per box: faces between cells of type >= 0 */
};

namespace {

/* cell (i, j, k) of box b handled by this thread in tile tl of the set (a
   workgroup's tile: its block index, or a tile of a persistent run's
   walk); false past the box edge */
template <int TX, int TY>
__device__ __forceinline__ bool my_cell(const BoxDev* __restrict__ boxes, const int* __restrict__ blockbox,
	int tl, int& b, int& i, int& j, int& k)
{
	b = blockbox[tl];
	const BoxDev& B = boxes[b];
	const int lb = tl - B.blk0;
	const int tx = lb % B.gx, ty = (lb / B.gx) % B.gy;
	k = lb / (B.gx * B.gy);
	i = tx * TX + (int)threadIdx.x;
	j = ty * TY + (int)threadIdx.y;
	return i < B.nx && j < B.ny;
}

__device__ __forceinline__ long long gidx(const BoxDev& B, int i, int j, int k)
{
	return B.off + i + (long long)B.nx * (j + (long long)B.ny * k);
}

__device__ __forceinline__ double pressure_of(const double* __restrict__ S, long long N, long long g,
	const Phys& ph)
{
	const double U[8] = {S[g], S[N + g], S[2 * N + g], S[3 * N + g], S[4 * N + g],
		S[5 * N + g], S[6 * N + g], S[7 * N + g]};
	return pressure64(U, ph.gm1, ph.half_inv_mu0);
}

/* update_vol_B with constant thermal pressure, source/mhd/solve.hpp:1646-1672,
   from the cell's face copies f (-x, +x, -y, +y, -z, +z) */
__device__ __forceinline__ void refit_cell_faces(double* __restrict__ S, long long N, long long g,
	const Phys& ph, const double f[6])
{
	const double rho = S[g];
	if (rho <= 0) return;
	const double p = pressure_of(S, N, g, ph);
	double vb[3];
	for (int d = 0; d < 3; d++) {
		vb[d] = 0.5 * (f[2 * d] + f[2 * d + 1]);
		S[(Q_VB + d) * N + g] = vb[d];
	}
	const double inv = 1 / rho;
	const double v[3] = {S[N + g] * inv, S[2 * N + g] * inv, S[3 * N + g] * inv};
	S[4 * N + g] = total_energy64(rho, v, p, vb, ph.gm1, ph.half_inv_mu0);
}

__device__ __forceinline__ void refit_cell(double* __restrict__ S, long long N, long long g, const Phys& ph)
{
	double f[6];
	for (int q = 0; q < 6; q++) f[q] = S[(Q_F + q) * N + g];
	refit_cell_faces(S, N, g, ph, f);
}

/* the two copies of one face after sync_magnetic_field's averaging
   (source/common_functions.hpp:184-269): a is the copy of the cell on the
   face's negative side, o the other */
__device__ __forceinline__ void sync_pair(double& a, double& o)
{
	const double diff = a - o;
	a -= diff / 2;
	o += diff / 2;
	const double diff2 = o - a;
	o -= diff2 / 2;
	a += diff2 / 2;
}

/* the six face copies of cell g (type ct) after that sync, from the copies
   at component fin: each face shared with a cell of the box of type >= 0
   is synced with the other copy (a cell of type < 0 keeps its copies) */
__device__ __forceinline__ void synced_faces(const BoxDev& B, long long N, const double* __restrict__ S,
	const signed char* __restrict__ ctype, long long g, int ct, int i, int j, int k, int fin, double f[6])
{
	const int id[3] = {i, j, k};
	const int dims[3] = {B.nx, B.ny, B.nz};
	const long long st[3] = {1, B.nx, (long long)B.nx * B.ny};
	for (int d = 0; d < 3; d++) {
		double lo = S[(fin + 2 * d) * N + g], hi = S[(fin + 2 * d + 1) * N + g];
		if (ct >= 0) {
			if (id[d] >= 1 && ctype[g - st[d]] >= 0) {
				double a = S[(fin + 2 * d + 1) * N + g - st[d]];
				sync_pair(a, lo);
			}
			if (id[d] + 1 < dims[d] && ctype[g + st[d]] >= 0) {
				double o = S[(fin + 2 * d) * N + g + st[d]];
				sync_pair(hi, o);
			}
		}
		f[2 * d] = lo;
		f[2 * d + 1] = hi;
	}
}

__device__ __forceinline__ void box_record_error(BoxStatus* st, int code, long long face)
{
	if (atomicCAS(&st->err, 0, code) == 0) st->err_face = face;
}

__device__ __forceinline__ void box_dt_body(const BoxDev* __restrict__ boxes, const BoxStatus* __restrict__ prev,
	BoxStatus* __restrict__ next, double* __restrict__ dt_out, double cfl, double max_dt, int b)
{
	const BoxDev& B = boxes[b];
	const double len[3] = {B.dx, B.dy, B.dz};
	double cfl_dt = DBL_MAX;
	for (int d = 0; d < 3; d++) {
		if (prev[b].zero[d] > 0) {
			cfl_dt = 0;
			break;
		}
		cfl_dt = smin(cfl_dt, cfl * len[d] / __longlong_as_double((long long)prev[b].smax_bits[d]));
	}
	dt_out[b] = (max_dt <= cfl_dt) ? max_dt : cfl_dt;
	for (int d = 0; d < 3; d++) {
		next[b].smax_bits[d] = 0;
		next[b].zero[d] = 0;
	}
	next[b].err = 0;
	next[b].err_face = -1;
}

__global__ void box_dt_kernel(const BoxDev* __restrict__ boxes, int nbox, const BoxStatus* __restrict__ prev,
	BoxStatus* __restrict__ next, double* __restrict__ dt_out, double cfl, double max_dt)
{
	const int b = blockIdx.x * blockDim.x + threadIdx.x;
	if (b < nbox) box_dt_body(boxes, prev, next, dt_out, cfl, max_dt, b);
}

/*
Flux pass of one face direction D over all boxes: the thread's cell c owns
the D-face on its negative side, between L = c - e_D and R = c; the flux is
computed when both cells have type >= 0 (get_fluxes,
source/mhd/solve.hpp:262-393) with bg = Bg_B(L)(+D), else the slot is 0.
*/
template <int SOLVER, int D, int TX, int TY>
__device__ __forceinline__ void box_flux_body(const BoxDev* __restrict__ boxes,
	const int* __restrict__ blockbox, long long N, const double* __restrict__ S,
	const signed char* __restrict__ ctype, const double* __restrict__ bg, double* __restrict__ P,
	const double* __restrict__ dts, BoxStatus* __restrict__ status, Phys ph, int tl)
{
	int b, i, j, k;
	const bool inside = my_cell<TX, TY>(boxes, blockbox, tl, b, i, j, k);
	double sp = 0;
	unsigned zero = 0;
	if (inside) {
		const BoxDev& B = boxes[b];
		const long long g = gidx(B, i, j, k);
		const int id = D == 0 ? i : (D == 1 ? j : k);
		double Fg[8] = {0, 0, 0, 0, 0, 0, 0, 0};
		if (id >= 1) {
			const long long gl = D == 0 ? g - 1 : (D == 1 ? g - B.nx : g - (long long)B.nx * B.ny);
			if (ctype[g] >= 0 && ctype[gl] >= 0) {
				double U[2][8];
				const long long cc[2] = {gl, g};
				for (int s = 0; s < 2; s++) {
					const long long c = cc[s];
					const double m[3] = {S[N + c], S[2 * N + c], S[3 * N + c]};
					const double vb[3] = {S[5 * N + c], S[6 * N + c], S[7 * N + c]};
					U[s][RHO] = S[c];
					U[s][NRJ] = S[4 * N + c];
					U[s][M0] = m[D];
					U[s][M1] = m[(D + 1) % 3];
					U[s][M2] = m[(D + 2) % 3];
					U[s][B0] = vb[D];
					U[s][B1] = vb[(D + 1) % 3];
					U[s][B2] = vb[(D + 2) % 3];
				}
				double bb[3] = {0, 0, 0};
				if (bg) {
					bb[0] = bg[(3 * D + D) * N + gl];
					bb[1] = bg[(3 * D + (D + 1) % 3) * N + gl];
					bb[2] = bg[(3 * D + (D + 2) % 3) * N + gl];
				}
				double F[8], s = 0;
				const int e = riemann<SOLVER, double>(U[0], U[1], bb, ph, F, s);
				if (e) {
					box_record_error(status + b, e, (long long)D * B.n + (g - B.off));
				} else {
					const double dt = dts[b];
					Fg[RHO] = dt * F[RHO];
					Fg[NRJ] = dt * F[NRJ];
					Fg[M0 + D] = dt * F[M0];
					Fg[M0 + (D + 1) % 3] = dt * F[M1];
					Fg[M0 + (D + 2) % 3] = dt * F[M2];
					Fg[B0 + D] = dt * F[B0];
					Fg[B0 + (D + 1) % 3] = dt * F[B1];
					Fg[B0 + (D + 2) % 3] = dt * F[B2];
					sp = s;
					zero = s <= 0 ? 1u : 0u;
				}
			}
		}
		double* Pd = P + (long long)D * 8 * N;
#pragma unroll
		for (int q = 0; q < 8; q++) Pd[q * N + g] = Fg[q];
	}
	/* per-box max speed and count of non-positive speeds; a block covers
	   cells of one box only */
	for (int off = 32; off > 0; off >>= 1) {
		sp = smax(sp, __shfl_xor(sp, off, 64));
		zero += __shfl_xor(zero, off, 64);
	}
	__shared__ double ms[TX * TY / 64];
	__shared__ unsigned mz[TX * TY / 64];
	const int tid = threadIdx.x + TX * threadIdx.y;
	if (tid % 64 == 0) {
		ms[tid / 64] = sp;
		mz[tid / 64] = zero;
	}
	__syncthreads();
	if (tid == 0) {
		double m = ms[0];
		unsigned z = mz[0];
		for (int w = 1; w < TX * TY / 64; w++) {
			m = smax(m, ms[w]);
			z += mz[w];
		}
		const int bb = blockbox[tl];
		if (m > 0) atomicMax(&status[bb].smax_bits[D], (unsigned long long)__double_as_longlong(m));
		if (z) atomicAdd(&status[bb].zero[D], z);
	}
}

template <int SOLVER, int D, int TX, int TY>
__global__ __launch_bounds__(TX * TY) void box_flux_kernel(const BoxDev* __restrict__ boxes,
	const int* __restrict__ blockbox, long long N, const double* __restrict__ S,
	const signed char* __restrict__ ctype, const double* __restrict__ bg, double* __restrict__ P,
	const double* __restrict__ dts, BoxStatus* __restrict__ status, Phys ph)
{
	box_flux_body<SOLVER, D, TX, TY>(boxes, blockbox, N, S, ctype, bg, P, dts, status, ph, (int)blockIdx.x);
}

/*
The three flux passes of box_flux_kernel in one launch: a thread computes
the fluxes of its cell's -x, -y and -z faces in turn (the cell's state read
once, the same operations per face as the per-direction kernel) and the
per-box speed reductions of the three directions.
*/
template <int SOLVER, int D>
__device__ __forceinline__ void box_face(const BoxDev& B, long long N, const double* __restrict__ S,
	const signed char* __restrict__ ctype, const double* __restrict__ bg, double* __restrict__ P,
	double dt, BoxStatus* __restrict__ st, int b, long long g, const int id[3], const double own[8],
	double& sp, unsigned& zero, const Phys& ph)
{
	double Fg[8] = {0, 0, 0, 0, 0, 0, 0, 0};
	if (id[D] >= 1) {
		const long long gl = D == 0 ? g - 1 : (D == 1 ? g - B.nx : g - (long long)B.nx * B.ny);
		if (ctype[g] >= 0 && ctype[gl] >= 0) {
			double U[2][8];
			const double lo[8] = {S[gl], S[N + gl], S[2 * N + gl], S[3 * N + gl], S[4 * N + gl],
				S[5 * N + gl], S[6 * N + gl], S[7 * N + gl]};
			const double* cs[2] = {lo, own};
			for (int s = 0; s < 2; s++) {
				const double* c = cs[s];
				U[s][RHO] = c[0];
				U[s][NRJ] = c[4];
				U[s][M0] = c[1 + D];
				U[s][M1] = c[1 + (D + 1) % 3];
				U[s][M2] = c[1 + (D + 2) % 3];
				U[s][B0] = c[5 + D];
				U[s][B1] = c[5 + (D + 1) % 3];
				U[s][B2] = c[5 + (D + 2) % 3];
			}
			double bb[3] = {0, 0, 0};
			if (bg) {
				bb[0] = bg[(3 * D + D) * N + gl];
				bb[1] = bg[(3 * D + (D + 1) % 3) * N + gl];
				bb[2] = bg[(3 * D + (D + 2) % 3) * N + gl];
			}
			double F[8], s = 0;
			const int e = riemann<SOLVER, double>(U[0], U[1], bb, ph, F, s);
			if (e) {
				box_record_error(st + b, e, (long long)D * B.n + (g - B.off));
			} else {
				Fg[RHO] = dt * F[RHO];
				Fg[NRJ] = dt * F[NRJ];
				Fg[M0 + D] = dt * F[M0];
				Fg[M0 + (D + 1) % 3] = dt * F[M1];
				Fg[M0 + (D + 2) % 3] = dt * F[M2];
				Fg[B0 + D] = dt * F[B0];
				Fg[B0 + (D + 1) % 3] = dt * F[B1];
				Fg[B0 + (D + 2) % 3] = dt * F[B2];
				sp = s;
				zero = s <= 0 ? 1u : 0u;
			}
		}
	}
	double* Pd = P + (long long)D * 8 * N;
#pragma unroll
	for (int q = 0; q < 8; q++) Pd[q * N + g] = Fg[q];
}

template <int SOLVER, int TX, int TY>
__device__ __forceinline__ void box_flux3_body(const BoxDev* __restrict__ boxes,
	const int* __restrict__ blockbox, long long N, const double* __restrict__ S,
	const signed char* __restrict__ ctype, const double* __restrict__ bg, double* __restrict__ P,
	const double* __restrict__ dts, BoxStatus* __restrict__ status, Phys ph, int tl)
{
	int b, i, j, k;
	const bool inside = my_cell<TX, TY>(boxes, blockbox, tl, b, i, j, k);
	double sp[3] = {0, 0, 0};
	unsigned zero[3] = {0, 0, 0};
	if (inside) {
		const BoxDev& B = boxes[b];
		const long long g = gidx(B, i, j, k);
		const int id[3] = {i, j, k};
		const double own[8] = {S[g], S[N + g], S[2 * N + g], S[3 * N + g], S[4 * N + g],
			S[5 * N + g], S[6 * N + g], S[7 * N + g]};
		const double dt = dts[b];
		box_face<SOLVER, 0>(B, N, S, ctype, bg, P, dt, status, b, g, id, own, sp[0], zero[0], ph);
		box_face<SOLVER, 1>(B, N, S, ctype, bg, P, dt, status, b, g, id, own, sp[1], zero[1], ph);
		box_face<SOLVER, 2>(B, N, S, ctype, bg, P, dt, status, b, g, id, own, sp[2], zero[2], ph);
	}
	/* per-box max speed and count of non-positive speeds per direction; a
	   block covers cells of one box only */
	__shared__ double ms[3][TX * TY / 64];
	__shared__ unsigned mz[3][TX * TY / 64];
	const int tid = threadIdx.x + TX * threadIdx.y;
	for (int d = 0; d < 3; d++) {
		double v = sp[d];
		unsigned z = zero[d];
		for (int off = 32; off > 0; off >>= 1) {
			v = smax(v, __shfl_xor(v, off, 64));
			z += __shfl_xor(z, off, 64);
		}
		if (tid % 64 == 0) {
			ms[d][tid / 64] = v;
			mz[d][tid / 64] = z;
		}
	}
	__syncthreads();
	if (tid < 3) {
		const int d = tid;
		double m = ms[d][0];
		unsigned z = mz[d][0];
		for (int w = 1; w < TX * TY / 64; w++) {
			m = smax(m, ms[d][w]);
			z += mz[d][w];
		}
		const int bb = blockbox[tl];
		if (m > 0) atomicMax(&status[bb].smax_bits[d], (unsigned long long)__double_as_longlong(m));
		if (z) atomicAdd(&status[bb].zero[d], z);
	}
}

template <int SOLVER, int TX, int TY>
__global__ __launch_bounds__(TX * TY) void box_flux3_kernel(const BoxDev* __restrict__ boxes,
	const int* __restrict__ blockbox, long long N, const double* __restrict__ S,
	const signed char* __restrict__ ctype, const double* __restrict__ bg, double* __restrict__ P,
	const double* __restrict__ dts, BoxStatus* __restrict__ status, Phys ph)
{
	box_flux3_body<SOLVER, TX, TY>(boxes, blockbox, N, S, ctype, bg, P, dts, status, ph, (int)blockIdx.x);
}

/*
Update of every cell in place (reads only the cell's own state and the
flux slots).  Normal cells: Face_dB of their six face copies from the
edge terms of assign_face_dBs_f{x,y,z} (source/mhd/solve.hpp:610-613,
:926-928, :1223-1225) in the order of the periodic step, face += dB / area,
conserved update of rho, m, E and Vol_B (update_mhd_state, :1551-1593).
Cells of type >= 0: refit at constant pressure (update_vol_B, :1904-1908).
With fsrc = Q_F2 the face copies come from the second set, synced here
(the previous step's last sync, which left them there), and every cell
writes its copies into the first set.
*/
template <int TX, int TY>
__device__ __forceinline__ void box_update_body(const BoxDev* __restrict__ boxes,
	const int* __restrict__ blockbox, long long N, double* __restrict__ S,
	const signed char* __restrict__ ctype, const double* __restrict__ P, Phys ph, int fsrc, int tl)
{
	int b, i, j, k;
	if (!my_cell<TX, TY>(boxes, blockbox, tl, b, i, j, k)) return;
	const BoxDev& B = boxes[b];
	const long long g = gidx(B, i, j, k);
	const int ct = ctype[g];
	const bool moved = fsrc != Q_F;
	if (ct < 0 && !moved) return;
	double f[6];
	if (moved) synced_faces(B, N, S, ctype, g, ct, i, j, k, fsrc, f);
	else
		for (int q = 0; q < 6; q++) f[q] = S[(Q_F + q) * N + g];
	if (ct > 0) {
		const double dx = B.dx, dy = B.dy, dz = B.dz;
		const long long sx = 1, sy = B.nx, sz = (long long)B.nx * B.ny;
		/* flux slot of the d-face owned by the cell at g + offset */
		auto PF = [&](int d, int q, long long o) { return P[((long long)d * 8 + q) * N + g + o]; };
		auto EZX = [&](long long o) { return +(PF(0, 6, o) * dz / 4); };
		auto EYX = [&](long long o) { return -(PF(0, 7, o) * dy / 4); };
		auto EXY = [&](long long o) { return +(PF(1, 7, o) * dx / 4); };
		auto EZY = [&](long long o) { return -(PF(1, 5, o) * dz / 4); };
		auto EXZ = [&](long long o) { return -(PF(2, 6, o) * dx / 4); };
		auto EYZ = [&](long long o) { return +(PF(2, 5, o) * dy / 4); };
		double dB[3][2];
		for (int side = 0; side < 2; side++) {
			long long I = side * sx;   /* x-face owned by (i + side, j, k) */
			double r = 0;
			r += EZX(I + sy); r -= EZX(I - sy); r -= EYX(I + sz); r += EYX(I - sz);
			r += EZY(I + sy); r += EZY(I - sx + sy); r -= EZY(I); r -= EZY(I - sx);
			r -= EYZ(I + sz); r -= EYZ(I - sx + sz); r += EYZ(I); r += EYZ(I - sx);
			dB[0][side] = r;
			const long long J = side * sy;   /* y-face owned by (i, j + side, k) */
			r = 0;
			r += EXY(J + sz); r -= EXY(J - sz); r -= EZY(J + sx); r += EZY(J - sx);
			r += EXZ(J + sz); r += EXZ(J - sy + sz); r -= EXZ(J); r -= EXZ(J - sy);
			r -= EZX(J + sx); r += EZX(J); r -= EZX(J + sx - sy); r += EZX(J - sy);
			dB[1][side] = r;
			const long long K = side * sz;   /* z-face owned by (i, j, k + side) */
			r = 0;
			r += EYZ(K + sx); r -= EYZ(K - sx); r -= EXZ(K + sy); r += EXZ(K - sy);
			r += EYX(K + sx); r += EYX(K + sx - sz); r -= EYX(K); r -= EYX(K - sz);
			r -= EXY(K + sy); r -= EXY(K + sy - sz); r += EXY(K); r += EXY(K - sz);
			dB[2][side] = r;
		}
		const double area[3] = {dy * dz, dx * dz, dx * dy};
		for (int d = 0; d < 3; d++)
		for (int side = 0; side < 2; side++) f[2 * d + side] = f[2 * d + side] + dB[d][side] / area[d];
		const double len[3] = {dx, dy, dz};
		const long long st[3] = {sx, sy, sz};
		double lo[3][8], hi[3][8];
		for (int d = 0; d < 3; d++)
		for (int q = 0; q < 8; q++) {
			lo[d][q] = 0.0 + PF(d, q, 0);
			hi[d][q] = 0.0 + PF(d, q, st[d]);
		}
		double rho = S[g];
		double m[3] = {S[N + g], S[2 * N + g], S[3 * N + g]};
		double E = S[4 * N + g];
		double vb[3] = {S[5 * N + g], S[6 * N + g], S[7 * N + g]};
		for (int d = 0; d < 3; d++) rho += (lo[d][RHO] - hi[d][RHO]) / len[d];
		for (int d = 0; d < 3; d++) {
			const double il = 1 / len[d];
			for (int q = 0; q < 3; q++) m[q] = m[q] + (lo[d][M0 + q] - hi[d][M0 + q]) * il;
		}
		for (int d = 0; d < 3; d++) E += (lo[d][NRJ] - hi[d][NRJ]) / len[d];
		for (int d = 0; d < 3; d++) {
			const double il = 1 / len[d];
			for (int q = 0; q < 3; q++) vb[q] = vb[q] + (lo[d][B0 + q] - hi[d][B0 + q]) * il;
		}
		S[g] = rho;
		for (int q = 0; q < 3; q++) S[(1 + q) * N + g] = m[q];
		S[4 * N + g] = E;
		for (int q = 0; q < 3; q++) S[(Q_VB + q) * N + g] = vb[q];
	}
	if (ct > 0 || moved)
		for (int q = 0; q < 6; q++) S[(Q_F + q) * N + g] = f[q];
	if (ct >= 0) refit_cell_faces(S, N, g, ph, f);
}

template <int TX, int TY>
__global__ __launch_bounds__(TX * TY) void box_update_kernel(const BoxDev* __restrict__ boxes,
	const int* __restrict__ blockbox, long long N, double* __restrict__ S,
	const signed char* __restrict__ ctype, const double* __restrict__ P, Phys ph, int fsrc)
{
	box_update_body<TX, TY>(boxes, blockbox, N, S, ctype, P, ph, fsrc, (int)blockIdx.x);
}

/*
sync_magnetic_field (source/common_functions.hpp:184-269) for the faces on
each cell's negative sides: the two copies of a face between cells of type
>= 0 are averaged as the reference's loop does it, first from the cell with
the smaller index (the negative-side one), then from the other.  One thread
per face writes both copies.
*/
template <int TX, int TY>
__device__ __forceinline__ void box_sync_body(const BoxDev* __restrict__ boxes,
	const int* __restrict__ blockbox, long long N, double* __restrict__ S,
	const signed char* __restrict__ ctype, int tl)
{
	int b, i, j, k;
	if (!my_cell<TX, TY>(boxes, blockbox, tl, b, i, j, k)) return;
	const BoxDev& B = boxes[b];
	const long long g = gidx(B, i, j, k);
	if (ctype[g] < 0) return;
	const int id[3] = {i, j, k};
	const long long st[3] = {1, B.nx, (long long)B.nx * B.ny};
	for (int d = 0; d < 3; d++) {
		if (id[d] < 1) continue;
		const long long ga = g - st[d];
		if (ctype[ga] < 0) continue;
		double a = S[(Q_F + 2 * d + 1) * N + ga];
		double o = S[(Q_F + 2 * d) * N + g];
		sync_pair(a, o);
		S[(Q_F + 2 * d + 1) * N + ga] = a;
		S[(Q_F + 2 * d) * N + g] = o;
	}
}

/*
The same sync from one set of face copies into another: each cell reads
its six copies and the other copy of each of its faces from the copies at
component fin and writes its six synced copies at component fout (cells of
type < 0 copy theirs through), so no thread reads what another writes and
one pass can also refit every cell of type >= 0 with its synced copies
(REFIT: the sync and refit stages of a step in one pass).  The two copies
of a face come out as box_sync_body leaves them: both cells apply
sync_pair to the same two values.
*/
template <int TX, int TY, bool REFIT>
__device__ __forceinline__ void box_sync_copy_body(const BoxDev* __restrict__ boxes,
	const int* __restrict__ blockbox, long long N, double* __restrict__ S,
	const signed char* __restrict__ ctype, Phys ph, int fin, int fout, int tl)
{
	int b, i, j, k;
	if (!my_cell<TX, TY>(boxes, blockbox, tl, b, i, j, k)) return;
	const BoxDev& B = boxes[b];
	const long long g = gidx(B, i, j, k);
	const int ct = ctype[g];
	double f[6];
	synced_faces(B, N, S, ctype, g, ct, i, j, k, fin, f);
	for (int q = 0; q < 6; q++) S[(fout + q) * N + g] = f[q];
	if (REFIT && ct >= 0) refit_cell_faces(S, N, g, ph, f);
}

template <int TX, int TY, bool REFIT>
__global__ __launch_bounds__(TX * TY) void box_sync_copy_kernel(const BoxDev* __restrict__ boxes,
	const int* __restrict__ blockbox, long long N, double* __restrict__ S,
	const signed char* __restrict__ ctype, Phys ph, int fin, int fout)
{
	box_sync_copy_body<TX, TY, REFIT>(boxes, blockbox, N, S, ctype, ph, fin, fout, (int)blockIdx.x);
}

template <int TX, int TY>
__global__ __launch_bounds__(TX * TY) void box_sync_kernel(const BoxDev* __restrict__ boxes,
	const int* __restrict__ blockbox, long long N, double* __restrict__ S,
	const signed char* __restrict__ ctype)
{
	box_sync_body<TX, TY>(boxes, blockbox, N, S, ctype, (int)blockIdx.x);
}

template <int TX, int TY>
__device__ __forceinline__ void box_refit_body(const BoxDev* __restrict__ boxes,
	const int* __restrict__ blockbox, long long N, double* __restrict__ S,
	const signed char* __restrict__ ctype, Phys ph, int tl)
{
	int b, i, j, k;
	if (!my_cell<TX, TY>(boxes, blockbox, tl, b, i, j, k)) return;
	const long long g = gidx(boxes[b], i, j, k);
	if (ctype[g] >= 0) refit_cell(S, N, g, ph);
}

template <int TX, int TY>
__global__ __launch_bounds__(TX * TY) void box_refit_kernel(const BoxDev* __restrict__ boxes,
	const int* __restrict__ blockbox, long long N, double* __restrict__ S,
	const signed char* __restrict__ ctype, Phys ph)
{
	box_refit_body<TX, TY>(boxes, blockbox, N, S, ctype, ph, (int)blockIdx.x);
}

/*
apply_boundaries_sw_box, source/mhd/solar_wind_box.hpp:306-541, and the
inflow cells of apply_solar_wind_boundaries, :219-273.  Outflow face, edge
and vertex cells copy rho, m and the thermal pressure of their interior
neighbour, with the face copies mirrored across the walls they touch;
planet cells get the inner-boundary state (energy computed with the
momentum in place of the velocity, as the reference does); inflow cells the
solar-wind state.  Copies read only normal cells, which no thread writes.
A thread per entry of the set's list of cells with a rule (cell g of box
b), the face copies at component fb.
*/
__device__ __forceinline__ void box_boundary_cell(const BoxDev* __restrict__ boxes, long long N,
	double* __restrict__ S, const unsigned char* __restrict__ rule, BoxStatus* __restrict__ status,
	Phys ph, int fb, int b, long long g)
{
	const BoxDev& B = boxes[b];
	const int r = rule[g];
	if (r == R_NONE) return;
	if (r == R_SW) {
		S[g] = B.sw_rho;
		for (int d = 0; d < 3; d++) S[(1 + d) * N + g] = B.sw_m[d];
		S[4 * N + g] = B.sw_E;
		for (int d = 0; d < 3; d++) {
			S[(Q_VB + d) * N + g] = B.sw_B[d];
			S[(fb + 2 * d) * N + g] = B.sw_B[d];
			S[(fb + 2 * d + 1) * N + g] = B.sw_B[d];
		}
		return;
	}
	if (r == R_PLANET) {
		S[g] = B.in_rho;
		for (int d = 0; d < 3; d++) S[(1 + d) * N + g] = B.in_m[d];
		S[4 * N + g] = B.in_E;
		for (int q = 0; q < 3; q++) S[(Q_VB + q) * N + g] = 0;
		for (int q = 0; q < 6; q++) S[(fb + q) * N + g] = 0;
		return;
	}
	const int o[3] = {(r - 1) % 3 - 1, ((r - 1) / 3) % 3 - 1, (r - 1) / 9 - 1};
	const long long src = g + o[0] + (long long)B.nx * (o[1] + (long long)B.ny * o[2]);
	if (S[src] <= 0) {
		box_record_error(status + b, ERR_DENSITY, 3LL * B.n + (g - B.off));
		return;
	}
	const double p = pressure_of(S, N, src, ph);
	S[g] = S[src];
	for (int q = 1; q <= 3; q++) S[q * N + g] = S[q * N + src];
	double vb[3];
	for (int d = 0; d < 3; d++) {
		const double lo = S[(fb + 2 * d) * N + src], hi = S[(fb + 2 * d + 1) * N + src];
		const double nlo = o[d] ? hi : lo, nhi = o[d] ? lo : hi;
		S[(fb + 2 * d) * N + g] = nlo;
		S[(fb + 2 * d + 1) * N + g] = nhi;
		vb[d] = 0.5 * (nlo + nhi);
		S[(Q_VB + d) * N + g] = vb[d];
	}
	const double rho = S[g];
	const double inv = 1 / rho;
	const double v[3] = {S[N + g] * inv, S[2 * N + g] * inv, S[3 * N + g] * inv};
	S[4 * N + g] = total_energy64(rho, v, p, vb, ph.gm1, ph.half_inv_mu0);
}

constexpr int LIST_BLOCK = 256;

__global__ __launch_bounds__(LIST_BLOCK) void box_boundary_kernel(const BoxDev* __restrict__ boxes,
	const CellRef* __restrict__ list, long long count, long long N, double* __restrict__ S,
	const unsigned char* __restrict__ rule, BoxStatus* __restrict__ status, Phys ph, int fb)
{
	const long long e = (long long)blockIdx.x * LIST_BLOCK + threadIdx.x;
	if (e < count) box_boundary_cell(boxes, N, S, rule, status, ph, fb, list[e].b, list[e].g);
}

/* the face each inflow cell shares with its interior neighbour gets the
   solar-wind normal field, with that neighbour's Vol_B component and energy
   refit at constant pressure (source/mhd/solar_wind_box.hpp:275-299); a
   thread per entry of the set's list of inflow cells */
__device__ __forceinline__ void box_inflow_face_cell(const BoxDev* __restrict__ boxes, long long N,
	double* __restrict__ S, const unsigned char* __restrict__ rule, BoxStatus* __restrict__ status,
	Phys ph, int fb, int b, long long g)
{
	const BoxDev& B = boxes[b];
	if (rule[g] != R_SW) return;
	const long long l = g - B.off;
	const int i = (int)(l % B.nx), j = (int)((l / B.nx) % B.ny), k = (int)(l / ((long long)B.nx * B.ny));
	const int id[3] = {i, j, k};
	const int dims[3] = {B.nx, B.ny, B.nz};
	const int ni = id[B.sw_dim] - B.sw_sg;
	if (ni < 0 || ni >= dims[B.sw_dim]) return;
	const long long st[3] = {1, B.nx, (long long)B.nx * B.ny};
	const long long nb = g - B.sw_sg * st[B.sw_dim];
	const double rho = S[nb];
	if (rho <= 0) {
		box_record_error(status + b, ERR_DENSITY, 3LL * B.n + (nb - B.off));
		return;
	}
	const double inv = 1 / rho;
	const double v[3] = {S[N + nb] * inv, S[2 * N + nb] * inv, S[3 * N + nb] * inv};
	const double p = pressure_of(S, N, nb, ph);
	const int toward = fb + 2 * B.sw_dim + (B.sw_sg > 0), away = fb + 2 * B.sw_dim + (B.sw_sg < 0);
	S[toward * N + nb] = B.sw_B[B.sw_dim];
	S[(Q_VB + B.sw_dim) * N + nb] = S[toward * N + nb] / 2 + S[away * N + nb] / 2;
	const double vb[3] = {S[5 * N + nb], S[6 * N + nb], S[7 * N + nb]};
	S[4 * N + nb] = total_energy64(rho, v, p, vb, ph.gm1, ph.half_inv_mu0);
}

__global__ __launch_bounds__(LIST_BLOCK) void box_inflow_face_kernel(const BoxDev* __restrict__ boxes,
	const CellRef* __restrict__ list, long long count, long long N, double* __restrict__ S,
	const unsigned char* __restrict__ rule, BoxStatus* __restrict__ status, Phys ph, int fb)
{
	const long long e = (long long)blockIdx.x * LIST_BLOCK + threadIdx.x;
	if (e < count) box_inflow_face_cell(boxes, N, S, rule, status, ph, fb, list[e].b, list[e].g);
}

/* sum of |div B| of normal cells per block, source/math/nabla.hpp:58-113 */
template <int TX, int TY>
__global__ __launch_bounds__(TX * TY) void box_div_kernel(const BoxDev* __restrict__ boxes,
	const int* __restrict__ blockbox, long long N, const double* __restrict__ S,
	const signed char* __restrict__ ctype, int box, double* __restrict__ partial)
{
	int b, i, j, k;
	double v = 0;
	const int tl = (int)blockIdx.x;
	if (my_cell<TX, TY>(boxes, blockbox, tl, b, i, j, k) && b == box) {
		const BoxDev& B = boxes[b];
		const long long g = gidx(B, i, j, k);
		if (ctype[g] > 0) {
			v = fabs((S[(Q_F + 1) * N + g] - S[Q_F * N + g]) / B.dx
				+ (S[(Q_F + 3) * N + g] - S[(Q_F + 2) * N + g]) / B.dy
				+ (S[(Q_F + 5) * N + g] - S[(Q_F + 4) * N + g]) / B.dz);
		}
	}
	for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
	__shared__ double w[TX * TY / 64];
	const int tid = threadIdx.x + TX * threadIdx.y;
	if (tid % 64 == 0) w[tid / 64] = v;
	__syncthreads();
	if (tid == 0) {
		double t = 0;
		for (int q = 0; q < TX * TY / 64; q++) t += w[q];
		partial[blockIdx.x] = t;
	}
}

/*
Persistent run: one cooperative launch advances every box by `steps`
main-loop steps (the stages of boxset_launch_step with boundaries), each
stage walked by the resident workgroups over the set's tiles and separated
from the next by a grid-wide barrier.  The barrier counts arrivals in
bar[0] with release / acquire fences at agent scope (a stage's writes
reach the other XCDs before the next stage reads them).  Its wait is
bounded: after RUN_SPIN_LIMIT polls a workgroup sets bar[1] and leaves, and
every workgroup that sees bar[1] leaves too, so all waves exit and the host
reports the error.  The per-tile work is the stage kernels' own, so the
results are theirs bit for bit.
*/
constexpr unsigned RUN_SPIN_LIMIT = 1u << 24;

__device__ __forceinline__ bool run_barrier(unsigned* bar, unsigned target, unsigned polls)
{
	__shared__ int ok;
	__syncthreads();
	if (threadIdx.x == 0 && threadIdx.y == 0) {
		__threadfence();
		atomicAdd(&bar[0], 1u);
		int good = 1;
		for (unsigned spin = 0;; spin++) {
			if (__hip_atomic_load(&bar[0], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) >= target) break;
			if (__hip_atomic_load(&bar[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
				good = 0;
				break;
			}
			if (spin >= polls) {
				atomicExch(&bar[1], 1u);
				good = 0;
				break;
			}
			__builtin_amdgcn_s_sleep(2);
		}
		__threadfence();
		ok = good;
	}
	__syncthreads();
	return ok != 0;
}

struct RunArgs {
	const BoxDev* boxes;
	const int* blockbox;
	long long N;
	double* S;
	const signed char* ctype;
	const unsigned char* rule;
	const CellRef* bnd;
	const CellRef* inflow;
	long long nbnd, ninflow;
	const double* bg;
	double* P;
	double* dt;
	BoxStatus* status;
	int nbox, nblocks;
	Phys ph;
	double cfl;
};

/* the kernel walks stage after stage (per step: dt, flux, update, sync +
   refit into the second face set, boundaries and inflow faces on it, and
   after the last step the sync back, which the next step's update does
   for the others) with the stage picked by a switch, so each stage's
   values live only in its own case: the kernel's registers are about the
   largest stage's, not every stage's at once */
template <int SOLVER, int TX, int TY>
__global__ __launch_bounds__(TX * TY) void box_run_kernel(RunArgs a, long steps, int last, unsigned* bar,
	unsigned polls)
{
	const int tid = threadIdx.x + TX * threadIdx.y;
	unsigned passed = 0;   /* grid barriers passed */
	for (long q = 0; q < 7 * steps; q++) {
		/* the last step syncs back into the first set; the others leave
		   it to the next step's update (no stage, no barrier) */
		if (q % 7 == 6 && q + 1 != 7 * steps) continue;
		const int lst = (int)((last + q / 7) & 1);
		BoxStatus* prev = a.status + (long long)lst * a.nbox;
		BoxStatus* next = a.status + (long long)(1 - lst) * a.nbox;
		switch ((int)(q % 7)) {
		case 0:
			for (int b = (int)blockIdx.x * TX * TY + tid; b < a.nbox; b += (int)gridDim.x * TX * TY)
				box_dt_body(a.boxes, prev, next, a.dt, a.cfl, 1e300, b);
			break;
		case 1:
			for (int t = blockIdx.x; t < a.nblocks; t += gridDim.x)
				box_flux3_body<SOLVER, TX, TY>(a.boxes, a.blockbox, a.N, a.S, a.ctype, a.bg, a.P, a.dt, next, a.ph, t);
			break;
		case 2: {
			/* the steps after the first take the face copies the previous
			   step left in the second set */
			const int fsrc = q >= 7 ? Q_F2 : Q_F;
			for (int t = blockIdx.x; t < a.nblocks; t += gridDim.x)
				box_update_body<TX, TY>(a.boxes, a.blockbox, a.N, a.S, a.ctype, a.P, a.ph, fsrc, t);
			break;
		}
		case 3:
			for (int t = blockIdx.x; t < a.nblocks; t += gridDim.x)
				box_sync_copy_body<TX, TY, true>(a.boxes, a.blockbox, a.N, a.S, a.ctype, a.ph, Q_F, Q_F2, t);
			break;
		case 4:
			for (long long e = (long long)blockIdx.x * TX * TY + tid; e < a.nbnd; e += (long long)gridDim.x * TX * TY)
				box_boundary_cell(a.boxes, a.N, a.S, a.rule, next, a.ph, Q_F2, a.bnd[e].b, a.bnd[e].g);
			break;
		case 5:
			for (long long e = (long long)blockIdx.x * TX * TY + tid; e < a.ninflow; e += (long long)gridDim.x * TX * TY)
				box_inflow_face_cell(a.boxes, a.N, a.S, a.rule, next, a.ph, Q_F2, a.inflow[e].b, a.inflow[e].g);
			break;
		default:
			for (int t = blockIdx.x; t < a.nblocks; t += gridDim.x)
				box_sync_copy_body<TX, TY, false>(a.boxes, a.blockbox, a.N, a.S, a.ctype, a.ph, Q_F2, Q_F, t);
			break;
		}
		if (!run_barrier(bar, ++passed * gridDim.x, polls)) return;
	}
}

__global__ void run_barrier_clear_kernel(unsigned* bar)
{
	if (threadIdx.x < 2) bar[threadIdx.x] = 0;
}

/* ---------------------------------------------------------------------- */
/* host side                                                              */
/* ---------------------------------------------------------------------- */

inline double cmin(const pamhd_grid_desc& G, int d, long i) { return G.start[d] + (double)i * G.cell_length[d]; }
inline double cmax(const pamhd_grid_desc& G, int d, long i) { return G.start[d] + (double)(i + 1) * G.cell_length[d]; }
inline double ccen(const pamhd_grid_desc& G, int d, long i) { return G.start[d] + ((double)i + 0.5) * G.cell_length[d]; }

/* constant field plus point dipoles, source/background_magnetic_field.hpp:331-494 */
void dipole_field(const pamhd_solar_wind_desc& w, const double pos[3], double mu0, double out[3])
{
	double r3[3] = {w.bg_value[0], w.bg_value[1], w.bg_value[2]};
	for (int q = 0; q < w.n_dipoles; q++) {
		const double* m = w.dipole_moment[q];
		const double* dp = w.dipole_position[q];
		const double r[3] = {pos[0] + (-dp[0]), pos[1] + (-dp[1]), pos[2] + (-dp[2])};
		const double r1 = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
		if (r1 < w.bg_min_distance) {
			const double f = mu0 / (2 * M_PI * std::pow(w.bg_min_distance, 3));
			for (int d = 0; d < 3; d++) r3[d] = r3[d] + m[d] * f;
		} else {
			const double inv = 1 / r1;
			const double u[3] = {r[0] * inv, r[1] * inv, r[2] * inv};
			const double pr = m[0] * u[0] + m[1] * u[1] + m[2] * u[2];
			const double f = mu0 / (4 * M_PI) / (r1 * r1 * r1);
			for (int d = 0; d < 3; d++) {
				const double proj = pr * u[d];
				r3[d] = r3[d] + (3 * proj + (-m[d])) * f;
			}
		}
	}
	for (int d = 0; d < 3; d++) out[d] = r3[d];
}

/* prepare_grid's classification for a uniform grid,
   source/grid/solar_wind_box.hpp:262-406 */
void classify(const pamhd_grid_desc& G, const pamhd_solar_wind_desc& w, signed char* ctype,
	unsigned char* rule)
{
	const long nx = G.cells[0], ny = G.cells[1], nz = G.cells[2];
	const long dims[3] = {nx, ny, nz};
	const int sw_dim = std::abs(w.direction) - 1;
	auto id = [&](long i, long j, long k) { return i + nx * (j + ny * k); };
	for (long k = 0; k < nz; k++)
	for (long j = 0; j < ny; j++)
	for (long i = 0; i < nx; i++) {
		const long idx[3] = {i, j, k};
		const long c = id(i, j, k);
		ctype[c] = 1;
		rule[c] = R_NONE;
		int side[3];
		for (int d = 0; d < 3; d++) {
			side[d] = idx[d] == 0 ? -1 : (idx[d] == dims[d] - 1 ? +1 : 0);
			if (dims[d] == 1) side[d] = 0;
		}
		if (side[0] || side[1] || side[2]) {
			ctype[c] = 0;
			rule[c] = (unsigned char)(1 + (1 - side[0]) + 3 * (1 - side[1]) + 9 * (1 - side[2]));
		}
		if (idx[sw_dim] == (w.direction < 0 ? 0 : dims[sw_dim] - 1)) {
			ctype[c] = 0;
			rule[c] = R_SW;
		}
	}
	/* planet: a corner within the inner radius, or the origin inside */
	for (long k = 0; k < nz; k++)
	for (long j = 0; j < ny; j++)
	for (long i = 0; i < nx; i++) {
		const double s[3] = {cmin(G, 0, i), cmin(G, 1, j), cmin(G, 2, k)};
		const double e[3] = {cmax(G, 0, i), cmax(G, 1, j), cmax(G, 2, k)};
		bool in = false;
		for (int q = 0; q < 8; q++) {
			const double x = (q & 4) ? e[0] : s[0], y = (q & 2) ? e[1] : s[1], z = (q & 1) ? e[2] : s[2];
			if (std::sqrt(x * x + y * y + z * z) <= w.inner_radius) in = true;
		}
		if (s[0] <= 0 && e[0] >= 0 && s[1] <= 0 && e[1] >= 0 && s[2] <= 0 && e[2] >= 0) in = true;
		if (in) {
			ctype[id(i, j, k)] = 0;
			rule[id(i, j, k)] = R_PLANET;
		}
	}
	for (long k = 0; k < nz; k++)
	for (long j = 0; j < ny; j++)
	for (long i = 0; i < nx; i++) {
		const long c = id(i, j, k);
		if (rule[c] != R_PLANET) continue;
		bool have_normal = false;
		for (int dz = -1; dz <= 1; dz++)
		for (int dy = -1; dy <= 1; dy++)
		for (int dx = -1; dx <= 1; dx++) {
			const int nn = (dx != 0) + (dy != 0) + (dz != 0);
			if (nn == 0 || nn == 3) continue;
			const long a = i + dx, bb = j + dy, cc = k + dz;
			if (a < 0 || bb < 0 || cc < 0 || a >= nx || bb >= ny || cc >= nz) continue;
			if (ctype[id(a, bb, cc)] == 1) have_normal = true;
		}
		if (!have_normal) {
			ctype[c] = -1;
			rule[c] = R_NONE;
		}
	}
}

/* the boundary rules, then the inflow faces, a thread per listed cell */
void launch_boundary_lists(BoxSet* s, hipStream_t st, BoxStatus* stat, int fb)
{
	if (s->nbnd > 0)
		box_boundary_kernel<<<(unsigned)((s->nbnd + LIST_BLOCK - 1) / LIST_BLOCK), LIST_BLOCK, 0, st>>>(
			s->boxes, s->bnd, s->nbnd, s->N, s->S, s->rule, stat, s->ph, fb);
	if (s->ninflow > 0)
		box_inflow_face_kernel<<<(unsigned)((s->ninflow + LIST_BLOCK - 1) / LIST_BLOCK), LIST_BLOCK, 0, st>>>(
			s->boxes, s->inflow, s->ninflow, s->N, s->S, s->rule, stat, s->ph, fb);
}

/* stage 4's flags: the update takes the face copies from the second set
   (the previous step of a chain left them there), and the step leaves its
   copies in the second set for the next step of the chain */
enum { CH_FROM2 = 1, CH_KEEP2 = 2 };

template <int TX, int TY> void launch_all(BoxSet* s, int solver, hipStream_t st, int stage,
	BoxStatus* stat, int chain)
{
	const dim3 grid(s->nblocks), block(TX, TY, 1);
	switch (stage) {
	case 0:
#define FLUX(SV) \
		if (s->flux3) { \
			box_flux3_kernel<SV, TX, TY><<<grid, block, 0, st>>>(s->boxes, s->blockbox, s->N, s->S, s->ctype, s->bg, s->P, s->dt, stat, s->ph); \
		} else { \
			box_flux_kernel<SV, 0, TX, TY><<<grid, block, 0, st>>>(s->boxes, s->blockbox, s->N, s->S, s->ctype, s->bg, s->P, s->dt, stat, s->ph); \
			box_flux_kernel<SV, 1, TX, TY><<<grid, block, 0, st>>>(s->boxes, s->blockbox, s->N, s->S, s->ctype, s->bg, s->P, s->dt, stat, s->ph); \
			box_flux_kernel<SV, 2, TX, TY><<<grid, block, 0, st>>>(s->boxes, s->blockbox, s->N, s->S, s->ctype, s->bg, s->P, s->dt, stat, s->ph); \
		}
		switch (solver) {
		case RUSANOV: FLUX(RUSANOV); break;
		case HLL_ATHENA: FLUX(HLL_ATHENA); break;
		case HLLD_ATHENA: FLUX(HLLD_ATHENA); break;
		case ROE_ATHENA: FLUX(ROE_ATHENA); break;
		}
#undef FLUX
		break;
	case 1:
		box_update_kernel<TX, TY><<<grid, block, 0, st>>>(s->boxes, s->blockbox, s->N, s->S, s->ctype, s->P, s->ph, Q_F);
		box_sync_kernel<TX, TY><<<grid, block, 0, st>>>(s->boxes, s->blockbox, s->N, s->S, s->ctype);
		box_refit_kernel<TX, TY><<<grid, block, 0, st>>>(s->boxes, s->blockbox, s->N, s->S, s->ctype, s->ph);
		break;
	case 2:
		launch_boundary_lists(s, st, stat, Q_F);
		box_sync_kernel<TX, TY><<<grid, block, 0, st>>>(s->boxes, s->blockbox, s->N, s->S, s->ctype);
		break;
	case 3:
		box_refit_kernel<TX, TY><<<grid, block, 0, st>>>(s->boxes, s->blockbox, s->N, s->S, s->ctype, s->ph);
		break;
	case 4:
		/* stages 1 and 2 with the second face set: sync + refit into it,
		   the boundaries on it, the last sync back (or, inside a chain,
		   in the next step's update) */
		box_update_kernel<TX, TY><<<grid, block, 0, st>>>(s->boxes, s->blockbox, s->N, s->S, s->ctype, s->P, s->ph,
			(chain & CH_FROM2) ? Q_F2 : Q_F);
		box_sync_copy_kernel<TX, TY, true><<<grid, block, 0, st>>>(s->boxes, s->blockbox, s->N, s->S, s->ctype, s->ph, Q_F, Q_F2);
		launch_boundary_lists(s, st, stat, Q_F2);
		if (!(chain & CH_KEEP2))
			box_sync_copy_kernel<TX, TY, false><<<grid, block, 0, st>>>(s->boxes, s->blockbox, s->N, s->S, s->ctype, s->ph, Q_F2, Q_F);
		break;
	}
}

void launch_stage(BoxSet* s, int solver, hipStream_t st, int stage, BoxStatus* stat, int chain = 0)
{
	switch (s->tile) {
	case 0: launch_all<64, 4>(s, solver, st, stage, stat, chain); break;
	case 1: launch_all<32, 8>(s, solver, st, stage, stat, chain); break;
	default: launch_all<16, 16>(s, solver, st, stage, stat, chain); break;
	}
}

/* block -> box table for the current tile shape */
hipError_t build_blocks(BoxSet* s)
{
	const Tile t = TILES[s->tile];
	std::vector<int> bb;
	for (int b = 0; b < s->nbox; b++) {
		BoxDev& B = s->host[b];
		B.gx = (B.nx + t.tx - 1) / t.tx;
		B.gy = (B.ny + t.ty - 1) / t.ty;
		B.blk0 = (int)bb.size();
		const long long nb = (long long)B.gx * B.gy * B.nz;
		for (long long q = 0; q < nb; q++) bb.push_back(b);
	}
	s->nblocks = (int)bb.size();
	(void)hipFree(s->blockbox);
	s->blockbox = nullptr;
	hipError_t e;
	if ((e = hipMalloc(&s->blockbox, sizeof(int) * bb.size())) != hipSuccess) return e;
	if ((e = hipMemcpy(s->blockbox, bb.data(), sizeof(int) * bb.size(), hipMemcpyHostToDevice)) != hipSuccess) return e;
	(void)hipFree(s->partial);
	s->partial = nullptr;
	if ((e = hipMalloc(&s->partial, sizeof(double) * bb.size())) != hipSuccess) return e;
	return hipMemcpy(s->boxes, s->host.data(), sizeof(BoxDev) * s->nbox, hipMemcpyHostToDevice);
}

/* copies box b's 14 * n values between a host array and the set */
hipError_t copy_box(BoxSet* s, int box, double* host, bool to_host)
{
	const BoxDev& B = s->host[box];
	for (int q = 0; q < NQ; q++) {
		double* dev = s->S + (long long)q * s->N + B.off;
		double* h = host + (long long)q * B.n;
		const hipError_t e = to_host
			? hipMemcpy(h, dev, sizeof(double) * B.n, hipMemcpyDeviceToHost)
			: hipMemcpy(dev, h, sizeof(double) * B.n, hipMemcpyHostToDevice);
		if (e != hipSuccess) return e;
	}
	return hipSuccess;
}

} // namespace

int boxset_count(const BoxSet* s) { return s->nbox; }
long long boxset_faces(const BoxSet* s, int box)
{
	return (box >= 0 && box < (int)s->faces.size()) ? s->faces[box] : 0;
}

long long boxset_cells(const BoxSet* s, int box) { return s->host[box].n; }
int boxset_tile(const BoxSet* s) { return s->tile; }

int boxset_set_tile(BoxSet* s, int tile)
{
	if (tile < 0 || tile > 2) return PAMHD_E_ARG;
	s->tile = tile;
	return build_blocks(s) == hipSuccess ? PAMHD_OK : PAMHD_E_DEVICE;
}

void boxset_destroy(BoxSet* s)
{
	if (!s) return;
	(void)hipFree(s->boxes);
	(void)hipFree(s->blockbox);
	(void)hipFree(s->S);
	(void)hipFree(s->P);
	(void)hipFree(s->bg);
	(void)hipFree(s->ctype);
	(void)hipFree(s->rule);
	(void)hipFree(s->status);
	(void)hipFree(s->dt);
	(void)hipFree(s->partial);
	(void)hipFree(s->bar);
	(void)hipFree(s->bnd);
	(void)hipFree(s->inflow);
	(void)hipHostFree(s->host_status);
	(void)hipHostFree(s->host_dt);
	delete s;
}

int boxset_create(BoxSet** out, int nbox, const pamhd_grid_desc* grids, const pamhd_physics* phys,
	const pamhd_solar_wind_desc* sw, std::string& err)
{
	*out = nullptr;
	if (nbox < 1 || !grids || !phys || !sw) {
		err = "null argument or no boxes";
		return PAMHD_E_ARG;
	}
	if (!(phys->adiabatic_index > 1) || !(phys->vacuum_permeability > 0) || !(phys->proton_mass > 0)) {
		err = "adiabatic index must be > 1, vacuum permeability and proton mass > 0";
		return PAMHD_E_ARG;
	}
	long long total = 0;
	for (int b = 0; b < nbox; b++) {
		const pamhd_grid_desc& G = grids[b];
		const pamhd_solar_wind_desc& w = sw[b];
		for (int d = 0; d < 3; d++) {
			if (G.cells[d] < 4) {
				err = "box " + std::to_string(b) + ": solar-wind boxes need >= 4 cells per dimension";
				return PAMHD_E_ARG;
			}
			if (G.periodic[d]) {
				err = "box " + std::to_string(b) + ": solar-wind boxes are not periodic";
				return PAMHD_E_ARG;
			}
			if (!(G.cell_length[d] > 0)) {
				err = "cell length must be > 0";
				return PAMHD_E_ARG;
			}
		}
		if (w.direction == 0 || std::abs(w.direction) > 3) {
			err = "solar-wind direction must be +-1, +-2 or +-3";
			return PAMHD_E_ARG;
		}
		/* source/mhd/solar_wind_box.hpp:95-103 */
		if (w.magnetic_field[std::abs(w.direction) - 1] != 0) {
			err = "solar-wind magnetic field along the inflow direction must be 0";
			return PAMHD_E_ARG;
		}
		if (!(w.number_density > 0) || !(w.inner_number_density > 0)) {
			err = "number densities must be > 0";
			return PAMHD_E_ARG;
		}
		if (w.n_dipoles < 0 || w.n_dipoles > PAMHD_MAX_DIPOLES) {
			err = "too many dipoles";
			return PAMHD_E_ARG;
		}
		total += (long long)G.cells[0] * G.cells[1] * G.cells[2];
	}
	if (total >= (1ll << 31)) {
		err = "more than 2^31 cells in a set";
		return PAMHD_E_UNSUPPORTED;
	}
	BoxSet* s = new BoxSet();
	s->nbox = nbox;
	s->N = total;
	s->ph = make_phys(phys->adiabatic_index, phys->vacuum_permeability);
	s->mp = phys->proton_mass;
	s->sw.assign(sw, sw + nbox);
	const long long N = total;
	auto bail = [&](hipError_t e, const char* what) {
		err = std::string(what) + ": " + hipGetErrorString(e);
		boxset_destroy(s);
		return PAMHD_E_DEVICE;
	};
	hipError_t e;
	if ((e = hipMalloc(&s->S, sizeof(double) * NQ_ALLOC * N)) != hipSuccess) return bail(e, "state");
	if ((e = hipMalloc(&s->P, sizeof(double) * 24 * N)) != hipSuccess) return bail(e, "flux");
	if ((e = hipMalloc(&s->bg, sizeof(double) * 9 * N)) != hipSuccess) return bail(e, "bg");
	if ((e = hipMalloc(&s->ctype, N)) != hipSuccess) return bail(e, "ctype");
	if ((e = hipMalloc(&s->rule, N)) != hipSuccess) return bail(e, "rule");
	if ((e = hipMalloc(&s->status, sizeof(BoxStatus) * 2 * nbox)) != hipSuccess) return bail(e, "status");
	if ((e = hipMalloc(&s->dt, sizeof(double) * nbox)) != hipSuccess) return bail(e, "dt");
	if ((e = hipMalloc(&s->bar, sizeof(unsigned) * 2)) != hipSuccess) return bail(e, "barrier");
	if ((e = hipMalloc(&s->boxes, sizeof(BoxDev) * nbox)) != hipSuccess) return bail(e, "boxes");
	if ((e = hipHostMalloc(&s->host_status, sizeof(BoxStatus) * nbox)) != hipSuccess) return bail(e, "pinned");
	if ((e = hipHostMalloc(&s->host_dt, sizeof(double) * nbox)) != hipSuccess) return bail(e, "pinned");

	/* per box: classification, background field, initialize_plasma
	   (source/mhd/solar_wind_box.hpp:62-205) */
	long long off = 0;
	const Phys& ph = s->ph;
	std::vector<CellRef> bnd, inflow;
	for (int b = 0; b < nbox; b++) {
		const pamhd_grid_desc& G = grids[b];
		const pamhd_solar_wind_desc& w = sw[b];
		BoxDev B{};
		B.nx = G.cells[0];
		B.ny = G.cells[1];
		B.nz = G.cells[2];
		B.n = B.nx * B.ny * B.nz;
		B.off = off;
		B.dx = G.cell_length[0];
		B.dy = G.cell_length[1];
		B.dz = G.cell_length[2];
		B.sw_dim = std::abs(w.direction) - 1;
		B.sw_sg = w.direction > 0 ? 1 : -1;
		const double mass = w.number_density * s->mp;
		B.sw_rho = mass;
		for (int d = 0; d < 3; d++) B.sw_m[d] = mass * w.velocity[d];
		B.sw_E = total_energy64(mass, w.velocity, w.pressure, w.magnetic_field, ph.gm1, ph.half_inv_mu0);
		for (int d = 0; d < 3; d++) B.sw_B[d] = w.magnetic_field[d];
		const double imass = w.inner_number_density * s->mp;
		B.in_rho = imass;
		for (int d = 0; d < 3; d++) B.in_m[d] = imass * w.inner_velocity[d];
		const double zero3[3] = {0, 0, 0};
		B.in_E = total_energy64(imass, B.in_m, w.inner_pressure, zero3, ph.gm1, ph.half_inv_mu0);
		s->host.push_back(B);

		const long n = B.n;
		std::vector<signed char> ct(n);
		std::vector<unsigned char> ru(n);
		classify(G, w, ct.data(), ru.data());
		for (long c = 0; c < n; c++) {
			if (ru[c] == R_NONE) continue;
			bnd.push_back(CellRef{off + c, b, 0});
			if (ru[c] == R_SW) inflow.push_back(CellRef{off + c, b, 0});
		}
		{
			long long nf = 0;
			for (long k = 0; k < B.nz; k++)
			for (long j = 0; j < B.ny; j++)
			for (long i = 0; i < B.nx; i++) {
				const long c = i + B.nx * (j + (long)B.ny * k);
				if (ct[c] < 0) continue;
				if (i > 0 && ct[c - 1] >= 0) nf++;
				if (j > 0 && ct[c - B.nx] >= 0) nf++;
				if (k > 0 && ct[c - (long)B.nx * B.ny] >= 0) nf++;
			}
			s->faces.push_back(nf);
		}
		std::vector<double> bg(9 * (size_t)n), st(NQ * (size_t)n);
		const double end_x = G.start[0] + (double)G.cells[0] * G.cell_length[0];
		for (long k = 0; k < B.nz; k++)
		for (long j = 0; j < B.ny; j++)
		for (long i = 0; i < B.nx; i++) {
			const long c = i + B.nx * (j + (long)B.ny * k);
			const long idx[3] = {i, j, k};
			for (int d = 0; d < 3; d++) {
				double p[3] = {ccen(G, 0, i), ccen(G, 1, j), ccen(G, 2, k)};
				p[d] = cmax(G, d, idx[d]);
				double v[3];
				dipole_field(w, p, phys->vacuum_permeability, v);
				for (int q = 0; q < 3; q++) bg[(3 * d + q) * (size_t)n + c] = v[q];
			}
			const double ratio = ccen(G, 0, i) / end_x;
			const double vf = (0.0 < ratio) ? ratio : 0.0;
			const double v[3] = {vf * w.velocity[0], vf * w.velocity[1], vf * w.velocity[2]};
			st[c] = mass;
			for (int d = 0; d < 3; d++) st[(1 + d) * (size_t)n + c] = mass * v[d];
			for (int d = 0; d < 3; d++) {
				st[(Q_VB + d) * (size_t)n + c] = w.magnetic_field[d];
				st[(Q_F + 2 * d) * (size_t)n + c] = w.magnetic_field[d];
				st[(Q_F + 2 * d + 1) * (size_t)n + c] = w.magnetic_field[d];
			}
			st[4 * (size_t)n + c] = total_energy64(mass, v, w.pressure, w.magnetic_field, ph.gm1, ph.half_inv_mu0);
		}
		if ((e = hipMemcpy(s->ctype + off, ct.data(), n, hipMemcpyHostToDevice)) != hipSuccess) return bail(e, "ctype");
		if ((e = hipMemcpy(s->rule + off, ru.data(), n, hipMemcpyHostToDevice)) != hipSuccess) return bail(e, "rule");
		for (int q = 0; q < 9; q++) {
			if ((e = hipMemcpy(s->bg + (long long)q * N + off, bg.data() + (size_t)q * n, sizeof(double) * n,
				hipMemcpyHostToDevice)) != hipSuccess) return bail(e, "bg");
		}
		for (int q = 0; q < NQ; q++) {
			if ((e = hipMemcpy(s->S + (long long)q * N + off, st.data() + (size_t)q * n, sizeof(double) * n,
				hipMemcpyHostToDevice)) != hipSuccess) return bail(e, "state");
		}
		off += n;
	}
	if ((e = build_blocks(s)) != hipSuccess) return bail(e, "blocks");
	s->nbnd = (long long)bnd.size();
	s->ninflow = (long long)inflow.size();
	if (s->nbnd > 0) {
		if ((e = hipMalloc(&s->bnd, sizeof(CellRef) * bnd.size())) != hipSuccess) return bail(e, "boundary list");
		if ((e = hipMemcpy(s->bnd, bnd.data(), sizeof(CellRef) * bnd.size(), hipMemcpyHostToDevice)) != hipSuccess)
			return bail(e, "boundary list");
	}
	if (s->ninflow > 0) {
		if ((e = hipMalloc(&s->inflow, sizeof(CellRef) * inflow.size())) != hipSuccess) return bail(e, "inflow list");
		if ((e = hipMemcpy(s->inflow, inflow.data(), sizeof(CellRef) * inflow.size(), hipMemcpyHostToDevice)) != hipSuccess)
			return bail(e, "inflow list");
	}
	if ((e = hipMemset(s->P, 0, sizeof(double) * 24 * N)) != hipSuccess) return bail(e, "flux");
	/* before the first flux pass every face's Max_v is -1, so the first
	   lagged step is 0 */
	std::vector<BoxStatus> init(2 * nbox);
	for (auto& q : init) {
		for (int d = 0; d < 3; d++) {
			q.smax_bits[d] = 0;
			q.zero[d] = 1;
		}
		q.err = 0;
		q.err_face = -1;
	}
	if ((e = hipMemcpy(s->status, init.data(), sizeof(BoxStatus) * 2 * nbox, hipMemcpyHostToDevice)) != hipSuccess) {
		return bail(e, "status");
	}
	/* update_vol_B(constant pressure) over cells of type >= 0, then
	   apply_boundaries_sw_box and sync_magnetic_field
	   (tests/mhd/solar_wind_box.cpp:441-461) */
	launch_stage(s, 0, nullptr, 3, s->status);
	launch_stage(s, 0, nullptr, 2, s->status);
	if ((e = hipGetLastError()) != hipSuccess) return bail(e, "init launch");
	if ((e = hipDeviceSynchronize()) != hipSuccess) return bail(e, "init");
	*out = s;
	return PAMHD_OK;
}

namespace {

/* dt, flux and the update stages of one step (chain: stage 4's flags) */
void launch_one_step(BoxSet* s, int solver, double max_dt, double cfl, int last_stage, int chain,
	hipStream_t st)
{
	BoxStatus* prev = s->status + (long long)s->last * s->nbox;
	BoxStatus* next = s->status + (long long)(1 - s->last) * s->nbox;
	box_dt_kernel<<<(s->nbox + 63) / 64, 64, 0, st>>>(s->boxes, s->nbox, prev, next, s->dt, cfl, max_dt);
	launch_stage(s, solver, st, 0, next);
	launch_stage(s, solver, st, last_stage, next, chain);
	s->last = 1 - s->last;
}

bool valid_solver(int solver, std::string& err)
{
	if (solver == RUSANOV || solver == HLL_ATHENA || solver == HLLD_ATHENA || solver == ROE_ATHENA) return true;
	err = "invalid solver " + std::to_string(solver);
	return false;
}

} // namespace

int boxset_launch_steps(BoxSet* s, int solver, double cfl, int steps, hipStream_t st, std::string& err)
{
	if (!valid_solver(solver, err)) return PAMHD_E_SOLVER;
	for (int q = 0; q < steps; q++)
		launch_one_step(s, solver, 1e300, cfl, 4, (q > 0 ? CH_FROM2 : 0) | (q + 1 < steps ? CH_KEEP2 : 0), st);
	const hipError_t e = hipGetLastError();
	if (e != hipSuccess) {
		err = std::string("launch: ") + hipGetErrorString(e);
		return PAMHD_E_DEVICE;
	}
	return PAMHD_OK;
}

int boxset_launch_step(BoxSet* s, int solver, double max_dt, double cfl, bool with_boundaries,
	hipStream_t st, std::string& err)
{
	if (!valid_solver(solver, err)) return PAMHD_E_SOLVER;
	launch_one_step(s, solver, max_dt, cfl, with_boundaries ? 4 : 1, 0, st);
	const hipError_t e = hipGetLastError();
	if (e != hipSuccess) {
		err = std::string("launch: ") + hipGetErrorString(e);
		return PAMHD_E_DEVICE;
	}
	return PAMHD_OK;
}

template <int SOLVER, int TX, int TY>
int run_persistent(BoxSet* s, double cfl, long steps, hipStream_t st, std::string& err)
{
	const void* kern = reinterpret_cast<const void*>(&box_run_kernel<SOLVER, TX, TY>);
	int per_cu = 0, dev = 0, ncu = 0;
	hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, box_run_kernel<SOLVER, TX, TY>, TX * TY, 0);
	if (e == hipSuccess) e = hipGetDevice(&dev);
	if (e == hipSuccess) e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
	if (e != hipSuccess || per_cu < 1 || ncu < 1) {
		err = std::string("persistent run: occupancy query failed: ") + hipGetErrorString(e);
		return PAMHD_E_DEVICE;
	}
	/* every workgroup resident at once (the cooperative launch checks it) */
	const long long resident = (long long)per_cu * ncu;
	const unsigned grid = (unsigned)(s->nblocks < resident ? s->nblocks : resident);
	RunArgs a;
	a.boxes = s->boxes;
	a.blockbox = s->blockbox;
	a.N = s->N;
	a.S = s->S;
	a.ctype = s->ctype;
	a.rule = s->rule;
	a.bnd = s->bnd;
	a.inflow = s->inflow;
	a.nbnd = s->nbnd;
	a.ninflow = s->ninflow;
	a.bg = s->bg;
	a.P = s->P;
	a.dt = s->dt;
	a.status = s->status;
	a.nbox = s->nbox;
	a.nblocks = s->nblocks;
	a.ph = s->ph;
	a.cfl = cfl;
	int last = s->last;
	unsigned* bar = s->bar;
	unsigned polls = s->bar_polls ? s->bar_polls : RUN_SPIN_LIMIT;
	void* args[] = {&a, &steps, &last, &bar, &polls};
	run_barrier_clear_kernel<<<1, 64, 0, st>>>(s->bar);
	e = hipLaunchCooperativeKernel(kern, dim3(grid), dim3(TX, TY, 1), args, 0, st);
	if (e != hipSuccess) {
		err = std::string("persistent run: cooperative launch: ") + hipGetErrorString(e);
		return PAMHD_E_DEVICE;
	}
	s->last = (int)((s->last + steps) & 1);
	unsigned flag[2] = {0, 0};
	if ((e = hipMemcpyAsync(flag, s->bar, sizeof(flag), hipMemcpyDeviceToHost, st)) != hipSuccess ||
		(e = hipStreamSynchronize(st)) != hipSuccess) {
		err = std::string("persistent run: ") + hipGetErrorString(e);
		return PAMHD_E_DEVICE;
	}
	if (flag[1]) {
		err = "persistent run: a grid barrier timed out (not every workgroup arrived); the state is incomplete";
		return PAMHD_E_DEVICE;
	}
	return PAMHD_OK;
}

template <int SOLVER>
int run_persistent_tile(BoxSet* s, double cfl, long steps, hipStream_t st, std::string& err)
{
	switch (s->tile) {
	case 0: return run_persistent<SOLVER, 64, 4>(s, cfl, steps, st, err);
	case 1: return run_persistent<SOLVER, 32, 8>(s, cfl, steps, st, err);
	default: return run_persistent<SOLVER, 16, 16>(s, cfl, steps, st, err);
	}
}

int boxset_run_persistent(BoxSet* s, int solver, double cfl, long steps, hipStream_t st, std::string& err)
{
	if (steps <= 0) return PAMHD_OK;
	switch (solver) {
	case RUSANOV: return run_persistent_tile<RUSANOV>(s, cfl, steps, st, err);
	case HLL_ATHENA: return run_persistent_tile<HLL_ATHENA>(s, cfl, steps, st, err);
	case HLLD_ATHENA: return run_persistent_tile<HLLD_ATHENA>(s, cfl, steps, st, err);
	case ROE_ATHENA: return run_persistent_tile<ROE_ATHENA>(s, cfl, steps, st, err);
	default:
		err = "invalid solver " + std::to_string(solver);
		return PAMHD_E_SOLVER;
	}
}

void boxset_set_barrier_polls(BoxSet* s, unsigned polls)
{
	s->bar_polls = polls;
}

int boxset_launch_boundaries(BoxSet* s, hipStream_t st)
{
	launch_stage(s, 0, st, 2, s->status + (long long)s->last * s->nbox);
	return hipGetLastError() == hipSuccess ? PAMHD_OK : PAMHD_E_DEVICE;
}

int boxset_readback(BoxSet* s, hipStream_t st, double* dt_per_box, std::string& err)
{
	hipError_t e;
	const BoxStatus* cur = s->status + (long long)s->last * s->nbox;
	if ((e = hipMemcpyAsync(s->host_status, cur, sizeof(BoxStatus) * s->nbox, hipMemcpyDeviceToHost, st)) != hipSuccess ||
		(e = hipMemcpyAsync(s->host_dt, s->dt, sizeof(double) * s->nbox, hipMemcpyDeviceToHost, st)) != hipSuccess ||
		(e = hipStreamSynchronize(st)) != hipSuccess) {
		err = std::string("readback: ") + hipGetErrorString(e);
		return PAMHD_E_DEVICE;
	}
	if (dt_per_box) for (int b = 0; b < s->nbox; b++) dt_per_box[b] = s->host_dt[b];
	for (int b = 0; b < s->nbox; b++) {
		const BoxStatus& q = s->host_status[b];
		if (!q.err) continue;
		const BoxDev& B = s->host[b];
		const long long d = q.err_face / B.n, c = q.err_face % B.n;
		const int i = (int)(c % B.nx), j = (int)((c / B.nx) % B.ny), k = (int)(c / ((long long)B.nx * B.ny));
		char buf[256];
		if (d < 3) {
			snprintf(buf, sizeof(buf), "box %d: solution failed between cell (%d,%d,%d) and its -%c neighbour (error %d)",
				b, i, j, k, "xyz"[d], q.err);
		} else {
			snprintf(buf, sizeof(buf), "box %d: boundary cell (%d,%d,%d) copies from a cell with non-positive density",
				b, i, j, k);
		}
		err = buf;
		return q.err;
	}
	return PAMHD_OK;
}

hipError_t boxset_download(BoxSet* s, int box, double* state) { return copy_box(s, box, state, true); }
hipError_t boxset_upload(BoxSet* s, int box, const double* state) { return copy_box(s, box, const_cast<double*>(state), false); }

hipError_t boxset_cell_types(BoxSet* s, int box, signed char* ctype, unsigned char* rule)
{
	const BoxDev& B = s->host[box];
	hipError_t e = hipSuccess;
	if (ctype) e = hipMemcpy(ctype, s->ctype + B.off, B.n, hipMemcpyDeviceToHost);
	if (e == hipSuccess && rule) e = hipMemcpy(rule, s->rule + B.off, B.n, hipMemcpyDeviceToHost);
	return e;
}

hipError_t boxset_background(BoxSet* s, int box, double* bg)
{
	const BoxDev& B = s->host[box];
	for (int q = 0; q < 9; q++) {
		const hipError_t e = hipMemcpy(bg + (long long)q * B.n, s->bg + (long long)q * s->N + B.off,
			sizeof(double) * B.n, hipMemcpyDeviceToHost);
		if (e != hipSuccess) return e;
	}
	return hipSuccess;
}

int boxset_autotune(BoxSet* s, int solver, double cfl, int steps, hipStream_t st,
	std::string& report, std::string& err)
{
	double* saveS = nullptr;
	BoxStatus* saveSt = nullptr;
	hipError_t e;
	const size_t nS = sizeof(double) * NQ * s->N, nSt = sizeof(BoxStatus) * 2 * s->nbox;
	if ((e = hipMalloc(&saveS, nS)) != hipSuccess || (e = hipMalloc(&saveSt, nSt)) != hipSuccess) {
		(void)hipFree(saveS);
		err = std::string("autotune buffers: ") + hipGetErrorString(e);
		return PAMHD_E_DEVICE;
	}
	hipEvent_t a, b;
	(void)hipEventCreate(&a);
	(void)hipEventCreate(&b);
	const int last0 = s->last;
	(void)hipMemcpyAsync(saveS, s->S, nS, hipMemcpyDeviceToDevice, st);
	(void)hipMemcpyAsync(saveSt, s->status, nSt, hipMemcpyDeviceToDevice, st);
	int best = s->tile;
	float best_ms = 1e30f;
	int rc = PAMHD_OK;
	report += "tiles:";
	for (int t = 0; t < 3 && rc == PAMHD_OK; t++) {
		if (boxset_set_tile(s, t) != PAMHD_OK) {
			rc = PAMHD_E_DEVICE;
			break;
		}
		rc = boxset_launch_step(s, solver, 1e300, cfl, true, st, err);
		(void)hipEventRecord(a, st);
		for (int q = 0; q < steps && rc == PAMHD_OK; q++) rc = boxset_launch_step(s, solver, 1e300, cfl, true, st, err);
		(void)hipEventRecord(b, st);
		(void)hipEventSynchronize(b);
		float ms = 0;
		(void)hipEventElapsedTime(&ms, a, b);
		ms /= steps;
		report += " " + std::to_string(TILES[t].tx) + "x" + std::to_string(TILES[t].ty) + "=" + std::to_string(ms) + "ms";
		if (ms < best_ms) {
			best_ms = ms;
			best = t;
		}
		(void)hipMemcpyAsync(s->S, saveS, nS, hipMemcpyDeviceToDevice, st);
		(void)hipMemcpyAsync(s->status, saveSt, nSt, hipMemcpyDeviceToDevice, st);
		s->last = last0;
	}
	(void)hipStreamSynchronize(st);
	if (rc == PAMHD_OK) rc = boxset_set_tile(s, best);
	report += "; chosen " + std::to_string(TILES[best].tx) + "x" + std::to_string(TILES[best].ty);
	(void)hipEventDestroy(a);
	(void)hipEventDestroy(b);
	(void)hipFree(saveS);
	(void)hipFree(saveSt);
	return rc;
}

hipError_t boxset_div_face2volume(BoxSet* s, int box, double* avg)
{
	const dim3 grid(s->nblocks);
	switch (s->tile) {
	case 0: box_div_kernel<64, 4><<<grid, dim3(64, 4, 1)>>>(s->boxes, s->blockbox, s->N, s->S, s->ctype, box, s->partial); break;
	case 1: box_div_kernel<32, 8><<<grid, dim3(32, 8, 1)>>>(s->boxes, s->blockbox, s->N, s->S, s->ctype, box, s->partial); break;
	default: box_div_kernel<16, 16><<<grid, dim3(16, 16, 1)>>>(s->boxes, s->blockbox, s->N, s->S, s->ctype, box, s->partial); break;
	}
	hipError_t e = hipGetLastError();
	if (e != hipSuccess) return e;
	std::vector<double> h(s->nblocks);
	if ((e = hipMemcpy(h.data(), s->partial, sizeof(double) * s->nblocks, hipMemcpyDeviceToHost)) != hipSuccess) return e;
	double t = 0;
	for (double v : h) t += v;
	*avg = t / (double)s->host[box].n;
	return hipSuccess;
}

} // namespace pamhd_dev
