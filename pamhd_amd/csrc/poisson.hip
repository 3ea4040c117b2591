/*
Multigrid Poisson solver for divergence cleaning on MI355X (gfx950).

Every level keeps phi, f and r as FP64 arrays, x fastest, or, on levels
smoothed red-black with an even x size, with the two colours as separate
halves (mg_index): a colour pass then reads the other colour and writes its
own in whole cache lines.  Kernels:
  rb_kernel        one colour of a red-black Gauss-Seidel sweep; a thread
                   owns one cell of the colour (x index 2t or 2t+1)
  jacobi_kernel    damped Jacobi into r (phi and r are then swapped)
  residual_kernel  r = f - L phi
  restrict_kernel  coarse f = mean of the children's r, coarse phi = 0
  restrict_residual_kernel  both of these at once, for pairs coarsened by 2
                   in every dimension (no r array)
  prolong_kernel   fine phi += trilinear interpolation of coarse phi
  prolong8_kernel  the same, a thread per coarse cell and its 8 children
  coarse_kernel    all sweeps of a small level inside one block, in LDS
  norm kernels     one-pass centred residual norm of the stopping test
  sum kernels      two-stage deterministic sums
Each value is computed in the same operation order as the CPU restatement,
so a given number of V-cycles gives bit-identical phi.
*/
#include <hip/hip_runtime.h>

#include <cmath>

#include "poisson.hpp"

namespace pamhd_dev {

namespace {

constexpr int PBX = 64, PBY = 4;
constexpr int COARSE_MAX = 2048;   /* cells of a level solved in one block */
constexpr int SUM_BLOCK = 256;

struct LevelArgs {
	int n0, n1, n2;
	double hh0, hh1, hh2;
	double diag;
	int packed, half;   /* colour-split layout, ncell / 2 */
};

LevelArgs args_of(const MgLevel& L)
{
	return LevelArgs{L.n[0], L.n[1], L.n[2], L.hh[0], L.hh[1], L.hh[2], L.diag, L.packed,
		L.ncell / 2};
}

/* the PBX x PBY thread tiles over (x, y) of each plane (or of each chunk
   of planes), launched as a 1-D grid.  Workgroups are handed to the 8 XCDs
   in turn (workgroup h to XCD h % 8), each XCD with its own L2: when the y
   tile count is a multiple of 8, XCD x takes the y range
   [x ty / 8, (x + 1) ty / 8) of every plane, so that the y and z neighbour
   rows a tile reads were mostly fetched by tiles on the same XCD. */
struct Tiles {
	int tx, ty, tz, swz;
};

Tiles tiles_of(int nx, int ny, int nz)
{
	Tiles T{(nx + PBX - 1) / PBX, (ny + PBY - 1) / PBY, nz, 0};
	T.swz = T.ty % 8 == 0 ? 1 : 0;
	return T;
}

int blocks_of(const Tiles& T)
{
	return T.tx * T.ty * T.tz;
}

__device__ __forceinline__ void tile_of(const Tiles& T, int& bx, int& by, int& bz)
{
	const int h = blockIdx.x;
	if (T.swz) {
		const int xcd = h & 7, r = h >> 3;
		const int per = T.tx * (T.ty >> 3);   /* one XCD's tiles of a plane */
		bz = r / per;
		const int l = r - bz * per;
		by = xcd * (T.ty >> 3) + l / T.tx;
		bx = l - (l / T.tx) * T.tx;
	} else {
		bx = h % T.tx;
		by = (h / T.tx) % T.ty;
		bz = h / (T.tx * T.ty);
	}
}

__device__ __forceinline__ int wrapi(int i, int n)
{
	return i < 0 ? i + n : (i >= n ? i - n : i);
}

/* in-range cell's index in the level's layout */
__device__ __forceinline__ int cidx(const LevelArgs& L, int i, int j, int k)
{
	return mg_index(L.packed, L.n0, L.n1, L.half, i, j, k);
}

/* periodic neighbour's index */
__device__ __forceinline__ int lidx(const LevelArgs& L, int i, int j, int k)
{
	return cidx(L, wrapi(i, L.n0), wrapi(j, L.n1), wrapi(k, L.n2));
}

/* indices of an in-range cell and its six periodic neighbours:
   [0] the cell, [1] x+1, [2] x-1, [3] y+1, [4] y-1, [5] z+1, [6] z-1 (the
   neighbours of a packed level's cell are all in the other colour's half) */
__device__ __forceinline__ void nb7(const LevelArgs& L, int i, int j, int k, int x[7])
{
	const int ip = i + 1 == L.n0 ? 0 : i + 1, im = i == 0 ? L.n0 - 1 : i - 1;
	const int jp = j + 1 == L.n1 ? 0 : j + 1, jm = j == 0 ? L.n1 - 1 : j - 1;
	const int kp = k + 1 == L.n2 ? 0 : k + 1, km = k == 0 ? L.n2 - 1 : k - 1;
	if (L.packed) {
		const int h0 = L.n0 >> 1;
		const int own = ((i + j + k) & 1) * L.half, other = L.half - own;
		const int row = h0 * (j + L.n1 * k);
		const int ih = i >> 1;
		x[0] = own + ih + row;
		x[1] = other + (ip >> 1) + row;
		x[2] = other + (im >> 1) + row;
		x[3] = other + ih + h0 * (jp + L.n1 * k);
		x[4] = other + ih + h0 * (jm + L.n1 * k);
		x[5] = other + ih + h0 * (j + L.n1 * kp);
		x[6] = other + ih + h0 * (j + L.n1 * km);
	} else {
		const int row = L.n0 * (j + L.n1 * k);
		x[0] = i + row;
		x[1] = ip + row;
		x[2] = im + row;
		x[3] = i + L.n0 * (jp + L.n1 * k);
		x[4] = i + L.n0 * (jm + L.n1 * k);
		x[5] = i + L.n0 * (j + L.n1 * kp);
		x[6] = i + L.n0 * (j + L.n1 * km);
	}
}

/* (sum of neighbours / h^2 - f) / diag, or phi itself without active
   dimensions */
__device__ __forceinline__ double relax_value(const LevelArgs& L, const double* phi,
	const double* f, int i, int j, int k)
{
	int x[7];
	nb7(L, i, j, k, x);
	double s = 0;
	if (L.n0 > 1) s += (phi[x[1]] + phi[x[2]]) / L.hh0;
	if (L.n1 > 1) s += (phi[x[3]] + phi[x[4]]) / L.hh1;
	if (L.n2 > 1) s += (phi[x[5]] + phi[x[6]]) / L.hh2;
	if (L.diag == 0) return phi[x[0]];
	return (s - f[x[0]]) / L.diag;
}

__device__ __forceinline__ double apply_lap(const LevelArgs& L, const double* phi, int i, int j, int k)
{
	int x[7];
	nb7(L, i, j, k, x);
	const double c = phi[x[0]];
	double s = 0;
	if (L.n0 > 1) s += (phi[x[1]] - 2 * c + phi[x[2]]) / L.hh0;
	if (L.n1 > 1) s += (phi[x[3]] - 2 * c + phi[x[4]]) / L.hh1;
	if (L.n2 > 1) s += (phi[x[5]] - 2 * c + phi[x[6]]) / L.hh2;
	return s;
}

__global__ __launch_bounds__(PBX * PBY) void rb_kernel(Tiles T, LevelArgs L, double* __restrict__ phi,
	const double* __restrict__ f, int color)
{
	int bx, by, k;
	tile_of(T, bx, by, k);
	const int t = bx * PBX + threadIdx.x;
	const int j = by * PBY + threadIdx.y;
	if (j >= L.n1) return;
	const int i = 2 * t + ((j + k + color) & 1);
	if (i >= L.n0) return;
	phi[cidx(L, i, j, k)] = relax_value(L, phi, f, i, j, k);
}

__global__ __launch_bounds__(PBX * PBY) void jacobi_kernel(Tiles T, LevelArgs L, const double* __restrict__ phi,
	const double* __restrict__ f, double* __restrict__ out)
{
	int bx, by, k;
	tile_of(T, bx, by, k);
	const int i = bx * PBX + threadIdx.x;
	const int j = by * PBY + threadIdx.y;
	if (i >= L.n0 || j >= L.n1) return;
	const int c = cidx(L, i, j, k);
	out[c] = phi[c] + 0.8 * (relax_value(L, phi, f, i, j, k) - phi[c]);
}

__global__ __launch_bounds__(PBX * PBY) void residual_kernel(Tiles T, LevelArgs L, const double* __restrict__ phi,
	const double* __restrict__ f, double* __restrict__ r)
{
	int bx, by, k;
	tile_of(T, bx, by, k);
	const int i = bx * PBX + threadIdx.x;
	const int j = by * PBY + threadIdx.y;
	if (i >= L.n0 || j >= L.n1) return;
	const int c = cidx(L, i, j, k);
	r[c] = f[c] - apply_lap(L, phi, i, j, k);
}

/* the coarsening ratios (1 or 2) of a level pair, and 1 / (r0 r1 r2) */
struct Ratio {
	int r0, r1, r2;
	double w;
};

Ratio ratio_of(const MgLevel& F, const MgLevel& C)
{
	Ratio q{F.n[0] / C.n[0], F.n[1] / C.n[1], F.n[2] / C.n[2], 0.0};
	q.w = 1.0 / (double)(q.r0 * q.r1 * q.r2);
	return q;
}

/* the sum of the children's r of coarse cell (i, j, k) */
__device__ __forceinline__ double restrict_sum(const LevelArgs& F, const Ratio& q, const double* r,
	int i, int j, int k)
{
	const int r0 = q.r0, r1 = q.r1, r2 = q.r2;
	double s = 0;
	for (int c = 0; c < r2; c++)
	for (int b = 0; b < r1; b++)
	for (int a = 0; a < r0; a++) {
		s += r[cidx(F, i * r0 + a, j * r1 + b, k * r2 + c)];
	}
	return s;
}

__global__ __launch_bounds__(PBX * PBY) void restrict_kernel(Tiles T, LevelArgs F, LevelArgs C, Ratio q,
	const double* __restrict__ r, double* __restrict__ cf, double* __restrict__ cphi)
{
	int bx, by, k;
	tile_of(T, bx, by, k);
	const int i = bx * PBX + threadIdx.x;
	const int j = by * PBY + threadIdx.y;
	if (i >= C.n0 || j >= C.n1) return;
	const int o = cidx(C, i, j, k);
	cf[o] = restrict_sum(F, q, r, i, j, k) * q.w;
	cphi[o] = 0;
}

/* restriction of the residual for level pairs coarsened by 2 in every
   dimension, without a residual pass: a thread computes the residuals
   f - L phi of the 8 children of its coarse cell (the operations of
   apply_lap) and sums them in the order of restrict_kernel; coarse phi = 0 */
__global__ __launch_bounds__(PBX * PBY) void restrict_residual_kernel(Tiles T, LevelArgs F, LevelArgs C,
	double w, const double* __restrict__ phi, const double* __restrict__ f,
	double* __restrict__ cf, double* __restrict__ cphi)
{
	int bx, by, k;
	tile_of(T, bx, by, k);
	const int i = bx * PBX + threadIdx.x;
	const int j = by * PBY + threadIdx.y;
	if (i >= C.n0 || j >= C.n1) return;
	/* fine coordinates 2i - 1 .. 2i + 2 per dimension, wrapped */
	int X[4], Y[4], Z[4];
	for (int d = 0; d < 4; d++) {
		X[d] = wrapi(2 * i - 1 + d, F.n0);
		Y[d] = wrapi(2 * j - 1 + d, F.n1);
		Z[d] = wrapi(2 * k - 1 + d, F.n2);
	}
	auto P = [&](int a, int b, int c) { return phi[cidx(F, X[a], Y[b], Z[c])]; };
	double s = 0;
	for (int c = 1; c <= 2; c++)
	for (int b = 1; b <= 2; b++)
	for (int a = 1; a <= 2; a++) {
		const double ctr = P(a, b, c);
		double l = 0;
		l += (P(a + 1, b, c) - 2 * ctr + P(a - 1, b, c)) / F.hh0;
		l += (P(a, b + 1, c) - 2 * ctr + P(a, b - 1, c)) / F.hh1;
		l += (P(a, b, c + 1) - 2 * ctr + P(a, b, c - 1)) / F.hh2;
		s += f[cidx(F, X[a], Y[b], Z[c])] - l;
	}
	const int o = cidx(C, i, j, k);
	cf[o] = s * w;
	cphi[o] = 0;
}

/* prolongation for level pairs coarsened by 2 in every dimension: a thread
   adds the interpolation to the 8 children of its coarse cell from the 27
   coarse values around it, each child's terms in the order of
   prolong_kernel */
__global__ __launch_bounds__(PBX * PBY) void prolong8_kernel(Tiles T, LevelArgs F, LevelArgs C,
	double* __restrict__ phi, const double* __restrict__ cphi)
{
	int bx, by, k;
	tile_of(T, bx, by, k);
	const int i = bx * PBX + threadIdx.x;
	const int j = by * PBY + threadIdx.y;
	if (i >= C.n0 || j >= C.n1) return;
	int X[3], Y[3], Z[3];   /* coarse i - 1, i, i + 1, wrapped */
	for (int d = 0; d < 3; d++) {
		X[d] = wrapi(i - 1 + d, C.n0);
		Y[d] = wrapi(j - 1 + d, C.n1);
		Z[d] = wrapi(k - 1 + d, C.n2);
	}
	double q[3][3][3];
	for (int c = 0; c < 3; c++)
	for (int b = 0; b < 3; b++)
	for (int a = 0; a < 3; a++) q[c][b][a] = cphi[cidx(C, X[a], Y[b], Z[c])];
	const double w[2] = {0.75, 0.25};
	for (int c = 0; c < 2; c++)
	for (int b = 0; b < 2; b++)
	for (int a = 0; a < 2; a++) {
		/* the parent (index 1) and its neighbour on the child's side */
		const int pa[2] = {1, a ? 2 : 0}, pb[2] = {1, b ? 2 : 0}, pc[2] = {1, c ? 2 : 0};
		double s = 0;
		for (int zc = 0; zc < 2; zc++)
		for (int yb = 0; yb < 2; yb++)
		for (int xa = 0; xa < 2; xa++) {
			s += w[xa] * w[yb] * w[zc] * q[pc[zc]][pb[yb]][pa[xa]];
		}
		phi[cidx(F, 2 * i + a, 2 * j + b, 2 * k + c)] += s;
	}
}

/* the trilinear interpolation of the coarse phi at fine cell (i, j, k) */
__device__ __forceinline__ double prolong_sum(const LevelArgs& C, const Ratio& q, const double* cphi,
	int i, int j, int k)
{
	const int idx[3] = {i, j, k};
	const int ratio[3] = {q.r0, q.r1, q.r2};
	const int cn[3] = {C.n0, C.n1, C.n2};
	/* per dimension the coarse cells (wrapped) and weights: 3/4 of the
	   parent and 1/4 of its neighbour on the child's side when the
	   dimension was coarsened, the parent alone otherwise */
	int ci[3][2];
	double cw[3][2];
	int cnt[3];
	for (int d = 0; d < 3; d++) {
		if (ratio[d] == 2) {
			const int I = idx[d] >> 1;
			const int J = (idx[d] & 1) ? (I + 1 == cn[d] ? 0 : I + 1) : (I == 0 ? cn[d] - 1 : I - 1);
			ci[d][0] = I;
			cw[d][0] = 0.75;
			ci[d][1] = J;
			cw[d][1] = 0.25;
			cnt[d] = 2;
		} else {
			ci[d][0] = idx[d];
			cw[d][0] = 1.0;
			ci[d][1] = idx[d];
			cw[d][1] = 0.0;
			cnt[d] = 1;
		}
	}
	double s = 0;
	for (int c = 0; c < cnt[2]; c++)
	for (int b = 0; b < cnt[1]; b++)
	for (int a = 0; a < cnt[0]; a++) {
		s += cw[0][a] * cw[1][b] * cw[2][c] * cphi[cidx(C, ci[0][a], ci[1][b], ci[2][c])];
	}
	return s;
}

__global__ __launch_bounds__(PBX * PBY) void prolong_kernel(Tiles T, LevelArgs F, LevelArgs C, Ratio q,
	double* __restrict__ phi, const double* __restrict__ cphi)
{
	int bx, by, k;
	tile_of(T, bx, by, k);
	const int i = bx * PBX + threadIdx.x;
	const int j = by * PBY + threadIdx.y;
	if (i >= F.n0 || j >= F.n1) return;
	phi[cidx(F, i, j, k)] += prolong_sum(C, q, cphi, i, j, k);
}

/* sweeps of a level held in LDS by one block (phi, f and a Jacobi buffer
   tmp in the level's layout), n <= PER * blockDim.x cells: a thread owns
   cells tid + m blockDim.x (m < PER) and keeps their neighbour indices,
   colours and right-hand sides in registers across the sweeps.  Ends with
   a barrier. */
template <int PER>
__device__ void block_sweeps(const LevelArgs& L, double* phi, const double* f, double* tmp, int n,
	int sweeps, int red_black)
{
	int x[PER][7], color[PER];
	double fc[PER];
	for (int m = 0; m < PER; m++) {
		const int c = threadIdx.x + m * blockDim.x;
		color[m] = -1;
		fc[m] = 0;
		if (c >= n) continue;
		const int i = c % L.n0, j = (c / L.n0) % L.n1, k = c / (L.n0 * L.n1);
		nb7(L, i, j, k, x[m]);
		color[m] = (i + j + k) & 1;
		fc[m] = f[x[m][0]];
	}
	/* relax_value on the cell's stored indices */
	auto relax = [&](int m) {
		double s = 0;
		if (L.n0 > 1) s += (phi[x[m][1]] + phi[x[m][2]]) / L.hh0;
		if (L.n1 > 1) s += (phi[x[m][3]] + phi[x[m][4]]) / L.hh1;
		if (L.n2 > 1) s += (phi[x[m][5]] + phi[x[m][6]]) / L.hh2;
		if (L.diag == 0) return phi[x[m][0]];
		return (s - fc[m]) / L.diag;
	};
	__syncthreads();
	for (int s = 0; s < sweeps; s++) {
		if (red_black) {
			for (int col = 0; col < 2; col++) {
				for (int m = 0; m < PER; m++) {
					if (color[m] == col) phi[x[m][0]] = relax(m);
				}
				__syncthreads();
			}
		} else {
			for (int m = 0; m < PER; m++) {
				if (color[m] < 0) continue;
				const double p = phi[x[m][0]];
				tmp[x[m][0]] = p + 0.8 * (relax(m) - p);
			}
			__syncthreads();
			for (int c = threadIdx.x; c < n; c += blockDim.x) phi[c] = tmp[c];
			__syncthreads();
		}
	}
}

/* sweeps of a small level (<= COARSE_MAX cells) in one block, in LDS */
constexpr int COARSE_PER = COARSE_MAX / 1024;

__global__ __launch_bounds__(1024) void coarse_kernel(LevelArgs L, double* __restrict__ phi_g,
	const double* __restrict__ f_g, int sweeps, int red_black)
{
	__shared__ double phi[COARSE_MAX], f[COARSE_MAX], tmp[COARSE_MAX];
	const int n = L.n0 * L.n1 * L.n2;
	for (int c = threadIdx.x; c < n; c += blockDim.x) {
		phi[c] = phi_g[c];
		f[c] = f_g[c];
	}
	__syncthreads();
	block_sweeps<COARSE_PER>(L, phi, f, tmp, n, sweeps, red_black);
	for (int c = threadIdx.x; c < n; c += blockDim.x) phi_g[c] = phi[c];
}

/*
The small levels at the bottom of the hierarchy in one block: from the
first tail level (restricted f and zero phi in global memory) a whole
V-cycle (pre-smoothing, residual, restriction, ..., the coarsest level's
sweeps, ..., prolongation, post-smoothing) runs in LDS with a barrier
between stages, each stage computing every value as the level kernels do,
and the first tail level's phi goes back to global memory.  Levels of up to
TAIL_PER * 1024 cells, TAIL_CELLS in all.
*/
constexpr int TAIL_LEVELS = 8, TAIL_CELLS = 5120, TAIL_PER = 4;

struct TailArgs {
	LevelArgs L[TAIL_LEVELS];
	Ratio q[TAIL_LEVELS];    /* level l to l + 1 */
	int off[TAIL_LEVELS];    /* the level's offset in each LDS array */
	int n[TAIL_LEVELS];      /* cells */
	int rb[TAIL_LEVELS];     /* red-black (else damped Jacobi) */
	int nlev;
};

__device__ __forceinline__ void cell_of(const LevelArgs& L, int c, int& i, int& j, int& k)
{
	i = c % L.n0;
	j = (c / L.n0) % L.n1;
	k = c / (L.n0 * L.n1);
}

__global__ __launch_bounds__(1024) void tail_kernel(TailArgs A, double* __restrict__ phi_g,
	const double* __restrict__ f_g)
{
	__shared__ double pool[3 * TAIL_CELLS];
	double* const P = pool;                 /* phi */
	double* const Fv = pool + TAIL_CELLS;   /* f */
	double* const R = pool + 2 * TAIL_CELLS;   /* residual, Jacobi buffer */
	for (int c = threadIdx.x; c < A.n[0]; c += blockDim.x) {
		P[c] = phi_g[c];
		Fv[c] = f_g[c];
	}
	__syncthreads();
	const int last = A.nlev - 1;
	for (int l = 0; l < last; l++) {
		const LevelArgs L = A.L[l], C = A.L[l + 1];
		const int o = A.off[l], oc = A.off[l + 1];
		block_sweeps<TAIL_PER>(L, P + o, Fv + o, R + o, A.n[l], 2, A.rb[l]);
		for (int c = threadIdx.x; c < A.n[l]; c += blockDim.x) {
			int i, j, k;
			cell_of(L, c, i, j, k);
			const int x = cidx(L, i, j, k);
			R[o + x] = Fv[o + x] - apply_lap(L, P + o, i, j, k);
		}
		__syncthreads();
		for (int c = threadIdx.x; c < A.n[l + 1]; c += blockDim.x) {
			int i, j, k;
			cell_of(C, c, i, j, k);
			const int x = cidx(C, i, j, k);
			Fv[oc + x] = restrict_sum(L, A.q[l], R + o, i, j, k) * A.q[l].w;
			P[oc + x] = 0;
		}
		__syncthreads();
	}
	block_sweeps<TAIL_PER>(A.L[last], P + A.off[last], Fv + A.off[last], R + A.off[last],
		A.n[last], 60, A.rb[last]);
	for (int l = last - 1; l >= 0; l--) {
		const LevelArgs L = A.L[l], C = A.L[l + 1];
		const int o = A.off[l], oc = A.off[l + 1];
		for (int c = threadIdx.x; c < A.n[l]; c += blockDim.x) {
			int i, j, k;
			cell_of(L, c, i, j, k);
			P[o + cidx(L, i, j, k)] += prolong_sum(C, A.q[l], P + oc, i, j, k);
		}
		__syncthreads();
		block_sweeps<TAIL_PER>(L, P + o, Fv + o, R + o, A.n[l], 2, A.rb[l]);
	}
	for (int c = threadIdx.x; c < A.n[0]; c += blockDim.x) phi_g[c] = P[c];
}

/*
Fused red-black sweep, z-marching with LDS windows.  A block owns a TX x TY
column tile and a chunk of planes.  One pass does a whole Gauss-Seidel
sweep in the order of the two colour passes: a red cell's new value depends
only on old black values, a black cell's on new red values.  Per plane q the
block keeps
  O[q]  old values of the tile with a two-cell ring (read from phi_in),
  M[q]  the tile with a one-cell ring after the red update (new red, old
        black values), computed from O[q-1], O[q], O[q+1];
and the black update of plane k reads M[k-1], M[k], M[k+1].  Planes wrap
periodically; a chunk starts by recomputing the M planes below it from old
values, which gives the same numbers as the block that owns them.  phi_in is
never written (phi is a separate buffer), so blocks are independent.

Requires every dimension even and >= 4 (checked by the caller).
*/
constexpr int TX = 32, TY = 8;
constexpr int WX = TX + 4, WY = TY + 4;     /* window with a 2-cell ring */
constexpr int NO = 4, NM = 3;               /* O planes k-1..k+2, M planes k-1..k+1 */

/* v >= -n wrapped into [0, n); the modulo only for windows wider than the
   level (levels narrower than a tile) */
__device__ __forceinline__ int wrapm(int v, int n)
{
	return v < 0 ? v + n : v < n ? v : v < 2 * n ? v - n : v % n;
}

__device__ __forceinline__ double rv_window(const LevelArgs& L, const double (*w)[WY][WX],
	int p, int pm, int pp, int x, int y, double fc)
{
	double s = 0;
	s += (w[p][y][x + 1] + w[p][y][x - 1]) / L.hh0;
	s += (w[p][y + 1][x] + w[p][y - 1][x]) / L.hh1;
	s += (w[pp][y][x] + w[pm][y][x]) / L.hh2;
	return (s - fc) / L.diag;
}

__device__ __forceinline__ int pslot(int q, int n) { return ((q % n) + n) % n; }

__global__ __launch_bounds__(TX * TY) void rb_fused_kernel(LevelArgs L, double* __restrict__ phi,
	const double* __restrict__ f, const double* __restrict__ phi_in, int kchunk)
{
	constexpr int NT = TX * TY;
	constexpr int FX = TX + 2, FY = TY + 2;           /* right-hand side over the red window */
	constexpr int NWO = (WX * WY + NT - 1) / NT;       /* window values per thread */
	constexpr int NWF = (FX * FY + NT - 1) / NT;
	__shared__ double O[NO][WY][WX];
	__shared__ double M[NM][WY][WX];
	__shared__ double Fw[2][FY][FX];                   /* f of planes q, slot q & 1 */
	const int i0 = blockIdx.x * TX, j0 = blockIdx.y * TY;
	const int k0 = blockIdx.z * kchunk;
	const int k1 = k0 + kchunk < L.n2 ? k0 + kchunk : L.n2;
	const int tid = threadIdx.x + TX * threadIdx.y;
	const int nz = L.n2;
	double po[NWO], pf[NWF];

	/* global loads of a plane's window into registers, issued a plane ahead
	   of their use so that their latency overlaps the sweeps of the plane
	   before, and their stores into LDS after the next barrier */
	auto fetch_old = [&](int q) {
		const int kk = wrapm(q, nz);
#pragma unroll
		for (int r = 0; r < NWO; r++) {
			const int e = tid + r * NT;
			if (e < WX * WY) {
				const int x = e % WX, y = e / WX;
				const int gi = wrapm(i0 + x - 2, L.n0), gj = wrapm(j0 + y - 2, L.n1);
				po[r] = phi_in[cidx(L, gi, gj, kk)];
			}
		}
	};
	auto store_old = [&](int q) {
		const int s = pslot(q, NO);
#pragma unroll
		for (int r = 0; r < NWO; r++) {
			const int e = tid + r * NT;
			if (e < WX * WY) O[s][e / WX][e % WX] = po[r];
		}
	};
	auto fetch_f = [&](int q) {
		const int kk = wrapm(q, nz);
#pragma unroll
		for (int r = 0; r < NWF; r++) {
			const int e = tid + r * NT;
			if (e < FX * FY) {
				const int x = e % FX + 1, y = e / FX + 1;
				const int gi = wrapm(i0 + x - 2, L.n0), gj = wrapm(j0 + y - 2, L.n1);
				pf[r] = f[cidx(L, gi, gj, kk)];
			}
		}
	};
	auto store_f = [&](int q) {
#pragma unroll
		for (int r = 0; r < NWF; r++) {
			const int e = tid + r * NT;
			if (e < FX * FY) Fw[q & 1][e / FX][e % FX] = pf[r];
		}
	};
	auto red = [&](int q) {
		const int kk = wrapm(q, nz);
		const int p = pslot(q, NO), pm = pslot(q - 1, NO), pp = pslot(q + 1, NO);
		const int s = pslot(q, NM);
		for (int e = tid; e < FX * FY; e += NT) {
			const int x = e % FX + 1, y = e / FX + 1;
			const int gi = wrapm(i0 + x - 2, L.n0), gj = wrapm(j0 + y - 2, L.n1);
			double v = O[p][y][x];
			if (((gi + gj + kk) & 1) == 0) v = rv_window(L, O, p, pm, pp, x, y, Fw[q & 1][y - 1][x - 1]);
			M[s][y][x] = v;
		}
	};
	auto black = [&](int q) {
		const int gi = i0 + (int)threadIdx.x, gj = j0 + (int)threadIdx.y;
		if (gi >= L.n0 || gj >= L.n1) return;
		const int x = threadIdx.x + 2, y = threadIdx.y + 2;
		const int p = pslot(q, NM), pm = pslot(q - 1, NM), pp = pslot(q + 1, NM);
		const int c = cidx(L, gi, gj, q);
		double v = M[p][y][x];
		if (((gi + gj + q) & 1) == 1) v = rv_window(L, M, p, pm, pp, x, y, Fw[q & 1][y - 1][x - 1]);
		phi[c] = v;
	};

	for (int q = k0 - 2; q <= k0 + 1; q++) {
		fetch_old(q);
		store_old(q);
	}
	fetch_f(k0 - 1);
	store_f(k0 - 1);
	fetch_f(k0);
	store_f(k0);
	__syncthreads();
	red(k0 - 1);
	red(k0);
	fetch_old(k0 + 2);
	fetch_f(k0 + 1);
	for (int k = k0; k < k1; k++) {
		__syncthreads();
		/* O[k-2] is no longer needed: M[k-1] and M[k] exist; f[k+1] goes
		   into the slot of f[k-1] */
		store_old(k + 2);
		store_f(k + 1);
		if (k + 1 < k1) {
			fetch_old(k + 3);
			fetch_f(k + 2);
		}
		__syncthreads();
		/* into the slot of M[k-2], no longer needed after black(k-1) */
		red(k + 1);
		__syncthreads();
		black(k);
	}
}

__global__ __launch_bounds__(SUM_BLOCK) void sum_partial_kernel(const double* __restrict__ a, int n,
	int mode, double shift, double* __restrict__ partial)
{
	__shared__ double sh[SUM_BLOCK];
	double s = 0;
	const int per = (n + gridDim.x - 1) / gridDim.x;
	const int lo = blockIdx.x * per;
	const int hi = lo + per < n ? lo + per : n;
	for (int c = lo + threadIdx.x; c < hi; c += SUM_BLOCK) {
		const double v = mode == 1 ? (a[c] - shift) * (a[c] - shift) : mode == 2 ? fabs(a[c]) : a[c];
		s += v;
	}
	sh[threadIdx.x] = s;
	__syncthreads();
	for (int off = SUM_BLOCK / 2; off > 0; off >>= 1) {
		if ((int)threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
		__syncthreads();
	}
	if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

/*
Centred 2-norm of the residual r = f - L phi (or of f alone) without an r
array: a block of PBX x PBY threads takes NORM_KZ planes of its tile, each
thread computing the residuals of its column of cells (the operations of
apply_lap, the z neighbours held in registers as the column advances),
and writes (count, mean, sum of squared deviations from that mean) of its
cells; norm_final_kernel merges the blocks with the pairwise update of
Chan, Golub and LeVeque (1979), so one pass gives the two-pass value to
rounding.  Blocks and merges follow a fixed order: the result does not vary
between runs.
*/
constexpr int NORM_KZ = 8;

struct Moments {
	double n, mean, m2;
};

__device__ __forceinline__ Moments merge(const Moments& a, const Moments& b)
{
	if (b.n == 0) return a;
	if (a.n == 0) return b;
	const double n = a.n + b.n;
	const double d = b.mean - a.mean;
	return Moments{n, a.mean + d * (b.n / n), a.m2 + b.m2 + d * d * (a.n * b.n / n)};
}

__device__ double block_sum(double v, double* sh)
{
	sh[threadIdx.x + PBX * threadIdx.y] = v;
	__syncthreads();
	for (int off = PBX * PBY / 2; off > 0; off >>= 1) {
		const int t = threadIdx.x + PBX * threadIdx.y;
		if (t < off) sh[t] += sh[t + off];
		__syncthreads();
	}
	const double r = sh[0];
	__syncthreads();
	return r;
}

__global__ __launch_bounds__(PBX * PBY) void norm_partial_kernel(Tiles T, LevelArgs L,
	const double* __restrict__ phi, const double* __restrict__ f, int with_lap,
	double* __restrict__ part)
{
	__shared__ double sh[PBX * PBY];
	int bx, by, bz;
	tile_of(T, bx, by, bz);
	const int i = bx * PBX + threadIdx.x;
	const int j = by * PBY + threadIdx.y;
	const int k0 = bz * NORM_KZ;
	const bool in = i < L.n0 && j < L.n1;
	double r[NORM_KZ];
	int cnt = 0;
	if (in) {
		const int ip = i + 1 == L.n0 ? 0 : i + 1, im = i == 0 ? L.n0 - 1 : i - 1;
		const int jp = j + 1 == L.n1 ? 0 : j + 1, jm = j == 0 ? L.n1 - 1 : j - 1;
		double below = 0, ctr = 0;
		if (with_lap) {
			below = phi[cidx(L, i, j, k0 == 0 ? L.n2 - 1 : k0 - 1)];
			ctr = phi[cidx(L, i, j, k0)];
		}
		for (int m = 0; m < NORM_KZ; m++) {
			const int k = k0 + m;
			r[m] = 0;
			if (k >= L.n2) continue;
			cnt++;
			const double fc = f[cidx(L, i, j, k)];
			if (!with_lap) {
				r[m] = fc;
				continue;
			}
			const double above = phi[cidx(L, i, j, k + 1 == L.n2 ? 0 : k + 1)];
			double l = 0;
			if (L.n0 > 1) l += (phi[cidx(L, ip, j, k)] - 2 * ctr + phi[cidx(L, im, j, k)]) / L.hh0;
			if (L.n1 > 1) l += (phi[cidx(L, i, jp, k)] - 2 * ctr + phi[cidx(L, i, jm, k)]) / L.hh1;
			if (L.n2 > 1) l += (above - 2 * ctr + below) / L.hh2;
			r[m] = fc - l;
			below = ctr;
			ctr = above;
		}
	}
	double s = 0;
	for (int m = 0; m < NORM_KZ; m++) s += r[m];
	const double n = block_sum((double)cnt, sh);
	const double total = block_sum(s, sh);
	const double mean = n > 0 ? total / n : 0.0;
	double q = 0;
	for (int m = 0; m < NORM_KZ; m++) {
		if (m < cnt) q += (r[m] - mean) * (r[m] - mean);
	}
	const double m2 = block_sum(q, sh);
	if (threadIdx.x == 0 && threadIdx.y == 0) {
		part[3 * blockIdx.x] = n;
		part[3 * blockIdx.x + 1] = mean;
		part[3 * blockIdx.x + 2] = m2;
	}
}

/* out[0] = the merged sum of squared deviations of nb block moments */
constexpr int NORM_FINAL = 1024;

__global__ __launch_bounds__(NORM_FINAL) void norm_final_kernel(const double* __restrict__ part, int nb,
	double* __restrict__ out)
{
	__shared__ Moments sh[NORM_FINAL];
	Moments a{0, 0, 0};
	for (int b = threadIdx.x; b < nb; b += NORM_FINAL) {
		a = merge(a, Moments{part[3 * b], part[3 * b + 1], part[3 * b + 2]});
	}
	sh[threadIdx.x] = a;
	__syncthreads();
	for (int off = NORM_FINAL / 2; off > 0; off >>= 1) {
		if ((int)threadIdx.x < off) sh[threadIdx.x] = merge(sh[threadIdx.x], sh[threadIdx.x + off]);
		__syncthreads();
	}
	if (threadIdx.x == 0) out[0] = sh[0].m2;
}

__global__ __launch_bounds__(SUM_BLOCK) void sum_final_kernel(const double* __restrict__ partial, int n,
	double* __restrict__ out)
{
	__shared__ double sh[SUM_BLOCK];
	double s = 0;
	for (int c = threadIdx.x; c < n; c += SUM_BLOCK) s += partial[c];
	sh[threadIdx.x] = s;
	__syncthreads();
	for (int off = SUM_BLOCK / 2; off > 0; off >>= 1) {
		if ((int)threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
		__syncthreads();
	}
	if (threadIdx.x == 0) *out = sh[0];
}

/* out[0], out[1] = the sums of the even and odd entries of pairs[0 .. 2 nb) */
__global__ __launch_bounds__(SUM_BLOCK) void pair_final_kernel(const double* __restrict__ pairs, int nb,
	double* __restrict__ out)
{
	__shared__ double sh[2][SUM_BLOCK];
	double a = 0, b = 0;
	for (int c = threadIdx.x; c < nb; c += SUM_BLOCK) {
		a += pairs[2 * c];
		b += pairs[2 * c + 1];
	}
	sh[0][threadIdx.x] = a;
	sh[1][threadIdx.x] = b;
	__syncthreads();
	for (int off = SUM_BLOCK / 2; off > 0; off >>= 1) {
		if ((int)threadIdx.x < off) {
			sh[0][threadIdx.x] += sh[0][threadIdx.x + off];
			sh[1][threadIdx.x] += sh[1][threadIdx.x + off];
		}
		__syncthreads();
	}
	if (threadIdx.x == 0) {
		out[0] = sh[0][0];
		out[1] = sh[1][0];
	}
}

Tiles level_tiles(const MgLevel& L, int xdiv = 1)
{
	return tiles_of((L.n[0] + xdiv - 1) / xdiv, L.n[1], L.n[2]);
}

bool rb_ok(const MgLevel& L)
{
	for (int d = 0; d < 3; d++) if (L.n[d] > 1 && L.n[d] % 2) return false;
	return true;
}

bool fused_ok(const MgLevel& L)
{
	for (int d = 0; d < 3; d++) if (L.n[d] < 4 || L.n[d] % 2) return false;
	return true;
}

} // namespace

hipError_t mg_create(PoissonMG& mg, const int dims[3], const double h[3])
{
	mg_destroy(mg);
	int n[3] = {dims[0], dims[1], dims[2]};
	double hh[3] = {h[0], h[1], h[2]};
	for (;;) {
		MgLevel L;
		for (int d = 0; d < 3; d++) {
			L.n[d] = n[d];
			L.h[d] = hh[d];
			L.hh[d] = hh[d] * hh[d];
		}
		L.diag = 0;
		for (int d = 0; d < 3; d++) if (n[d] > 1) L.diag += 2 / (hh[d] * hh[d]);
		L.ncell = n[0] * n[1] * n[2];
		L.packed = (rb_ok(L) && n[0] % 2 == 0) ? 1 : 0;
		hipError_t e;
		if ((e = hipMalloc(&L.phi, sizeof(double) * L.ncell)) != hipSuccess) return e;
		mg.lv.push_back(L);
		if ((e = hipMalloc(&mg.lv.back().f, sizeof(double) * L.ncell)) != hipSuccess) return e;
		if ((e = hipMalloc(&mg.lv.back().r, sizeof(double) * L.ncell)) != hipSuccess) return e;
		double hmin = 1e300;
		for (int d = 0; d < 3; d++) if (n[d] > 1 && hh[d] < hmin) hmin = hh[d];
		bool coarsened = false;
		for (int d = 0; d < 3; d++) {
			if (n[d] % 2 == 0 && n[d] >= 4 && hh[d] <= 1.5 * hmin) {
				n[d] /= 2;
				hh[d] *= 2;
				coarsened = true;
			}
		}
		if (!coarsened || L.ncell <= 64 || mg.lv.size() == 32) break;
	}
	/* the bottom levels solved in one block (at least two, none of them the
	   finest) */
	{
		const int nl = (int)mg.lv.size();
		int total = 0, l0 = nl;
		for (int l = nl - 1; l >= 1; l--) {
			const int c = mg.lv[l].ncell;
			if (c > TAIL_PER * 1024 || total + c > TAIL_CELLS || nl - l > TAIL_LEVELS) break;
			total += c;
			l0 = l;
		}
		mg.tail = nl - l0 >= 2 ? l0 : -1;
	}
	mg.partial_len = 1024;
	hipError_t e;
	if ((e = hipMalloc(&mg.partial, sizeof(double) * mg.partial_len)) != hipSuccess) return e;
	if ((e = hipMalloc(&mg.result, 2 * sizeof(double))) != hipSuccess) return e;
	{
		const MgLevel& F = mg.lv[0];
		const int nb = blocks_of(tiles_of(F.n[0], F.n[1], (F.n[2] + NORM_KZ - 1) / NORM_KZ));
		mg.norm_part_len = (size_t)3 * nb;
		if ((e = hipMalloc(&mg.norm_part, sizeof(double) * mg.norm_part_len)) != hipSuccess) return e;
	}
	if ((e = hipHostMalloc(&mg.host_result, 2 * sizeof(double))) != hipSuccess) return e;
	return hipSuccess;
}

void mg_destroy(PoissonMG& mg)
{
	for (auto& L : mg.lv) {
		(void)hipFree(L.phi);
		(void)hipFree(L.f);
		(void)hipFree(L.r);
	}
	mg.lv.clear();
	(void)hipFree(mg.partial);
	(void)hipFree(mg.result);
	(void)hipFree(mg.norm_part);
	mg.norm_part = nullptr;
	mg.norm_part_len = 0;
	(void)hipHostFree(mg.host_result);
	mg.partial = mg.result = mg.host_result = nullptr;
	mg.partial_len = 0;
}

hipError_t mg_sum_pairs(PoissonMG& mg, hipStream_t st, const double* pairs, int nb, double out[2])
{
	pair_final_kernel<<<1, SUM_BLOCK, 0, st>>>(pairs, nb, mg.result);
	hipError_t e = hipGetLastError();
	if (e != hipSuccess) return e;
	if ((e = hipMemcpyAsync(mg.host_result, mg.result, 2 * sizeof(double), hipMemcpyDeviceToHost, st)) != hipSuccess) return e;
	if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
	out[0] = mg.host_result[0];
	out[1] = mg.host_result[1];
	return hipSuccess;
}

hipError_t mg_sum(PoissonMG& mg, hipStream_t st, const double* a, int n, int mode, double shift,
	double* out)
{
	int blocks = (n + SUM_BLOCK * 8 - 1) / (SUM_BLOCK * 8);
	if (blocks < 1) blocks = 1;
	if (blocks > (int)mg.partial_len) blocks = (int)mg.partial_len;
	sum_partial_kernel<<<blocks, SUM_BLOCK, 0, st>>>(a, n, mode, shift, mg.partial);
	sum_final_kernel<<<1, SUM_BLOCK, 0, st>>>(mg.partial, blocks, mg.result);
	hipError_t e = hipGetLastError();
	if (e != hipSuccess) return e;
	if ((e = hipMemcpyAsync(mg.host_result, mg.result, sizeof(double), hipMemcpyDeviceToHost, st)) != hipSuccess) return e;
	if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
	*out = *mg.host_result;
	return hipSuccess;
}

void mg_smooth(PoissonMG& mg, hipStream_t st, int level, int sweeps)
{
	MgLevel& L = mg.lv[level];
	const LevelArgs A = args_of(L);
	const bool rb = rb_ok(L);
	if (L.ncell <= COARSE_MAX) {
		/* enough threads for COARSE_PER cells each, whole waves */
		int threads = (L.ncell + COARSE_PER - 1) / COARSE_PER;
		threads = threads < 1024 ? ((threads + 63) / 64) * 64 : 1024;
		coarse_kernel<<<1, threads, 0, st>>>(A, L.phi, L.f, sweeps, rb ? 1 : 0);
		return;
	}
	if (rb && mg.fused_smoother && fused_ok(L)) {
		/* about 1024 blocks (4 per CU, all resident at once): split the
		   march into z-chunks */
		const int tiles = ((L.n[0] + TX - 1) / TX) * ((L.n[1] + TY - 1) / TY);
		int chunks = (1024 + tiles - 1) / tiles;
		if (chunks > L.n[2] / 4) chunks = L.n[2] / 4;
		if (chunks < 1) chunks = 1;
		const int kc = (L.n[2] + chunks - 1) / chunks;
		const dim3 grid((L.n[0] + TX - 1) / TX, (L.n[1] + TY - 1) / TY, (L.n[2] + kc - 1) / kc);
		for (int s = 0; s < sweeps; s++) {
			/* reads the old values from phi, writes the new ones to r,
			   then the buffers swap */
			rb_fused_kernel<<<grid, dim3(TX, TY, 1), 0, st>>>(A, L.r, L.f, L.phi, kc);
			std::swap(L.phi, L.r);
		}
		return;
	}
	const Tiles TL = level_tiles(L), TL2 = level_tiles(L, 2);
	for (int s = 0; s < sweeps; s++) {
		if (rb) {
			for (int color = 0; color < 2; color++) {
				rb_kernel<<<blocks_of(TL2), dim3(PBX, PBY, 1), 0, st>>>(TL2, A, L.phi, L.f, color);
			}
		} else {
			jacobi_kernel<<<blocks_of(TL), dim3(PBX, PBY, 1), 0, st>>>(TL, A, L.phi, L.f, L.r);
			std::swap(L.phi, L.r);
		}
	}
}

void mg_vcycle(PoissonMG& mg, hipStream_t st, int l)
{
	const int nlev = (int)mg.lv.size();
	if (l == mg.tail) {
		TailArgs A{};
		A.nlev = nlev - l;
		int off = 0;
		for (int m = 0; m < A.nlev; m++) {
			const MgLevel& L = mg.lv[l + m];
			A.L[m] = args_of(L);
			if (m + 1 < A.nlev) A.q[m] = ratio_of(L, mg.lv[l + m + 1]);
			A.off[m] = off;
			A.n[m] = L.ncell;
			A.rb[m] = rb_ok(L) ? 1 : 0;
			off += L.ncell;
		}
		tail_kernel<<<1, 1024, 0, st>>>(A, mg.lv[l].phi, mg.lv[l].f);
		return;
	}
	if (l == nlev - 1) {
		mg_smooth(mg, st, l, 60);
		return;
	}
	MgLevel& F = mg.lv[l];
	MgLevel& C = mg.lv[l + 1];
	const Ratio q = ratio_of(F, C);
	const Tiles TF = level_tiles(F), TC = level_tiles(C);
	const bool by2 = q.r0 == 2 && q.r1 == 2 && q.r2 == 2;
	mg_smooth(mg, st, l, 2);
	if (by2) {
		restrict_residual_kernel<<<blocks_of(TC), dim3(PBX, PBY, 1), 0, st>>>(TC, args_of(F), args_of(C),
			q.w, F.phi, F.f, C.f, C.phi);
	} else {
		residual_kernel<<<blocks_of(TF), dim3(PBX, PBY, 1), 0, st>>>(TF, args_of(F), F.phi, F.f, F.r);
		restrict_kernel<<<blocks_of(TC), dim3(PBX, PBY, 1), 0, st>>>(TC, args_of(F), args_of(C), q,
			F.r, C.f, C.phi);
	}
	mg_vcycle(mg, st, l + 1);
	if (by2) {
		prolong8_kernel<<<blocks_of(TC), dim3(PBX, PBY, 1), 0, st>>>(TC, args_of(F), args_of(C), F.phi, C.phi);
	} else {
		prolong_kernel<<<blocks_of(TF), dim3(PBX, PBY, 1), 0, st>>>(TF, args_of(F), args_of(C), q,
			F.phi, C.phi);
	}
	mg_smooth(mg, st, l, 2);
}

/* ||r - mean(r)||_2 of the finest level's residual r = f - L phi, or of f
   alone (with_lap 0) */
static hipError_t centred_norm(PoissonMG& mg, hipStream_t st, int with_lap, double* out)
{
	const MgLevel& F = mg.lv[0];
	Tiles T = tiles_of(F.n[0], F.n[1], (F.n[2] + NORM_KZ - 1) / NORM_KZ);
	const int nb = blocks_of(T);
	norm_partial_kernel<<<nb, dim3(PBX, PBY, 1), 0, st>>>(T, args_of(F), F.phi, F.f, with_lap,
		mg.norm_part);
	norm_final_kernel<<<1, NORM_FINAL, 0, st>>>(mg.norm_part, nb, mg.result);
	hipError_t e = hipGetLastError();
	if (e != hipSuccess) return e;
	if ((e = hipMemcpyAsync(mg.host_result, mg.result, sizeof(double), hipMemcpyDeviceToHost, st)) != hipSuccess) return e;
	if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
	*out = std::sqrt(*mg.host_result);
	return hipSuccess;
}

__global__ void fill_kernel(double* a, int n, double v)
{
	const int c = blockIdx.x * blockDim.x + threadIdx.x;
	if (c < n) a[c] = v;
}

hipError_t mg_solve(PoissonMG& mg, hipStream_t st, double tol, double atol, int max_cycles,
	double* rel, int* cycles, bool* converged)
{
	MgLevel& F = mg.lv[0];
	fill_kernel<<<(F.ncell + 255) / 256, 256, 0, st>>>(F.phi, F.ncell, 0.0);
	hipError_t e = hipGetLastError();
	if (e != hipSuccess) return e;
	double fnorm;
	if ((e = centred_norm(mg, st, 0, &fnorm)) != hipSuccess) return e;
	const double target = tol * fnorm > atol ? tol * fnorm : atol;
	double rnorm = fnorm;
	int it = 0;
	if (fnorm > target) {
		for (it = 1; it <= max_cycles; it++) {
			mg_vcycle(mg, st, 0);
			if ((e = centred_norm(mg, st, 1, &rnorm)) != hipSuccess) return e;
			if (rnorm <= target) break;
		}
		if (it > max_cycles) it = max_cycles;
	}
	*rel = fnorm > 0 ? rnorm / fnorm : 0;
	*cycles = it;
	*converged = !(rnorm > target);
	return hipSuccess;
}

} // namespace pamhd_dev
