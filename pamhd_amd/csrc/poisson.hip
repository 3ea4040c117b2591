/*
Multigrid Poisson solver for divergence cleaning on MI355X (gfx950).

Every level keeps phi, f and r as FP64 arrays (x fastest).  Kernels:
  rb_kernel        one colour of a red-black Gauss-Seidel sweep; a thread
                   owns one cell of the colour (x index 2t or 2t+1)
  jacobi_kernel    damped Jacobi into r (phi and r are then swapped)
  residual_kernel  r = f - L phi
  restrict_kernel  coarse f = mean of the children's r, coarse phi = 0
  prolong_kernel   fine phi += trilinear interpolation of coarse phi
  coarse_kernel    all sweeps of a small level inside one block, in LDS
  sum kernels      two-stage deterministic sums for the stopping test
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
};

LevelArgs args_of(const MgLevel& L)
{
	return LevelArgs{L.n[0], L.n[1], L.n[2], L.hh[0], L.hh[1], L.hh[2], L.diag};
}

__device__ __forceinline__ int wrapi(int i, int n)
{
	return i < 0 ? i + n : (i >= n ? i - n : i);
}

__device__ __forceinline__ int lidx(const LevelArgs& L, int i, int j, int k)
{
	return wrapi(i, L.n0) + L.n0 * (wrapi(j, L.n1) + L.n1 * wrapi(k, L.n2));
}

/* (sum of neighbours / h^2 - f) / diag, or phi itself without active
   dimensions */
__device__ __forceinline__ double relax_value(const LevelArgs& L, const double* phi,
	const double* f, int i, int j, int k)
{
	double s = 0;
	if (L.n0 > 1) s += (phi[lidx(L, i + 1, j, k)] + phi[lidx(L, i - 1, j, k)]) / L.hh0;
	if (L.n1 > 1) s += (phi[lidx(L, i, j + 1, k)] + phi[lidx(L, i, j - 1, k)]) / L.hh1;
	if (L.n2 > 1) s += (phi[lidx(L, i, j, k + 1)] + phi[lidx(L, i, j, k - 1)]) / L.hh2;
	const int c = lidx(L, i, j, k);
	if (L.diag == 0) return phi[c];
	return (s - f[c]) / L.diag;
}

__device__ __forceinline__ double apply_lap(const LevelArgs& L, const double* phi, int i, int j, int k)
{
	const double c = phi[lidx(L, i, j, k)];
	double s = 0;
	if (L.n0 > 1) s += (phi[lidx(L, i + 1, j, k)] - 2 * c + phi[lidx(L, i - 1, j, k)]) / L.hh0;
	if (L.n1 > 1) s += (phi[lidx(L, i, j + 1, k)] - 2 * c + phi[lidx(L, i, j - 1, k)]) / L.hh1;
	if (L.n2 > 1) s += (phi[lidx(L, i, j, k + 1)] - 2 * c + phi[lidx(L, i, j, k - 1)]) / L.hh2;
	return s;
}

__global__ __launch_bounds__(PBX * PBY) void rb_kernel(LevelArgs L, double* __restrict__ phi,
	const double* __restrict__ f, int color)
{
	const int t = blockIdx.x * PBX + threadIdx.x;
	const int j = blockIdx.y * PBY + threadIdx.y;
	const int k = blockIdx.z;
	if (j >= L.n1) return;
	const int i = 2 * t + ((j + k + color) & 1);
	if (i >= L.n0) return;
	phi[i + L.n0 * (j + L.n1 * k)] = relax_value(L, phi, f, i, j, k);
}

__global__ __launch_bounds__(PBX * PBY) void jacobi_kernel(LevelArgs L, const double* __restrict__ phi,
	const double* __restrict__ f, double* __restrict__ out)
{
	const int i = blockIdx.x * PBX + threadIdx.x;
	const int j = blockIdx.y * PBY + threadIdx.y;
	const int k = blockIdx.z;
	if (i >= L.n0 || j >= L.n1) return;
	const int c = i + L.n0 * (j + L.n1 * k);
	out[c] = phi[c] + 0.8 * (relax_value(L, phi, f, i, j, k) - phi[c]);
}

__global__ __launch_bounds__(PBX * PBY) void residual_kernel(LevelArgs L, const double* __restrict__ phi,
	const double* __restrict__ f, double* __restrict__ r)
{
	const int i = blockIdx.x * PBX + threadIdx.x;
	const int j = blockIdx.y * PBY + threadIdx.y;
	const int k = blockIdx.z;
	if (i >= L.n0 || j >= L.n1) return;
	const int c = i + L.n0 * (j + L.n1 * k);
	r[c] = f[c] - apply_lap(L, phi, i, j, k);
}

__global__ __launch_bounds__(PBX * PBY) void restrict_kernel(LevelArgs F, LevelArgs C,
	const double* __restrict__ r, double* __restrict__ cf, double* __restrict__ cphi)
{
	const int i = blockIdx.x * PBX + threadIdx.x;
	const int j = blockIdx.y * PBY + threadIdx.y;
	const int k = blockIdx.z;
	if (i >= C.n0 || j >= C.n1) return;
	const int r0 = F.n0 / C.n0, r1 = F.n1 / C.n1, r2 = F.n2 / C.n2;
	const double w = 1.0 / (double)(r0 * r1 * r2);
	double s = 0;
	for (int c = 0; c < r2; c++)
	for (int b = 0; b < r1; b++)
	for (int a = 0; a < r0; a++) {
		s += r[(i * r0 + a) + F.n0 * ((j * r1 + b) + F.n1 * (k * r2 + c))];
	}
	const int o = i + C.n0 * (j + C.n1 * k);
	cf[o] = s * w;
	cphi[o] = 0;
}

__global__ __launch_bounds__(PBX * PBY) void prolong_kernel(LevelArgs F, LevelArgs C,
	double* __restrict__ phi, const double* __restrict__ cphi)
{
	const int i = blockIdx.x * PBX + threadIdx.x;
	const int j = blockIdx.y * PBY + threadIdx.y;
	const int k = blockIdx.z;
	if (i >= F.n0 || j >= F.n1) return;
	const int idx[3] = {i, j, k};
	const int ratio[3] = {F.n0 / C.n0, F.n1 / C.n1, F.n2 / C.n2};
	int ci[3][2];
	double cw[3][2];
	int cnt[3];
	for (int d = 0; d < 3; d++) {
		if (ratio[d] == 2) {
			const int I = idx[d] / 2, side = (idx[d] % 2) ? 1 : -1;
			ci[d][0] = I;
			cw[d][0] = 0.75;
			ci[d][1] = I + side;
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
		s += cw[0][a] * cw[1][b] * cw[2][c] * cphi[lidx(C, ci[0][a], ci[1][b], ci[2][c])];
	}
	phi[i + F.n0 * (j + F.n1 * k)] += s;
}

/* sweeps of a small level (<= COARSE_MAX cells) in one block, in LDS */
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
	for (int s = 0; s < sweeps; s++) {
		if (red_black) {
			for (int color = 0; color < 2; color++) {
				for (int c = threadIdx.x; c < n; c += blockDim.x) {
					const int i = c % L.n0, j = (c / L.n0) % L.n1, k = c / (L.n0 * L.n1);
					if (((i + j + k) & 1) != color) continue;
					phi[c] = relax_value(L, phi, f, i, j, k);
				}
				__syncthreads();
			}
		} else {
			for (int c = threadIdx.x; c < n; c += blockDim.x) {
				const int i = c % L.n0, j = (c / L.n0) % L.n1, k = c / (L.n0 * L.n1);
				tmp[c] = phi[c] + 0.8 * (relax_value(L, phi, f, i, j, k) - phi[c]);
			}
			__syncthreads();
			for (int c = threadIdx.x; c < n; c += blockDim.x) phi[c] = tmp[c];
			__syncthreads();
		}
	}
	for (int c = threadIdx.x; c < n; c += blockDim.x) phi_g[c] = phi[c];
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
				po[r] = phi_in[gi + L.n0 * (gj + L.n1 * kk)];
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
				pf[r] = f[gi + L.n0 * (gj + L.n1 * kk)];
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
		const int c = gi + L.n0 * (gj + L.n1 * q);
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

dim3 level_grid(const MgLevel& L, int xdiv = 1)
{
	const int nx = (L.n[0] + xdiv - 1) / xdiv;
	return dim3((nx + PBX - 1) / PBX, (L.n[1] + PBY - 1) / PBY, L.n[2]);
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
	mg.partial_len = 1024;
	hipError_t e;
	if ((e = hipMalloc(&mg.partial, sizeof(double) * mg.partial_len)) != hipSuccess) return e;
	if ((e = hipMalloc(&mg.result, sizeof(double))) != hipSuccess) return e;
	if ((e = hipHostMalloc(&mg.host_result, sizeof(double))) != hipSuccess) return e;
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
	(void)hipHostFree(mg.host_result);
	mg.partial = mg.result = mg.host_result = nullptr;
	mg.partial_len = 0;
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
		coarse_kernel<<<1, 1024, 0, st>>>(A, L.phi, L.f, sweeps, rb ? 1 : 0);
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
	for (int s = 0; s < sweeps; s++) {
		if (rb) {
			for (int color = 0; color < 2; color++) {
				rb_kernel<<<level_grid(L, 2), dim3(PBX, PBY, 1), 0, st>>>(A, L.phi, L.f, color);
			}
		} else {
			jacobi_kernel<<<level_grid(L), dim3(PBX, PBY, 1), 0, st>>>(A, L.phi, L.f, L.r);
			std::swap(L.phi, L.r);
		}
	}
}

void mg_vcycle(PoissonMG& mg, hipStream_t st, int l)
{
	const int nlev = (int)mg.lv.size();
	if (l == nlev - 1) {
		mg_smooth(mg, st, l, 60);
		return;
	}
	MgLevel& F = mg.lv[l];
	MgLevel& C = mg.lv[l + 1];
	mg_smooth(mg, st, l, 2);
	residual_kernel<<<level_grid(F), dim3(PBX, PBY, 1), 0, st>>>(args_of(F), F.phi, F.f, F.r);
	restrict_kernel<<<level_grid(C), dim3(PBX, PBY, 1), 0, st>>>(args_of(F), args_of(C), F.r, C.f, C.phi);
	mg_vcycle(mg, st, l + 1);
	prolong_kernel<<<level_grid(F), dim3(PBX, PBY, 1), 0, st>>>(args_of(F), args_of(C), F.phi, C.phi);
	mg_smooth(mg, st, l, 2);
}

static hipError_t l2_centred(PoissonMG& mg, hipStream_t st, const double* a, int n, double* out)
{
	double s;
	hipError_t e = mg_sum(mg, st, a, n, 0, 0.0, &s);
	if (e != hipSuccess) return e;
	const double m = s / (double)n;
	if ((e = mg_sum(mg, st, a, n, 1, m, &s)) != hipSuccess) return e;
	*out = std::sqrt(s);
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
	if ((e = l2_centred(mg, st, F.f, F.ncell, &fnorm)) != hipSuccess) return e;
	const double target = tol * fnorm > atol ? tol * fnorm : atol;
	double rnorm = fnorm;
	int it = 0;
	if (fnorm > target) {
		for (it = 1; it <= max_cycles; it++) {
			mg_vcycle(mg, st, 0);
			residual_kernel<<<level_grid(F), dim3(PBX, PBY, 1), 0, st>>>(args_of(F), F.phi, F.f, F.r);
			if ((e = hipGetLastError()) != hipSuccess) return e;
			if ((e = l2_centred(mg, st, F.r, F.ncell, &rnorm)) != hipSuccess) return e;
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
