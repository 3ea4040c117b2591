/*
Hand-written adjoint of the periodic MHD step with a frozen time step
(HLL fluxes, no background field), for gradients of a loss of the final
state with respect to the initial state.

The forward step restated as a map U^n -> U^{n+1} (dt given):
  face states    L, R of each d-face: rho, m, E of the two cells, B the
                 cell-centred field vb = (face + next face) / 2, rotated
  fluxes         P = dt * F_HLL(L, R), rotated back
  CT             face' = face + dB(P) / area      (12 edge terms per face)
  update         rho' = rho + sum_d (P(d, c) - P(d, c + e_d)) / len_d, same
                 for m, E and vb_flux = vb + sum_d (...)
  refits         E'' = E' + (|vb_new|^2 - |vb_flux|^2) / (2 mu0) with
                 vb_new the average of the new faces (the thermal-pressure
                 refit of update_vol_B, source/mhd/solve.hpp:1646-1672; the
                 later refits with vb_old = vb_new are the identity, and the
                 kinetic terms cancel exactly)
The reverse step takes lambda^{n+1} (the loss's gradient with respect to
U^{n+1}, same layout as the state) to lambda^n in four kernels, with
16 n doubles of scratch in 2-D (about 22 n in 3-D):
  pe_kernel      the forward step's six magnetic-flux (edge field) parts of
                 the three faces of each cell: Pe (6 n); a thread per face
  g_kernel       mu_new = lambda_E vb_new / mu0 at the cell and at its lower
                 neighbours, G = lambda_face + (mu_new(c) + mu_new(c - e_d)) / 2
                 (3 n), and mu_flux = -lambda_E vb_flux / mu0 (3 n); a thread
                 per cell and direction
  tile_kernel    per tile of 256 cells: lambda_P of each face owned by a cell
                 from the update (conserved and vb_flux) and the transposed
                 12-term CT sums over G, then lambda_L, lambda_R = J^T (dt
                 lambda_P) (the HLL Jacobian by reverse accumulation, or by
                 forward-mode dual numbers; a face that joins a cell to
                 itself, in a grid one cell thick, by the physical flux's
                 Jacobian, both sides at once); lambda_R is added to the cell's
                 accumulator, lambda_L goes to the left cell through LDS when
                 that cell is in the tile and to a per-tile boundary buffer
                 otherwise; the accumulators (lambda rho, m, E and lambda_vb,
                 8 n) are written out (over Pe, which is dead by then)
  final_kernel   adds the boundary buffer's terms, writes lambda rho, m, E and
                 lambda_face = G + (lambda_vb(c) + lambda_vb(c - e_d)) / 2;
                 in a run of reverse steps whose magnetic fluxes are given it
                 also makes the next step's G and mu_flux (final_g_kernel),
                 taking that step's new faces from the state it ended at
Each flux Jacobian is evaluated once per face; every sum has a fixed order,
so a reverse step is reproducible bit for bit.
*/
#include <hip/hip_runtime.h>

#include <cmath>

#include "pamhd_mi355x.h"
#include "riemann.hpp"

namespace pamhd_dev {

namespace {

constexpr int ABX = 64, ABY = 4;
constexpr int JN = 4;   /* seeded directions per dual-number pass of the flux Jacobian */

/* value and derivatives along N seeded input directions */
template <int N> struct Dual {
	double v;
	double d[N];
};

template <int N> __device__ __forceinline__ Dual<N> cst(double a)
{
	Dual<N> r;
	r.v = a;
	for (int q = 0; q < N; q++) r.d[q] = 0;
	return r;
}
template <int N> __device__ __forceinline__ Dual<N> operator+(const Dual<N>& a, const Dual<N>& b)
{
	Dual<N> r;
	r.v = a.v + b.v;
	for (int q = 0; q < N; q++) r.d[q] = a.d[q] + b.d[q];
	return r;
}
template <int N> __device__ __forceinline__ Dual<N> operator-(const Dual<N>& a, const Dual<N>& b)
{
	Dual<N> r;
	r.v = a.v - b.v;
	for (int q = 0; q < N; q++) r.d[q] = a.d[q] - b.d[q];
	return r;
}
template <int N> __device__ __forceinline__ Dual<N> operator-(const Dual<N>& a)
{
	Dual<N> r;
	r.v = -a.v;
	for (int q = 0; q < N; q++) r.d[q] = -a.d[q];
	return r;
}
template <int N> __device__ __forceinline__ Dual<N> operator*(const Dual<N>& a, const Dual<N>& b)
{
	Dual<N> r;
	r.v = a.v * b.v;
	for (int q = 0; q < N; q++) r.d[q] = a.d[q] * b.v + a.v * b.d[q];
	return r;
}
template <int N> __device__ __forceinline__ Dual<N> operator/(const Dual<N>& a, const Dual<N>& b)
{
	Dual<N> r;
	r.v = a.v / b.v;
	const double ib = 1 / b.v;
	for (int q = 0; q < N; q++) r.d[q] = (a.d[q] - r.v * b.d[q]) * ib;
	return r;
}
template <int N> __device__ __forceinline__ Dual<N> dsqrt(const Dual<N>& a)
{
	Dual<N> r;
	r.v = __builtin_sqrt(a.v);
	const double h = 0.5 / r.v;
	for (int q = 0; q < N; q++) r.d[q] = a.d[q] * h;
	return r;
}
template <int N> __device__ __forceinline__ Dual<N> dabs(const Dual<N>& a)
{
	return a.v < 0 ? -a : a;
}
/* std::max / std::min tie rules (first argument wins ties) */
template <int N> __device__ __forceinline__ const Dual<N>& dmax(const Dual<N>& a, const Dual<N>& b)
{
	return (a.v < b.v) ? b : a;
}
template <int N> __device__ __forceinline__ const Dual<N>& dmin(const Dual<N>& a, const Dual<N>& b)
{
	return (b.v < a.v) ? b : a;
}

/*
HLL flux of source/mhd/hll_athena.hpp:57-275 without background field on
dual numbers, in the operation order of the FP64 device solver (the value
parts equal its results).  Returns false where the forward solver would
report an error; such faces get no gradient.
*/
template <int N> __device__ bool hll_dual(const Dual<N> L[8], const Dual<N> R[8], const Phys& ph,
	Dual<N> F[8])
{
	using D = Dual<N>;
	const D half = cst<N>(0.5), one = cst<N>(1.0), two = cst<N>(2.0), zero = cst<N>(0.0);
	const D gamma = cst<N>(ph.gamma), mu0 = cst<N>(ph.mu0), inv_mu0 = cst<N>(ph.inv_mu0);
	const D gm1 = cst<N>(ph.gm1), hinv = cst<N>(ph.half_inv_mu0);
	if (L[RHO].v <= 0 || R[RHO].v <= 0) return false;
	D p[2], v[2][3], cf[2];
	const D* U[2] = {L, R};
	for (int s = 0; s < 2; s++) {
		const D* u = U[s];
		/* pressure64 */
		const D kinetic = half / u[RHO] * (u[M0] * u[M0] + u[M1] * u[M1] + u[M2] * u[M2]);
		const D magnetic = hinv * (u[B0] * u[B0] + u[B1] * u[B1] + u[B2] * u[B2]);
		p[s] = (u[NRJ] - kinetic - magnetic) * gm1;
		if (!(p[s].v > 0) || !__builtin_isnormal(p[s].v)) return false;
		const D inv = one / u[RHO];
		for (int i = 0; i < 3; i++) v[s][i] = u[M0 + i] * inv;
		/* fast_speed */
		const D mag = dsqrt(u[B0] * u[B0] + u[B1] * u[B1] + u[B2] * u[B2]);
		const D sound = dsqrt(gamma * p[s] / u[RHO]);
		if (mag.v <= 0) {
			cf[s] = sound;
		} else {
			const D alfven = mag / dsqrt(mu0 * u[RHO]);
			const D sound2 = sound * sound, alfven2 = alfven * alfven, speeds2 = sound2 + alfven2;
			const D ratio = u[B0] / mag;
			const D to_sqrt = sound2 * sound2 + alfven2 * alfven2
				+ two * sound2 * alfven2 * (one - two * (ratio * ratio));
			if (!(to_sqrt.v > 0)) return false;
			cf[s] = dsqrt(half * (speeds2 + dsqrt(to_sqrt)));
		}
	}
	const D ms = dmax(cf[0], cf[1]);
	const bool lo = v[0][0].v <= v[1][0].v;
	const D msn = lo ? v[0][0] - ms : v[1][0] - ms;
	const D msp = lo ? v[1][0] + ms : v[0][0] + ms;
	const D bm = dmin(msn, zero), bp = dmax(msp, zero);
	const double den0 = bp.v - bm.v;
	if (!__builtin_isnormal(den0) || den0 < 0) {
		for (int q = 0; q < 8; q++) F[q] = zero;
		return true;
	}
	D FF[2][8];
	for (int s = 0; s < 2; s++) {
		const D* u = U[s];
		const D* B = u + B0;
		const D pB1 = half * inv_mu0 * (B[0] * B[0] + B[1] * B[1] + B[2] * B[2]);
		D* f = FF[s];
		f[RHO] = u[M0];
		for (int i = 0; i < 3; i++) f[M0 + i] = u[M0 + i] * v[s][0] + (B[i] * B[0]) * (-inv_mu0);
		f[M0] = f[M0] + (p[s] + pB1);
		f[NRJ] = v[s][0] * (u[NRJ] + p[s] + pB1)
			- B[0] * (v[s][0] * B[0] + v[s][1] * B[1] + v[s][2] * B[2]) * inv_mu0;
		for (int i = 0; i < 3; i++) f[B0 + i] = B[i] * v[s][0] + v[s][i] * (-B[0]);
		f[B0] = zero;
	}
	D* FL = FF[0];
	D* FR = FF[1];
	FL[RHO] = FL[RHO] - L[RHO] * bm;
	FL[NRJ] = FL[NRJ] - L[NRJ] * bm;
	FR[RHO] = FR[RHO] - R[RHO] * bp;
	FR[NRJ] = FR[NRJ] - R[NRJ] * bp;
	for (int i = 0; i < 3; i++) {
		FL[M0 + i] = FL[M0 + i] + L[M0 + i] * (-bm);
		FL[B0 + i] = FL[B0 + i] + L[B0 + i] * (-bm);
		FR[M0 + i] = FR[M0 + i] + R[M0 + i] * (-bp);
		FR[B0 + i] = FR[B0 + i] + R[B0 + i] * (-bp);
	}
	FR[B0] = FL[B0] = zero;
	const D den = bp - bm, sum = bp + bm;
	F[RHO] = (FL[RHO] + FR[RHO]) / two + (FL[RHO] - FR[RHO]) * sum / den / two;
	F[NRJ] = (FL[NRJ] + FR[NRJ]) / two + (FL[NRJ] - FR[NRJ]) * sum / den / two;
	const D w = sum / den / two;
	for (int i = 0; i < 3; i++) {
		F[M0 + i] = (FL[M0 + i] + FR[M0 + i]) * half + (FL[M0 + i] - FR[M0 + i]) * w;
		F[B0 + i] = (FL[B0 + i] + FR[B0 + i]) * half + (FL[B0 + i] - FR[B0 + i]) * w;
	}
	return true;
}

/*
J^T g of the same HLL flux by reverse accumulation: the forward pass
recomputes the intermediate values, the reverse pass carries the adjoint
g of the face-frame flux back to both sides' inputs (gL, gR).  It returns
what hll_dual returns (false: no gradient), and the same derivatives up to
rounding, for about three flux evaluations instead of the 16 seeded
directions of the dual-number form.  The fast speed is differentiated
through s2 = gamma p / rho, a2 = |B|^2 / (mu0 rho) and r2 = B_n^2 / |B|^2
(cf^2 = (s2 + a2 + sqrt(ts)) / 2, ts = s2^2 + a2^2 + 2 s2 a2 (1 - 2 r2)),
which needs two square roots and two reciprocals per side, and every
quotient is a product with a reciprocal made once: each differs from the
solver's own value by a few ulp.
*/
__device__ bool hll_adjoint(const double L[8], const double R[8], const Phys& ph,
	const double gF[8], double gL[8], double gR[8])
{
	for (int q = 0; q < 8; q++) gL[q] = gR[q] = 0;
	if (L[RHO] <= 0 || R[RHO] <= 0) return false;
	const double gamma = ph.gamma, inv_mu0 = ph.inv_mu0;
	const double gm1 = ph.gm1, hinv = ph.half_inv_mu0;
	const double* U[2] = {L, R};
	double* gU[2] = {gL, gR};
	double p[2], v[2][3], cf[2], s2[2], a2[2] = {0, 0}, r2[2] = {0, 0}, rts[2] = {1, 1};
	double ibsq[2] = {0, 0}, bsq[2], msq[2], irho[2];
	bool hasB[2];
	for (int s = 0; s < 2; s++) {
		const double* u = U[s];
		const double ir = 1 / u[RHO];
		irho[s] = ir;
		msq[s] = u[M0] * u[M0] + u[M1] * u[M1] + u[M2] * u[M2];
		bsq[s] = u[B0] * u[B0] + u[B1] * u[B1] + u[B2] * u[B2];
		p[s] = (u[NRJ] - 0.5 * ir * msq[s] - hinv * bsq[s]) * gm1;
		if (!(p[s] > 0) || !__builtin_isnormal(p[s])) return false;
		for (int i = 0; i < 3; i++) v[s][i] = u[M0 + i] * ir;
		s2[s] = gamma * p[s] * ir;
		hasB[s] = bsq[s] > 0;
		if (!hasB[s]) {
			cf[s] = __builtin_sqrt(s2[s]);
		} else {
			a2[s] = bsq[s] * inv_mu0 * ir;
			ibsq[s] = 1 / bsq[s];
			r2[s] = u[B0] * u[B0] * ibsq[s];
			const double ts = s2[s] * s2[s] + a2[s] * a2[s] + 2 * s2[s] * a2[s] * (1 - 2 * r2[s]);
			if (!(ts > 0)) return false;
			rts[s] = __builtin_sqrt(ts);
			cf[s] = __builtin_sqrt(0.5 * (s2[s] + a2[s] + rts[s]));
		}
	}
	/* wave speeds with the tie rules of std::max / std::min */
	const int ims = (cf[0] < cf[1]) ? 1 : 0;
	const double ms = cf[ims];
	const bool lo = v[0][0] <= v[1][0];
	const int sn = lo ? 0 : 1, sp = lo ? 1 : 0;
	const double msn = v[sn][0] - ms, msp = v[sp][0] + ms;
	const bool take_n = !(0.0 < msn), take_p = !(msp < 0.0);
	const double bm = take_n ? msn : 0.0, bp = take_p ? msp : 0.0;
	const double den0 = bp - bm;
	if (!__builtin_isnormal(den0) || den0 < 0) return true;   /* zero flux */
	/* physical fluxes, then the HLL combination */
	double FF[2][8], pB1[2], vB[2];
	for (int s = 0; s < 2; s++) {
		const double* u = U[s];
		const double* B = u + B0;
		pB1[s] = 0.5 * inv_mu0 * bsq[s];
		vB[s] = v[s][0] * B[0] + v[s][1] * B[1] + v[s][2] * B[2];
		double* f = FF[s];
		f[RHO] = u[M0];
		for (int i = 0; i < 3; i++) f[M0 + i] = u[M0 + i] * v[s][0] - B[i] * B[0] * inv_mu0;
		f[M0] += p[s] + pB1[s];
		f[NRJ] = v[s][0] * (u[NRJ] + p[s] + pB1[s]) - B[0] * vB[s] * inv_mu0;
		for (int i = 0; i < 3; i++) f[B0 + i] = B[i] * v[s][0] - v[s][i] * B[0];
		f[B0] = 0;
	}
	const double den = bp - bm, sum = bp + bm, rden = 1 / den, i2den = 0.5 * rden, w = sum * i2den;
	double gFL[8], gFR[8], gw = 0;
	for (int q = 0; q < 8; q++) {
		if (q == B0) { gFL[q] = gFR[q] = 0; continue; }
		const double fl = FF[0][q] - L[q] * bm, fr = FF[1][q] - R[q] * bp;
		gFL[q] = gF[q] * (0.5 + w);
		gFR[q] = gF[q] * (0.5 - w);
		gw += gF[q] * (fl - fr);
	}
	const double gsum = gw * i2den, gden = -gw * sum * i2den * rden;
	double gbp = gsum + gden, gbm = gsum - gden;
	for (int q = 0; q < 8; q++) {
		if (q == B0) continue;
		gL[q] -= gFL[q] * bm;
		gbm -= gFL[q] * L[q];
		gR[q] -= gFR[q] * bp;
		gbp -= gFR[q] * R[q];
	}
	double gp[2] = {0, 0}, gv[2][3] = {{0, 0, 0}, {0, 0, 0}}, gcf[2] = {0, 0};
	for (int s = 0; s < 2; s++) {
		const double* gf = s ? gFR : gFL;
		const double* u = U[s];
		const double* B = u + B0;
		double* gu = gU[s];
		double gpB1 = 0;
		gu[M0] += gf[RHO];
		for (int i = 0; i < 3; i++) {
			gu[M0 + i] += gf[M0 + i] * v[s][0];
			gv[s][0] += gf[M0 + i] * u[M0 + i];
			gu[B0 + i] -= gf[M0 + i] * B[0] * inv_mu0;
			gu[B0] -= gf[M0 + i] * B[i] * inv_mu0;
		}
		gp[s] += gf[M0];
		gpB1 += gf[M0];
		gv[s][0] += gf[NRJ] * (u[NRJ] + p[s] + pB1[s]);
		gu[NRJ] += gf[NRJ] * v[s][0];
		gp[s] += gf[NRJ] * v[s][0];
		gpB1 += gf[NRJ] * v[s][0];
		gu[B0] -= gf[NRJ] * vB[s] * inv_mu0;
		const double gvB = -gf[NRJ] * B[0] * inv_mu0;
		for (int i = 0; i < 3; i++) {
			gv[s][i] += gvB * B[i];
			gu[B0 + i] += gvB * v[s][i];
		}
		for (int i = 1; i < 3; i++) {
			gu[B0 + i] += gf[B0 + i] * v[s][0];
			gv[s][0] += gf[B0 + i] * B[i];
			gv[s][i] -= gf[B0 + i] * B[0];
			gu[B0] -= gf[B0 + i] * v[s][i];
		}
		for (int i = 0; i < 3; i++) gu[B0 + i] += gpB1 * inv_mu0 * B[i];
	}
	double gms = 0;
	if (take_n) { gv[sn][0] += gbm; gms -= gbm; }
	if (take_p) { gv[sp][0] += gbp; gms += gbp; }
	gcf[ims] += gms;
	for (int s = 0; s < 2; s++) {
		const double* u = U[s];
		double* gu = gU[s];
		const double ir = irho[s];
		/* cf = sqrt(Q): gQ = gcf / (2 cf) */
		const double gQ = gcf[s] * 0.5 / cf[s];
		double gs2, gbsq = 0;
		if (!hasB[s]) {
			gs2 = gQ;
		} else {
			/* Q = (s2 + a2 + rts) / 2, rts = sqrt(ts) */
			gs2 = 0.5 * gQ;
			double ga2 = 0.5 * gQ;
			const double gts = 0.25 * gQ / rts[s];
			const double t = 1 - 2 * r2[s];
			gs2 += gts * (2 * s2[s] + 2 * a2[s] * t);
			ga2 += gts * (2 * a2[s] + 2 * s2[s] * t);
			const double gr2 = -4 * gts * s2[s] * a2[s];
			/* r2 = B_n^2 / bsq, a2 = bsq / (mu0 rho) */
			gu[B0] += gr2 * 2 * u[B0] * ibsq[s];
			gbsq += ga2 * inv_mu0 * ir - gr2 * r2[s] * ibsq[s];
			gu[RHO] -= ga2 * a2[s] * ir;
		}
		/* s2 = gamma p / rho */
		gp[s] += gs2 * gamma * ir;
		gu[RHO] -= gs2 * s2[s] * ir;
		for (int i = 0; i < 3; i++) gu[B0 + i] += 2 * u[B0 + i] * gbsq;
		for (int i = 0; i < 3; i++) {
			gu[M0 + i] += gv[s][i] * ir;
			gu[RHO] -= gv[s][i] * v[s][i] * ir;
		}
		const double gpg = gp[s] * gm1;
		gu[NRJ] += gpg;
		for (int i = 0; i < 3; i++) {
			gu[M0 + i] -= gpg * v[s][i];
			gu[B0 + i] -= gpg * 2 * hinv * u[B0 + i];
		}
		gu[RHO] += gpg * 0.5 * msq[s] * ir * ir;
	}
	return true;
}

struct AdjGrid {
	int nx, ny, nz, n;
	double len[3];
	double inv_len[3];    /* 1 / len */
	double inv_area[3];   /* 1 / area of the d-faces */
	double dt;
	Phys ph;
};

/* tiles of tile_kernel: tx x ty x tz cells (256), ntx x nty x ntz of them;
   a tile keeps boundary slots for the faces of its lower x, y and z layers
   whose left cell lies in another tile (none in a dimension one tile spans) */
struct AdjTiles {
	int tx, ty, tz;
	int lx, ly, lz;   /* log2 of tx, ty, tz (powers of two) */
	int ntx, nty, ntz;
	int sx, sy, sz, slots;
};

AdjTiles tiles_for(int nx, int ny, int nz)
{
	AdjTiles T;
	T.tx = 64;
	T.ty = nz > 1 ? 2 : 4;
	T.tz = nz > 1 ? 2 : 1;
	T.lx = 6;
	T.ly = nz > 1 ? 1 : 2;
	T.lz = nz > 1 ? 1 : 0;
	T.ntx = (nx + T.tx - 1) / T.tx;
	T.nty = (ny + T.ty - 1) / T.ty;
	T.ntz = (nz + T.tz - 1) / T.tz;
	T.sx = T.ntx > 1 ? T.ty * T.tz : 0;
	T.sy = T.nty > 1 ? T.tx * T.tz : 0;
	T.sz = T.ntz > 1 ? T.tx * T.ty : 0;
	T.slots = T.sx + T.sy + T.sz;
	return T;
}

constexpr int TILE_CELLS = 256;

/* i in [-2n, 2n) wrapped into [0, n): new_face offsets a neighbour's
   coordinates by one and face_dB by one more, and n may be 1 */
__device__ __forceinline__ int wrapi(int i, int n)
{
	if (i < 0) i += n;
	if (i < 0) i += n;
	if (i >= n) i -= n;
	if (i >= n) i -= n;
	return i;
}

__device__ __forceinline__ int cell_at(const AdjGrid& g, int i, int j, int k)
{
	return wrapi(i, g.nx) + g.nx * (wrapi(j, g.ny) + g.ny * wrapi(k, g.nz));
}

/* neighbour of (i, j, k) at offset (a, b, c) */
__device__ __forceinline__ int nb(const AdjGrid& g, int i, int j, int k, int a, int b, int c)
{
	return cell_at(g, i + a, j + b, k + c);
}

__device__ __forceinline__ bool thread_cell(const AdjGrid& g, int& i, int& j, int& k)
{
	i = blockIdx.x * ABX + threadIdx.x;
	j = blockIdx.y * ABY + threadIdx.y;
	k = blockIdx.z;
	return i < g.nx && j < g.ny;
}

/* state of cell c rotated into the frame of a d-face (vb = face average) */
__device__ __forceinline__ void face_state(const AdjGrid& g, const double* __restrict__ s, int d,
	int i, int j, int k, double U[8])
{
	const int n = g.n;
	const int c = cell_at(g, i, j, k);
	const double m[3] = {s[n + c], s[2 * n + c], s[3 * n + c]};
	const double B[3] = {
		0.5 * (s[5 * n + c] + s[5 * n + nb(g, i, j, k, 1, 0, 0)]),
		0.5 * (s[6 * n + c] + s[6 * n + nb(g, i, j, k, 0, 1, 0)]),
		0.5 * (s[7 * n + c] + s[7 * n + nb(g, i, j, k, 0, 0, 1)])};
	U[RHO] = s[c];
	U[NRJ] = s[4 * n + c];
	for (int q = 0; q < 3; q++) {
		U[M0 + q] = m[(d + q) % 3];
		U[B0 + q] = B[(d + q) % 3];
	}
}

/* slot in Pe of the d-flux (times dt) of global field component 5 + b
   (b != d): the face-frame component q = (b - d + 3) % 3 is 1 or 2 */
__device__ __forceinline__ int pe_slot(int d, int b)
{
	return 2 * d + (b - d + 3) % 3 - 1;
}

template <int D>
__device__ __forceinline__ void pe_face(const AdjGrid& g, const double* __restrict__ s,
	double* __restrict__ Pe, int i, int j, int k)
{
	const int c = cell_at(g, i, j, k);
	const double bg[3] = {0, 0, 0};
	double Ld[8], Rd[8], F[8], sp;
	face_state(g, s, D, i - (D == 0), j - (D == 1), k - (D == 2), Ld);
	face_state(g, s, D, i, j, k, Rd);
	const bool ok = flux_hll<double>(Ld, Rd, bg, g.ph, F, sp) == ERR_OK;
	for (int q = 1; q < 3; q++) Pe[(2 * D + q - 1) * g.n + c] = g.dt * (ok ? F[B0 + q] : 0.0);
}

/* the forward step's magnetic fluxes (global frame, times dt) of every face:
   the normal component's flux is zero, the other two go to Pe; a thread
   per face (blockIdx.z = k + nz d) */
__global__ __launch_bounds__(ABX * ABY) void adj_pe_kernel(AdjGrid g, const double* __restrict__ s,
	double* __restrict__ Pe)
{
	const int i = blockIdx.x * ABX + threadIdx.x;
	const int j = blockIdx.y * ABY + threadIdx.y;
	const int k = blockIdx.z % g.nz, d = blockIdx.z / g.nz;
	if (i >= g.nx || j >= g.ny) return;
	/* d is uniform in a block: one branch per block, and inside it the
	   rotation and the tables resolve at compile time */
	if (d == 0) pe_face<0>(g, s, Pe, i, j, k);
	else if (d == 1) pe_face<1>(g, s, Pe, i, j, k);
	else pe_face<2>(g, s, Pe, i, j, k);
}

/*
Edge terms of the constrained-transport face change (the forward's order
and signs): face dim t of cell c gains sign * E_e(c + offset) / area_t, with
E_e = esign * P(pd, pq) * len_l / 4 for the six edge fields
  e: 0 EZX = +P(0,6) dz/4   1 EYX = -P(0,7) dy/4   2 EXY = +P(1,7) dx/4
     3 EZY = -P(1,5) dz/4   4 EXZ = -P(2,6) dx/4   5 EYZ = +P(2,5) dy/4
*/
struct CtTerm { signed char t, e, sign, o0, o1, o2; };
/* compile-time tables: the loops over them unroll into straight-line code */
constexpr CtTerm CT_TERMS[36] = {
	{0, 0, +1, 0, 1, 0}, {0, 0, -1, 0, -1, 0}, {0, 1, -1, 0, 0, 1}, {0, 1, +1, 0, 0, -1},
	{0, 3, +1, 0, 1, 0}, {0, 3, +1, -1, 1, 0}, {0, 3, -1, 0, 0, 0}, {0, 3, -1, -1, 0, 0},
	{0, 5, -1, 0, 0, 1}, {0, 5, -1, -1, 0, 1}, {0, 5, +1, 0, 0, 0}, {0, 5, +1, -1, 0, 0},
	{1, 2, +1, 0, 0, 1}, {1, 2, -1, 0, 0, -1}, {1, 3, -1, 1, 0, 0}, {1, 3, +1, -1, 0, 0},
	{1, 4, +1, 0, 0, 1}, {1, 4, +1, 0, -1, 1}, {1, 4, -1, 0, 0, 0}, {1, 4, -1, 0, -1, 0},
	{1, 0, -1, 1, 0, 0}, {1, 0, +1, 0, 0, 0}, {1, 0, -1, 1, -1, 0}, {1, 0, +1, 0, -1, 0},
	{2, 5, +1, 1, 0, 0}, {2, 5, -1, -1, 0, 0}, {2, 4, -1, 0, 1, 0}, {2, 4, +1, 0, -1, 0},
	{2, 1, +1, 1, 0, 0}, {2, 1, +1, 1, 0, -1}, {2, 1, -1, 0, 0, 0}, {2, 1, -1, 0, 0, -1},
	{2, 2, -1, 0, 1, 0}, {2, 2, -1, 0, 1, -1}, {2, 2, +1, 0, 0, 0}, {2, 2, +1, 0, 0, -1}};
/* per edge field: flux dim, component, sign, length dim */
constexpr signed char CT_EMF[6][4] = {
	{0, 6, +1, 2}, {0, 7, -1, 1}, {1, 7, +1, 0}, {1, 5, -1, 2}, {2, 6, -1, 0}, {2, 5, +1, 1}};

__device__ __forceinline__ double emf(const AdjGrid& g, const double* __restrict__ Pe, int e, int c)
{
	const int pd = CT_EMF[e][0], pq = CT_EMF[e][1], sg = CT_EMF[e][2], ld = CT_EMF[e][3];
	return sg * (Pe[pe_slot(pd, pq - 5) * g.n + c] * g.len[ld] / 4);
}

/* dB of the t-face owned by (i, j, k) */
__device__ double face_dB(const AdjGrid& g, const double* __restrict__ Pe, int t, int i, int j, int k)
{
	double r = 0;
#pragma unroll
	for (int q = 12 * t; q < 12 * t + 12; q++) {
		const CtTerm& T = CT_TERMS[q];
		r += T.sign * emf(g, Pe, T.e, nb(g, i, j, k, T.o0, T.o1, T.o2));
	}
	return r;
}

__device__ __forceinline__ double area(const AdjGrid& g, int t)
{
	return g.len[(t + 1) % 3] * g.len[(t + 2) % 3];
}

/* new d-face value (after CT) of the face owned by (i, j, k) */
__device__ __forceinline__ double new_face(const AdjGrid& g, const double* __restrict__ s,
	const double* __restrict__ Pe, int d, int i, int j, int k)
{
	return s[(5 + d) * g.n + cell_at(g, i, j, k)] + face_dB(g, Pe, d, i, j, k) / area(g, d);
}

/* G and mu_flux of the D-faces of cell (i, j, k) from lambda_E of the cell
   and of its lower D neighbour and lambda of the cell's D-face.  The step's
   new faces (after CT) are those of s_new, the state the step ended at,
   when it is given: the forward step sums the same edge terms in the same
   order, so they equal new_face's bit for bit */
template <int D, bool GIVEN>
__device__ __forceinline__ void g_face_vals(const AdjGrid& g, const double* __restrict__ s,
	const double* __restrict__ Pe, const double* __restrict__ s_new, double lE, double lEm,
	double lB, double* __restrict__ G, double* __restrict__ MU, int i, int j, int k)
{
	const int n = g.n;
	const int c = cell_at(g, i, j, k);
	const int di = D == 0, dj = D == 1, dk = D == 2;
	const int cp = nb(g, i, j, k, di, dj, dk);
	double f0, f1, fm;
	if constexpr (GIVEN) {
		const double* F = s_new + (5 + D) * n;
		f0 = F[c];
		f1 = F[cp];
		fm = F[nb(g, i, j, k, -di, -dj, -dk)];
	} else {
		f0 = new_face(g, s, Pe, D, i, j, k);
		f1 = new_face(g, s, Pe, D, i + di, j + dj, k + dk);
		fm = new_face(g, s, Pe, D, i - di, j - dj, k - dk);
	}
	const double mu_c = lE * (0.5 * (f0 + f1)) * g.ph.inv_mu0;
	const double mu_m = lEm * (0.5 * (fm + f0)) * g.ph.inv_mu0;
	G[D * n + c] = lB + 0.5 * mu_c + 0.5 * mu_m;
	double vb_flux = 0.5 * (s[(5 + D) * n + c] + s[(5 + D) * n + cp]);
	for (int e = 0; e < 3; e++) {
		if (e == D) continue;   /* the normal component has no flux */
		const int ce = nb(g, i, j, k, e == 0, e == 1, e == 2);
		const int sl = pe_slot(e, D);
		vb_flux += (Pe[sl * n + c] - Pe[sl * n + ce]) / g.len[e];
	}
	MU[D * n + c] = -lE * vb_flux * g.ph.inv_mu0;
}

template <int D>
__device__ __forceinline__ void g_face(const AdjGrid& g, const double* __restrict__ s,
	const double* __restrict__ Pe, const double* __restrict__ lam, double* __restrict__ G,
	double* __restrict__ MU, int i, int j, int k)
{
	const int n = g.n;
	const int c = cell_at(g, i, j, k);
	const int cm = nb(g, i, j, k, -(D == 0), -(D == 1), -(D == 2));
	g_face_vals<D, false>(g, s, Pe, nullptr, lam[4 * n + c], lam[4 * n + cm], lam[(5 + D) * n + c], G, MU,
		i, j, k);
}

/* G = lambda_face + (mu_new(c) + mu_new(c - e_d)) / 2 and
   mu_flux = -lambda_E vb_flux / mu0 of each cell, a thread per cell and
   direction d (blockIdx.z = k + nz d) */
__global__ __launch_bounds__(ABX * ABY) void adj_g_kernel(AdjGrid g, const double* __restrict__ s,
	const double* __restrict__ Pe, const double* __restrict__ lam, double* __restrict__ G,
	double* __restrict__ MU)
{
	const int i = blockIdx.x * ABX + threadIdx.x;
	const int j = blockIdx.y * ABY + threadIdx.y;
	const int k = blockIdx.z % g.nz, d = blockIdx.z / g.nz;
	if (i >= g.nx || j >= g.ny) return;
	if (d == 0) g_face<0>(g, s, Pe, lam, G, MU, i, j, k);
	else if (d == 1) g_face<1>(g, s, Pe, lam, G, MU, i, j, k);
	else g_face<2>(g, s, Pe, lam, G, MU, i, j, k);
}

/* lambda of the flux of the d-face owned by (i, j, k), global frame */
__device__ __forceinline__ void face_lambda(const AdjGrid& g, const double* __restrict__ lam,
	const double* __restrict__ MU, const double* __restrict__ G, int d, int i, int j, int k,
	double lp[8])
{
	const int n = g.n;
	const int c = cell_at(g, i, j, k);
	const int cm = nb(g, i, j, k, -(d == 0), -(d == 1), -(d == 2));
	const double il = g.inv_len[d];
	lp[RHO] = (lam[c] - lam[cm]) * il;
	for (int q = 0; q < 3; q++) lp[M0 + q] = (lam[(1 + q) * n + c] - lam[(1 + q) * n + cm]) * il;
	lp[NRJ] = (lam[4 * n + c] - lam[4 * n + cm]) * il;
	for (int q = 0; q < 3; q++) lp[B0 + q] = (MU[q * n + c] - MU[q * n + cm]) * il;
	/* transposed edge sums: E_e(c) enters face t of cell c - offset */
#pragma unroll
	for (int e = 0; e < 6; e++) {
		if (CT_EMF[e][0] != d) continue;
		double lE = 0;
#pragma unroll
		for (int q = 0; q < 36; q++) {
			const CtTerm& T = CT_TERMS[q];
			if (T.e != e) continue;
			const int tc = nb(g, i, j, k, -T.o0, -T.o1, -T.o2);
			lE += T.sign * G[T.t * n + tc] * g.inv_area[T.t];
		}
		lp[CT_EMF[e][1]] += CT_EMF[e][2] * lE * g.len[CT_EMF[e][3]] / 4;
	}
}

/* J^T (dt lambda_P) of the d-face owned by (i, j, k) for its left and right
   cells, global frame */
template <bool REVERSE>
__device__ __forceinline__ void face_jt(const AdjGrid& g, const double* __restrict__ s,
	const double lp[8], int d, int i, int j, int k, double wl[8], double wr[8])
{
	double Ld[8], Rd[8], lF[8], out[2][8];
	face_state(g, s, d, i - (d == 0), j - (d == 1), k - (d == 2), Ld);
	face_state(g, s, d, i, j, k, Rd);
	/* lambda of the face-frame flux: dt * lambda_P, rotated into the frame */
	lF[RHO] = g.dt * lp[RHO];
	lF[NRJ] = g.dt * lp[NRJ];
	for (int q = 0; q < 3; q++) {
		lF[M0 + q] = g.dt * lp[M0 + (d + q) % 3];
		lF[B0 + q] = g.dt * lp[B0 + (d + q) % 3];
	}
	if constexpr (REVERSE) {
		hll_adjoint(Ld, Rd, g.ph, lF, out[0], out[1]);
	} else {
#pragma unroll
		for (int side = 0; side < 2; side++) {
			/* J^T lambda_F for the 8 inputs of this side, JN inputs per pass */
#pragma unroll
			for (int a0 = 0; a0 < 8; a0 += JN) {
				Dual<JN> L[8], R[8], F[8];
				for (int q = 0; q < 8; q++) {
					L[q] = cst<JN>(Ld[q]);
					R[q] = cst<JN>(Rd[q]);
				}
				for (int q = 0; q < JN; q++) {
					if (side) R[a0 + q].d[q] = 1;
					else L[a0 + q].d[q] = 1;
				}
				const bool ok = hll_dual<JN>(L, R, g.ph, F);
				for (int a = 0; a < JN; a++) {
					double t = 0;
					if (ok) for (int q = 0; q < 8; q++) t += lF[q] * F[q].d[a];
					out[side][a0 + a] = t;
				}
			}
		}
	}
	/* back to the global frame: rho, m (3), E, B (3) */
	double* w[2] = {wl, wr};
	for (int side = 0; side < 2; side++) {
		w[side][RHO] = out[side][RHO];
		w[side][NRJ] = out[side][NRJ];
		for (int q = 0; q < 3; q++) {
			w[side][M0 + (d + q) % 3] = out[side][M0 + q];
			w[side][B0 + (d + q) % 3] = out[side][B0 + q];
		}
	}
}

/*
J^T g summed over both sides of a face whose two cells are one cell (the
d-faces of a grid one cell thick in d, periodic): with L = R = U the HLL
flux is the physical flux F(U), and the sum of its two side Jacobians is
F's own (the dissipation's and the wave speeds' derivatives cancel between
the sides).  Faces where hll_adjoint gives no gradient (a failing pressure
or fast speed, a zero flux) give none here either.
*/
__device__ bool self_face_adjoint(const double u[8], const Phys& ph, const double g[8], double gu[8])
{
	for (int q = 0; q < 8; q++) gu[q] = 0;
	if (u[RHO] <= 0) return false;
	const double inv_mu0 = ph.inv_mu0, gm1 = ph.gm1, hinv = ph.half_inv_mu0;
	const double msq = u[M0] * u[M0] + u[M1] * u[M1] + u[M2] * u[M2];
	const double bsq = u[B0] * u[B0] + u[B1] * u[B1] + u[B2] * u[B2];
	const double p = (u[NRJ] - 0.5 / u[RHO] * msq - hinv * bsq) * gm1;
	if (!(p > 0) || !__builtin_isnormal(p)) return false;
	const double ir = 1 / u[RHO];
	const double v[3] = {u[M0] * ir, u[M1] * ir, u[M2] * ir};
	const double mag = __builtin_sqrt(bsq);
	const double sound = __builtin_sqrt(ph.gamma * p / u[RHO]);
	double cf = sound;
	if (mag > 0) {
		const double alfven = mag / __builtin_sqrt(ph.mu0 * u[RHO]);
		const double s2 = sound * sound, a2 = alfven * alfven, ratio = u[B0] / mag;
		const double ts = s2 * s2 + a2 * a2 + 2 * s2 * a2 * (1 - 2 * (ratio * ratio));
		if (!(ts > 0)) return false;
		cf = __builtin_sqrt(0.5 * (s2 + a2 + __builtin_sqrt(ts)));
	}
	const double msn = v[0] - cf, msp = v[0] + cf;
	const double den0 = (!(msp < 0.0) ? msp : 0.0) - (!(0.0 < msn) ? msn : 0.0);
	if (!__builtin_isnormal(den0) || den0 < 0) return true;   /* zero flux */
	const double* B = u + B0;
	const double pB1 = 0.5 * inv_mu0 * bsq;
	const double vB = v[0] * B[0] + v[1] * B[1] + v[2] * B[2];
	double gv[3] = {0, 0, 0}, gp = 0, gpB1 = 0;
	gu[M0] += g[RHO];
	for (int i = 0; i < 3; i++) {
		gu[M0 + i] += g[M0 + i] * v[0];
		gv[0] += g[M0 + i] * u[M0 + i];
		gu[B0 + i] -= g[M0 + i] * B[0] * inv_mu0;
		gu[B0] -= g[M0 + i] * B[i] * inv_mu0;
	}
	gp += g[M0];
	gpB1 += g[M0];
	gv[0] += g[NRJ] * (u[NRJ] + p + pB1);
	gu[NRJ] += g[NRJ] * v[0];
	gp += g[NRJ] * v[0];
	gpB1 += g[NRJ] * v[0];
	gu[B0] -= g[NRJ] * vB * inv_mu0;
	const double gvB = -g[NRJ] * B[0] * inv_mu0;
	for (int i = 0; i < 3; i++) {
		gv[i] += gvB * B[i];
		gu[B0 + i] += gvB * v[i];
	}
	for (int i = 1; i < 3; i++) {
		gu[B0 + i] += g[B0 + i] * v[0];
		gv[0] += g[B0 + i] * B[i];
		gv[i] -= g[B0 + i] * B[0];
		gu[B0] -= g[B0 + i] * v[i];
	}
	for (int i = 0; i < 3; i++) gu[B0 + i] += gpB1 * inv_mu0 * B[i];
	for (int i = 0; i < 3; i++) {
		gu[M0 + i] += gv[i] * ir;
		gu[RHO] -= gv[i] * v[i] * ir;
	}
	const double gpg = gp * gm1;
	gu[NRJ] += gpg;
	for (int i = 0; i < 3; i++) {
		gu[M0 + i] -= gpg * v[i];
		gu[B0 + i] -= gpg * 2 * hinv * u[B0 + i];
	}
	gu[RHO] += gpg * 0.5 * msq * ir * ir;
	return true;
}

/* J^T (dt lambda_P) of the d-face owned by (i, j, k) when its two cells are
   the cell itself, global frame */
__device__ __forceinline__ void self_face_jt(const AdjGrid& g, const double* __restrict__ s,
	const double lp[8], int d, int i, int j, int k, double w[8])
{
	double U[8], lF[8], out[8];
	face_state(g, s, d, i, j, k, U);
	lF[RHO] = g.dt * lp[RHO];
	lF[NRJ] = g.dt * lp[NRJ];
	for (int q = 0; q < 3; q++) {
		lF[M0 + q] = g.dt * lp[M0 + (d + q) % 3];
		lF[B0 + q] = g.dt * lp[B0 + (d + q) % 3];
	}
	self_face_adjoint(U, g.ph, lF, out);
	w[RHO] = out[RHO];
	w[NRJ] = out[NRJ];
	for (int q = 0; q < 3; q++) {
		w[M0 + (d + q) % 3] = out[M0 + q];
		w[B0 + (d + q) % 3] = out[B0 + q];
	}
}

/* whether the grid is one cell thick in d (its d-faces join a cell to itself) */
__device__ __forceinline__ bool one_thick(const AdjGrid& g, int d)
{
	return (d == 0 ? g.nx : d == 1 ? g.ny : g.nz) == 1;
}

__device__ __forceinline__ void tile_of_cell(const AdjTiles& T, int i, int j, int k, int& ti, int& tj,
	int& tk, int& li, int& lj, int& lk)
{
	ti = i / T.tx;
	tj = j / T.ty;
	tk = k / T.tz;
	li = i - ti * T.tx;
	lj = j - tj * T.ty;
	lk = k - tk * T.tz;
}

/* boundary slot (in its tile) of the d-face owned by a cell on the tile's
   lower d layer */
__device__ __forceinline__ int bnd_slot(const AdjTiles& T, int d, int li, int lj, int lk)
{
	if (d == 0) return lj + T.ty * lk;
	if (d == 1) return T.sx + li + T.tx * lk;
	return T.sx + T.sy + li + T.tx * lj;
}

/* whether the left cell of a d-face owned by a cell at local (li, lj, lk)
   is in the same tile */
__device__ __forceinline__ bool left_in_tile(const AdjTiles& T, int d, int li, int lj, int lk)
{
	if (d == 0) return li > 0 || T.ntx == 1;
	if (d == 1) return lj > 0 || T.nty == 1;
	return lk > 0 || T.ntz == 1;
}

/* the d-face owned by cell (i, j, k) at local (li, lj, lk) of tile `tile`:
   lambda_R into the cell's accumulator (in LDS, which keeps registers for
   the Jacobian), lambda_L to the left cell's LDS slot or to the boundary
   buffer */
template <bool REVERSE, int D>
__device__ __forceinline__ void tile_face(const AdjGrid& g, const AdjTiles& T,
	const double* __restrict__ s, const double* __restrict__ lam, const double* __restrict__ MU,
	const double* __restrict__ G, double* __restrict__ BND, int tile, int i, int j, int k,
	int li, int lj, int lk, double (*acc_s)[TILE_CELLS], double (*wl_s)[8][TILE_CELLS])
{
	double lp[8], wl[8], wr[8];
	face_lambda(g, lam, MU, G, D, i, j, k, lp);
	const int t = threadIdx.x;
	if (one_thick(g, D)) {
		/* both sides are this cell: the physical flux's Jacobian, no left
		   term (its hydro lambda_P is zero; its transverse field components
		   still carry the edge sums) */
		self_face_jt(g, s, lp, D, i, j, k, wr);
		for (int q = 0; q < 8; q++) acc_s[q][t] += wr[q];
		return;
	}
	face_jt<REVERSE>(g, s, lp, D, i, j, k, wl, wr);
	for (int q = 0; q < 8; q++) acc_s[q][t] += wr[q];
	if (left_in_tile(T, D, li, lj, lk)) {
		/* the left cell's local index (wrapping inside a tile that spans the
		   dimension) */
		int mi = li - (D == 0), mj = lj - (D == 1), mk = lk - (D == 2);
		if (mi < 0) mi = g.nx - 1;
		if (mj < 0) mj = g.ny - 1;
		if (mk < 0) mk = g.nz - 1;
		const int lm = mi + T.tx * (mj + T.ty * mk);
		for (int q = 0; q < 8; q++) wl_s[D][q][lm] = wl[q];
	} else {
		double* b = BND + ((size_t)tile * T.slots + bnd_slot(T, D, li, lj, lk)) * 8;
		for (int q = 0; q < 8; q++) b[q] = wl[q];
	}
}

template <bool REVERSE>
__global__ __launch_bounds__(TILE_CELLS) void adj_tile_kernel(AdjGrid g, AdjTiles T,
	const double* __restrict__ s, const double* __restrict__ lam, const double* __restrict__ MU,
	const double* __restrict__ G, double* __restrict__ OUT, double* __restrict__ BND)
{
	__shared__ double wl_s[3][8][TILE_CELLS];
	__shared__ double acc_s[8][TILE_CELLS];
	const int t = threadIdx.x;
	const int li = t % T.tx, lj = (t / T.tx) % T.ty, lk = t / (T.tx * T.ty);
	const int i = blockIdx.x * T.tx + li, j = blockIdx.y * T.ty + lj, k = blockIdx.z * T.tz + lk;
	const bool valid = i < g.nx && j < g.ny && k < g.nz;
	const int n = g.n;
	const int tile = blockIdx.x + T.ntx * (blockIdx.y + T.nty * blockIdx.z);
	int c = 0;
	if (valid) {
		c = cell_at(g, i, j, k);
		for (int q = 0; q < 5; q++) acc_s[q][t] = lam[q * n + c];
		for (int q = 0; q < 3; q++) acc_s[5 + q][t] = MU[q * n + c];
		tile_face<REVERSE, 0>(g, T, s, lam, MU, G, BND, tile, i, j, k, li, lj, lk, acc_s, wl_s);
		/* one face at a time: no loads of the next face hoisted above this one */
		__asm__ volatile("" ::: "memory");
		tile_face<REVERSE, 1>(g, T, s, lam, MU, G, BND, tile, i, j, k, li, lj, lk, acc_s, wl_s);
		__asm__ volatile("" ::: "memory");
		tile_face<REVERSE, 2>(g, T, s, lam, MU, G, BND, tile, i, j, k, li, lj, lk, acc_s, wl_s);
	}
	__syncthreads();
	if (!valid) return;
	double acc[8];
	for (int q = 0; q < 8; q++) acc[q] = acc_s[q][t];
	for (int d = 0; d < 3; d++) {
		if (one_thick(g, d)) continue;   /* that face was summed whole above */
		/* the face owned by the right neighbour: its left cell is this one */
		int pi = li + (d == 0), pj = lj + (d == 1), pk = lk + (d == 2);
		const int ip = i + (d == 0), jp = j + (d == 1), kp = k + (d == 2);
		if (ip == g.nx) pi = 0;
		if (jp == g.ny) pj = 0;
		if (kp == g.nz) pk = 0;
		if (pi >= T.tx || pj >= T.ty || pk >= T.tz) continue;   /* in another tile */
		if ((d == 0 && ip == g.nx && T.ntx > 1) || (d == 1 && jp == g.ny && T.nty > 1)
			|| (d == 2 && kp == g.nz && T.ntz > 1)) continue;
		for (int q = 0; q < 8; q++) acc[q] += wl_s[d][q][t];
	}
	for (int q = 0; q < 8; q++) OUT[q * n + c] = acc[q];
}

/* the boundary buffer's terms of cell (i, j, k) (in range): for each d
   whose right neighbour lies in another tile (at that tile's lower d layer),
   the slot of the neighbour's d-face, else null */
__device__ __forceinline__ void bnd_terms(const AdjGrid& g, const AdjTiles& T,
	const double* __restrict__ BND, int i, int j, int k, const double* b[3])
{
	const int ip = i + 1 == g.nx ? 0 : i + 1, jp = j + 1 == g.ny ? 0 : j + 1,
		kp = k + 1 == g.nz ? 0 : k + 1;
	const int tx = i >> T.lx, ty = j >> T.ly, tz = k >> T.lz;
	const int li = i & (T.tx - 1), lj = j & (T.ty - 1), lk = k & (T.tz - 1);
	b[0] = b[1] = b[2] = nullptr;
	if (T.ntx > 1 && (ip & (T.tx - 1)) == 0) {
		const int tile = (ip >> T.lx) + T.ntx * (ty + T.nty * tz);
		b[0] = BND + ((size_t)tile * T.slots + bnd_slot(T, 0, 0, lj, lk)) * 8;
	}
	if (T.nty > 1 && (jp & (T.ty - 1)) == 0) {
		const int tile = tx + T.ntx * ((jp >> T.ly) + T.nty * tz);
		b[1] = BND + ((size_t)tile * T.slots + bnd_slot(T, 1, li, 0, lk)) * 8;
	}
	if (T.ntz > 1 && (kp & (T.tz - 1)) == 0) {
		const int tile = tx + T.ntx * (ty + T.nty * (kp >> T.lz));
		b[2] = BND + ((size_t)tile * T.slots + bnd_slot(T, 2, li, lj, 0)) * 8;
	}
}

/* component q of cell (i, j, k)'s accumulator plus its boundary terms in
   the order d = 0, 1, 2 */
__device__ __forceinline__ double final_acc(const AdjGrid& g, const AdjTiles& T,
	const double* __restrict__ OUT, const double* __restrict__ BND, int i, int j, int k, int q)
{
	const double* b[3];
	bnd_terms(g, T, BND, i, j, k, b);
	double a = OUT[q * g.n + cell_at(g, i, j, k)];
	for (int d = 0; d < 3; d++) {
		if (b[d]) a += b[d][q];
	}
	return a;
}

__global__ __launch_bounds__(ABX * ABY) void adj_final_kernel(AdjGrid g, AdjTiles T,
	double* __restrict__ lam, const double* __restrict__ G, const double* __restrict__ OUT,
	const double* __restrict__ BND)
{
	int i, j, k;
	if (!thread_cell(g, i, j, k)) return;
	const int n = g.n;
	const int c = cell_at(g, i, j, k);
	const double* b[3];
	bnd_terms(g, T, BND, i, j, k, b);
	double v[3];
	for (int q = 0; q < 8; q++) {
		double a = OUT[q * n + c];
		for (int d = 0; d < 3; d++) {
			if (b[d]) a += b[d][q];
		}
		if (q < 5) lam[q * n + c] = a;
		else v[q - 5] = a;
	}
	for (int d = 0; d < 3; d++) {
		const int im = d == 0 ? (i == 0 ? g.nx - 1 : i - 1) : i;
		const int jm = d == 1 ? (j == 0 ? g.ny - 1 : j - 1) : j;
		const int km = d == 2 ? (k == 0 ? g.nz - 1 : k - 1) : k;
		const double vm = final_acc(g, T, OUT, BND, im, jm, km, 5 + d);
		lam[(5 + d) * n + c] = G[d * n + c] + 0.5 * v[d] + 0.5 * vm;
	}
}

/*
adj_final_kernel of one reverse step followed by adj_g_kernel of the next
(the step before it in time, state s2 and magnetic fluxes Pe2): a thread
finishes lambda of its cell, writes it, and from the values it holds, with
lambda_E of its lower neighbours recomputed from their accumulators, makes
the next step's G and mu_flux of its cell.  G of a cell is read only by the
cell's own thread here, so the next step's G replaces it in place.  The
next step ended at s_new (the state this reverse step started from), whose
faces are that step's new faces.
*/
__global__ __launch_bounds__(ABX * ABY) void adj_final_g_kernel(AdjGrid g, AdjTiles T,
	double* __restrict__ lam, double* __restrict__ G, const double* __restrict__ OUT,
	const double* __restrict__ BND, const double* __restrict__ s2, const double* __restrict__ Pe2,
	const double* __restrict__ s_new, double* __restrict__ MU)
{
	int i, j, k;
	if (!thread_cell(g, i, j, k)) return;
	const int n = g.n;
	const int c = cell_at(g, i, j, k);
	const double* b[3];
	bnd_terms(g, T, BND, i, j, k, b);
	double v[8];
	for (int q = 0; q < 8; q++) {
		double a = OUT[q * n + c];
		for (int d = 0; d < 3; d++) {
			if (b[d]) a += b[d][q];
		}
		v[q] = a;
	}
	for (int q = 0; q < 5; q++) lam[q * n + c] = v[q];
	double lB[3], lEm[3];
	for (int d = 0; d < 3; d++) {
		const int im = d == 0 ? (i == 0 ? g.nx - 1 : i - 1) : i;
		const int jm = d == 1 ? (j == 0 ? g.ny - 1 : j - 1) : j;
		const int km = d == 2 ? (k == 0 ? g.nz - 1 : k - 1) : k;
		/* the lower neighbour's accumulator, components 5 + d and 4, with
		   its boundary terms in the order of final_acc */
		const double* bm[3];
		bnd_terms(g, T, BND, im, jm, km, bm);
		const int cmi = cell_at(g, im, jm, km);
		double vm = OUT[(5 + d) * n + cmi], em = OUT[4 * n + cmi];
		for (int e = 0; e < 3; e++) {
			if (bm[e]) {
				vm += bm[e][5 + d];
				em += bm[e][4];
			}
		}
		lB[d] = G[d * n + c] + 0.5 * v[5 + d] + 0.5 * vm;
		lam[(5 + d) * n + c] = lB[d];
		lEm[d] = em;
	}
	g_face_vals<0, true>(g, s2, Pe2, s_new, v[4], lEm[0], lB[0], G, MU, i, j, k);
	g_face_vals<1, true>(g, s2, Pe2, s_new, v[4], lEm[1], lB[1], G, MU, i, j, k);
	g_face_vals<2, true>(g, s2, Pe2, s_new, v[4], lEm[2], lB[2], G, MU, i, j, k);
}

} // namespace

/* scratch: G (3 n), mu_flux (3 n), then Pe (6 n) for the first two kernels
   and the accumulators (8 n) with the tiles' boundary terms for the last two */
size_t adjoint_scratch_doubles(int nx, int ny, int nz)
{
	const size_t n = (size_t)nx * ny * nz;
	const AdjTiles T = tiles_for(nx, ny, nz);
	const size_t bnd = (size_t)T.ntx * T.nty * T.ntz * T.slots * 8;
	const size_t tail = 8 * n + bnd > 6 * n ? 8 * n + bnd : 6 * n;
	return 6 * n + tail;
}

void adjoint_steps_launch(int nx, int ny, int nz, const double len[3], int nsteps, const double* dts,
	const double* const* states, const double* const* pes, const Phys& ph, double* lam,
	double* scratch, hipStream_t st, int reverse)
{
	AdjGrid g;
	g.nx = nx;
	g.ny = ny;
	g.nz = nz;
	g.n = nx * ny * nz;
	for (int d = 0; d < 3; d++) {
		g.len[d] = len[d];
		g.inv_len[d] = 1 / len[d];
		g.inv_area[d] = 1 / (len[(d + 1) % 3] * len[(d + 2) % 3]);
	}
	g.ph = ph;
	const size_t n = g.n;
	const AdjTiles T = tiles_for(nx, ny, nz);
	double* G = scratch;
	double* MU = G + 3 * n;
	double* Pe_own = MU + 3 * n;
	double* OUT = Pe_own;      /* the scratch Pe is dead once G and mu_flux are made */
	double* BND = OUT + 8 * n;
	const dim3 grid((nx + ABX - 1) / ABX, (ny + ABY - 1) / ABY, nz), block(ABX, ABY, 1);
	const dim3 fgrid((nx + ABX - 1) / ABX, (ny + ABY - 1) / ABY, 3 * nz);   /* a thread per face */
	const dim3 tgrid(T.ntx, T.nty, T.ntz);
	bool have_g = false;   /* G and mu_flux of step q made by the previous step's last kernel */
	for (int q = 0; q < nsteps; q++) {
		g.dt = dts[q];
		const double* state = states[q];
		if (!have_g) {
			/* the step's magnetic fluxes: given by the caller (the forward
			   step's own), or recomputed from the state */
			const double* Pe = pes && pes[q] ? pes[q] : Pe_own;
			if (Pe == Pe_own) adj_pe_kernel<<<fgrid, block, 0, st>>>(g, state, Pe_own);
			adj_g_kernel<<<fgrid, block, 0, st>>>(g, state, Pe, lam, G, MU);
		}
		if (reverse) adj_tile_kernel<true><<<tgrid, TILE_CELLS, 0, st>>>(g, T, state, lam, MU, G, OUT, BND);
		else adj_tile_kernel<false><<<tgrid, TILE_CELLS, 0, st>>>(g, T, state, lam, MU, G, OUT, BND);
		have_g = q + 1 < nsteps && pes && pes[q + 1];
		if (have_g) {
			adj_final_g_kernel<<<grid, block, 0, st>>>(g, T, lam, G, OUT, BND, states[q + 1], pes[q + 1],
				state, MU);
		} else {
			adj_final_kernel<<<grid, block, 0, st>>>(g, T, lam, G, OUT, BND);
		}
	}
}

void adjoint_launch(int nx, int ny, int nz, const double len[3], double dt, const Phys& ph,
	const double* state, double* lam, double* scratch, hipStream_t st, int reverse,
	const double* pe_given)
{
	adjoint_steps_launch(nx, ny, nz, len, 1, &dt, &state, &pe_given, ph, lam, scratch, st, reverse);
}

} // namespace pamhd_dev
