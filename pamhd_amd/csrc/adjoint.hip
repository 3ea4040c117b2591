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
U^{n+1}, same layout as the state) to lambda^n:
  mu_kernel      mu_new = lambda_E vb_new / mu0, mu_flux = -lambda_E vb_flux / mu0
  lface_kernel   G = lambda_face + (mu_new(c) + mu_new(c - e_d)) / 2
  lflux_kernel   lambda_P of each face from the update (conserved and vb_flux)
                 and the transposed 12-term CT sums over G
  jac_kernel     lambda_L, lambda_R = J^T (dt lambda_P): the HLL Jacobian
                 by forward-mode dual numbers, four seeded inputs per pass
                 (four passes), in the same operation order as the forward
  cell_kernel    lambda rho, m, E of each cell, lambda_vb (face-state and
                 vb_flux parts)
  face_kernel    lambda_face = G + (lambda_vb(c) + lambda_vb(c - e_d)) / 2
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
directions of the dual-number form.
*/
__device__ bool hll_adjoint(const double L[8], const double R[8], const Phys& ph,
	const double gF[8], double gL[8], double gR[8])
{
	for (int q = 0; q < 8; q++) gL[q] = gR[q] = 0;
	if (L[RHO] <= 0 || R[RHO] <= 0) return false;
	const double gamma = ph.gamma, mu0 = ph.mu0, inv_mu0 = ph.inv_mu0;
	const double gm1 = ph.gm1, hinv = ph.half_inv_mu0;
	const double* U[2] = {L, R};
	double* gU[2] = {gL, gR};
	double p[2], v[2][3], cf[2], mag[2], sound[2], alfven[2] = {0, 0}, s2[2] = {0, 0};
	double a2[2] = {0, 0}, ratio[2] = {0, 0}, rts[2] = {1, 1}, bsq[2], msq[2];
	bool hasB[2];
	for (int s = 0; s < 2; s++) {
		const double* u = U[s];
		msq[s] = u[M0] * u[M0] + u[M1] * u[M1] + u[M2] * u[M2];
		bsq[s] = u[B0] * u[B0] + u[B1] * u[B1] + u[B2] * u[B2];
		p[s] = (u[NRJ] - 0.5 / u[RHO] * msq[s] - hinv * bsq[s]) * gm1;
		if (!(p[s] > 0) || !__builtin_isnormal(p[s])) return false;
		const double inv = 1 / u[RHO];
		for (int i = 0; i < 3; i++) v[s][i] = u[M0 + i] * inv;
		mag[s] = __builtin_sqrt(bsq[s]);
		sound[s] = __builtin_sqrt(gamma * p[s] / u[RHO]);
		hasB[s] = mag[s] > 0;
		if (!hasB[s]) {
			cf[s] = sound[s];
		} else {
			alfven[s] = mag[s] / __builtin_sqrt(mu0 * u[RHO]);
			s2[s] = sound[s] * sound[s];
			a2[s] = alfven[s] * alfven[s];
			ratio[s] = u[B0] / mag[s];
			const double ts = s2[s] * s2[s] + a2[s] * a2[s]
				+ 2 * s2[s] * a2[s] * (1 - 2 * (ratio[s] * ratio[s]));
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
	const double den = bp - bm, sum = bp + bm, w = sum / den / 2;
	double gFL[8], gFR[8], gw = 0;
	for (int q = 0; q < 8; q++) {
		if (q == B0) { gFL[q] = gFR[q] = 0; continue; }
		const double fl = FF[0][q] - L[q] * bm, fr = FF[1][q] - R[q] * bp;
		gFL[q] = gF[q] * (0.5 + w);
		gFR[q] = gF[q] * (0.5 - w);
		gw += gF[q] * (fl - fr);
	}
	const double gsum = gw / (2 * den), gden = -gw * sum / (2 * den * den);
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
		const double rho = u[RHO];
		double gsound = 0, gmag = 0;
		if (!hasB[s]) {
			gsound += gcf[s];
		} else {
			const double gQ = gcf[s] / (2 * cf[s]);
			double gs2 = 0.5 * gQ, ga2 = 0.5 * gQ;
			const double gts = 0.5 * gQ / (2 * rts[s]);
			const double r2 = ratio[s] * ratio[s];
			gs2 += gts * (2 * s2[s] + 2 * a2[s] * (1 - 2 * r2));
			ga2 += gts * (2 * a2[s] + 2 * s2[s] * (1 - 2 * r2));
			const double gratio = gts * 2 * s2[s] * a2[s] * (-4 * ratio[s]);
			gsound += 2 * sound[s] * gs2;
			const double galf = 2 * alfven[s] * ga2;
			gu[B0] += gratio / mag[s];
			gmag -= gratio * u[B0] / (mag[s] * mag[s]);
			gmag += galf / __builtin_sqrt(mu0 * rho);
			gu[RHO] -= 0.5 * alfven[s] / rho * galf;
			for (int i = 0; i < 3; i++) gu[B0 + i] += gmag * u[B0 + i] / mag[s];
		}
		gp[s] += gsound * gamma / (2 * rho * sound[s]);
		gu[RHO] -= gsound * sound[s] / (2 * rho);
		for (int i = 0; i < 3; i++) {
			gu[M0 + i] += gv[s][i] / rho;
			gu[RHO] -= gv[s][i] * u[M0 + i] / (rho * rho);
		}
		gu[NRJ] += gp[s] * gm1;
		for (int i = 0; i < 3; i++) {
			gu[M0 + i] -= gp[s] * gm1 * u[M0 + i] / rho;
			gu[B0 + i] -= gp[s] * gm1 * 2 * hinv * u[B0 + i];
		}
		gu[RHO] += gp[s] * gm1 * 0.5 * msq[s] / (rho * rho);
	}
	return true;
}

struct AdjGrid {
	int nx, ny, nz, n;
	double len[3];
	double dt;
	Phys ph;
};

__device__ __forceinline__ int wrapi(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

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

/* forward fluxes P (global frame, times dt) of every face */
__global__ __launch_bounds__(ABX * ABY) void adj_flux_kernel(AdjGrid g, const double* __restrict__ s,
	double* __restrict__ P)
{
	int i, j, k;
	if (!thread_cell(g, i, j, k)) return;
	const int c = cell_at(g, i, j, k);
	for (int d = 0; d < 3; d++) {
		double Ld[8], Rd[8];
		face_state(g, s, d, i - (d == 0), j - (d == 1), k - (d == 2), Ld);
		face_state(g, s, d, i, j, k, Rd);
		Dual<1> L[8], R[8], F[8];
		for (int q = 0; q < 8; q++) {
			L[q] = cst<1>(Ld[q]);
			R[q] = cst<1>(Rd[q]);
		}
		double Fg[8] = {0, 0, 0, 0, 0, 0, 0, 0};
		if (hll_dual<1>(L, R, g.ph, F)) {
			Fg[RHO] = F[RHO].v;
			Fg[NRJ] = F[NRJ].v;
			for (int q = 0; q < 3; q++) {
				Fg[M0 + (d + q) % 3] = F[M0 + q].v;
				Fg[B0 + (d + q) % 3] = F[B0 + q].v;
			}
		}
		for (int q = 0; q < 8; q++) P[(d * 8 + q) * g.n + c] = g.dt * Fg[q];
	}
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

__device__ __forceinline__ double emf(const AdjGrid& g, const double* __restrict__ P, int e, int c)
{
	const int pd = CT_EMF[e][0], pq = CT_EMF[e][1], sg = CT_EMF[e][2], ld = CT_EMF[e][3];
	return sg * (P[(pd * 8 + pq) * g.n + c] * g.len[ld] / 4);
}

/* dB of the t-face owned by (i, j, k) */
__device__ double face_dB(const AdjGrid& g, const double* __restrict__ P, int t, int i, int j, int k)
{
	double r = 0;
#pragma unroll
	for (int q = 12 * t; q < 12 * t + 12; q++) {
		const CtTerm& T = CT_TERMS[q];
		r += T.sign * emf(g, P, T.e, nb(g, i, j, k, T.o0, T.o1, T.o2));
	}
	return r;
}

__device__ __forceinline__ double area(const AdjGrid& g, int t)
{
	return g.len[(t + 1) % 3] * g.len[(t + 2) % 3];
}

__global__ __launch_bounds__(ABX * ABY) void adj_mu_kernel(AdjGrid g, const double* __restrict__ s,
	const double* __restrict__ P, const double* __restrict__ lam, double* __restrict__ mu)
{
	int i, j, k;
	if (!thread_cell(g, i, j, k)) return;
	const int n = g.n;
	const int c = cell_at(g, i, j, k);
	const double lE = lam[4 * n + c];
#pragma unroll
	for (int d = 0; d < 3; d++) {
		const int cp = nb(g, i, j, k, d == 0, d == 1, d == 2);
		const double f0 = s[(5 + d) * n + c] + face_dB(g, P, d, i, j, k) / area(g, d);
		const double f1 = s[(5 + d) * n + cp]
			+ face_dB(g, P, d, i + (d == 0), j + (d == 1), k + (d == 2)) / area(g, d);
		const double vb_new = 0.5 * (f0 + f1);
		double vb_flux = 0.5 * (s[(5 + d) * n + c] + s[(5 + d) * n + cp]);
		for (int e = 0; e < 3; e++) {
			const int ce = nb(g, i, j, k, e == 0, e == 1, e == 2);
			vb_flux += (P[(e * 8 + 5 + d) * n + c] - P[(e * 8 + 5 + d) * n + ce]) / g.len[e];
		}
		mu[d * n + c] = lE * vb_new * g.ph.inv_mu0;
		mu[(3 + d) * n + c] = -lE * vb_flux * g.ph.inv_mu0;
	}
}

__global__ __launch_bounds__(ABX * ABY) void adj_lface_kernel(AdjGrid g, const double* __restrict__ lam,
	const double* __restrict__ mu, double* __restrict__ G)
{
	int i, j, k;
	if (!thread_cell(g, i, j, k)) return;
	const int n = g.n;
	const int c = cell_at(g, i, j, k);
	for (int d = 0; d < 3; d++) {
		const int cm = nb(g, i, j, k, -(d == 0), -(d == 1), -(d == 2));
		G[d * n + c] = lam[(5 + d) * n + c] + 0.5 * mu[d * n + c] + 0.5 * mu[d * n + cm];
	}
}

/* lambda_P of the three faces owned by the thread's cell */
__global__ __launch_bounds__(ABX * ABY) void adj_lflux_kernel(AdjGrid g, const double* __restrict__ lam,
	const double* __restrict__ mu, const double* __restrict__ G, double* __restrict__ LP)
{
	int i, j, k;
	if (!thread_cell(g, i, j, k)) return;
	const int n = g.n;
	const int c = cell_at(g, i, j, k);
#pragma unroll
	for (int d = 0; d < 3; d++) {
		const int cm = nb(g, i, j, k, -(d == 0), -(d == 1), -(d == 2));
		const double il = 1 / g.len[d];
		double lp[8];
		lp[RHO] = (lam[c] - lam[cm]) * il;
		for (int q = 0; q < 3; q++) lp[M0 + q] = (lam[(1 + q) * n + c] - lam[(1 + q) * n + cm]) * il;
		lp[NRJ] = (lam[4 * n + c] - lam[4 * n + cm]) * il;
		for (int q = 0; q < 3; q++) lp[B0 + q] = (mu[(3 + q) * n + c] - mu[(3 + q) * n + cm]) * il;
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
				lE += T.sign * G[T.t * n + tc] / area(g, T.t);
			}
			lp[CT_EMF[e][1]] += CT_EMF[e][2] * lE * g.len[CT_EMF[e][3]] / 4;
		}
		for (int q = 0; q < 8; q++) LP[(d * 8 + q) * n + c] = lp[q];
	}
}

/* lambda_L, lambda_R (global frame) of the faces owned by the cell */
__global__ __launch_bounds__(ABX * ABY) void adj_jac_kernel(AdjGrid g, const double* __restrict__ s,
	const double* __restrict__ LP, double* __restrict__ W)
{
	int i, j, k;
	if (!thread_cell(g, i, j, k)) return;
	const int n = g.n;
	const int c = cell_at(g, i, j, k);
	for (int d = 0; d < 3; d++) {
		double Ld[8], Rd[8], lF[8];
		face_state(g, s, d, i - (d == 0), j - (d == 1), k - (d == 2), Ld);
		face_state(g, s, d, i, j, k, Rd);
		/* lambda of the face-frame flux: dt * lambda_P, rotated into the frame */
		lF[RHO] = g.dt * LP[(d * 8 + RHO) * n + c];
		lF[NRJ] = g.dt * LP[(d * 8 + NRJ) * n + c];
		for (int q = 0; q < 3; q++) {
			lF[M0 + q] = g.dt * LP[(d * 8 + M0 + (d + q) % 3) * n + c];
			lF[B0 + q] = g.dt * LP[(d * 8 + B0 + (d + q) % 3) * n + c];
		}
#pragma unroll
		for (int side = 0; side < 2; side++) {
			/* J^T lambda_F for the 8 inputs of this side, JN inputs per pass */
			double out[8] = {0, 0, 0, 0, 0, 0, 0, 0};
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
				if (hll_dual<JN>(L, R, g.ph, F)) {
					for (int a = 0; a < JN; a++) {
						double t = 0;
						for (int q = 0; q < 8; q++) t += lF[q] * F[q].d[a];
						out[a0 + a] = t;
					}
				}
			}
			/* back to the global frame: rho, m (3), E, B (3) */
			double* w = W + (size_t)(d * 16 + side * 8) * n;
			w[RHO * n + c] = out[RHO];
			w[NRJ * n + c] = out[NRJ];
			for (int q = 0; q < 3; q++) {
				w[(M0 + (d + q) % 3) * n + c] = out[M0 + q];
				w[(B0 + (d + q) % 3) * n + c] = out[B0 + q];
			}
		}
	}
}

/* the same W by hll_adjoint */
__global__ __launch_bounds__(ABX * ABY) void adj_jacr_kernel(AdjGrid g, const double* __restrict__ s,
	const double* __restrict__ LP, double* __restrict__ W)
{
	int i, j, k;
	if (!thread_cell(g, i, j, k)) return;
	const int n = g.n;
	const int c = cell_at(g, i, j, k);
	for (int d = 0; d < 3; d++) {
		double Ld[8], Rd[8], lF[8], out[2][8];
		face_state(g, s, d, i - (d == 0), j - (d == 1), k - (d == 2), Ld);
		face_state(g, s, d, i, j, k, Rd);
		lF[RHO] = g.dt * LP[(d * 8 + RHO) * n + c];
		lF[NRJ] = g.dt * LP[(d * 8 + NRJ) * n + c];
		for (int q = 0; q < 3; q++) {
			lF[M0 + q] = g.dt * LP[(d * 8 + M0 + (d + q) % 3) * n + c];
			lF[B0 + q] = g.dt * LP[(d * 8 + B0 + (d + q) % 3) * n + c];
		}
		hll_adjoint(Ld, Rd, g.ph, lF, out[0], out[1]);
		for (int side = 0; side < 2; side++) {
			double* w = W + (size_t)(d * 16 + side * 8) * n;
			w[RHO * n + c] = out[side][RHO];
			w[NRJ * n + c] = out[side][NRJ];
			for (int q = 0; q < 3; q++) {
				w[(M0 + (d + q) % 3) * n + c] = out[side][M0 + q];
				w[(B0 + (d + q) % 3) * n + c] = out[side][B0 + q];
			}
		}
	}
}

/* lambda^n of rho, m, E (in place) and lambda_vb of each cell */
__global__ __launch_bounds__(ABX * ABY) void adj_cell_kernel(AdjGrid g, double* __restrict__ lam,
	const double* __restrict__ mu, const double* __restrict__ W, double* __restrict__ V)
{
	int i, j, k;
	if (!thread_cell(g, i, j, k)) return;
	const int n = g.n;
	const int c = cell_at(g, i, j, k);
	double acc[8];
	for (int q = 0; q < 5; q++) acc[q] = lam[q * n + c];
	for (int q = 0; q < 3; q++) acc[5 + q] = mu[(3 + q) * n + c];
	for (int d = 0; d < 3; d++) {
		const int cp = nb(g, i, j, k, d == 0, d == 1, d == 2);
		const double* wr = W + (size_t)(d * 16 + 8) * n;   /* the cell is R of its own face */
		const double* wl = W + (size_t)(d * 16) * n;       /* and L of the face owned by cp */
		for (int q = 0; q < 8; q++) acc[q] += wr[q * n + c] + wl[q * n + cp];
	}
	for (int q = 0; q < 5; q++) lam[q * n + c] = acc[q];
	for (int q = 0; q < 3; q++) V[q * n + c] = acc[5 + q];
}

__global__ __launch_bounds__(ABX * ABY) void adj_face_kernel(AdjGrid g, double* __restrict__ lam,
	const double* __restrict__ G, const double* __restrict__ V)
{
	int i, j, k;
	if (!thread_cell(g, i, j, k)) return;
	const int n = g.n;
	const int c = cell_at(g, i, j, k);
	for (int d = 0; d < 3; d++) {
		const int cm = nb(g, i, j, k, -(d == 0), -(d == 1), -(d == 2));
		lam[(5 + d) * n + c] = G[d * n + c] + 0.5 * V[d * n + c] + 0.5 * V[d * n + cm];
	}
}

} // namespace

/* scratch: 24 n (P) + 6 n (mu) + 3 n (G) + 24 n (lambda_P) + 48 n (W) + 3 n (V) */
size_t adjoint_scratch_doubles(long long n) { return (size_t)108 * n; }

void adjoint_launch(int nx, int ny, int nz, const double len[3], double dt, const Phys& ph,
	const double* state, double* lam, double* scratch, hipStream_t st, int reverse)
{
	AdjGrid g;
	g.nx = nx;
	g.ny = ny;
	g.nz = nz;
	g.n = nx * ny * nz;
	for (int d = 0; d < 3; d++) g.len[d] = len[d];
	g.dt = dt;
	g.ph = ph;
	const size_t n = g.n;
	double* P = scratch;
	double* mu = P + 24 * n;
	double* G = mu + 6 * n;
	double* LP = G + 3 * n;
	double* W = LP + 24 * n;
	double* V = W + 48 * n;
	const dim3 grid((nx + ABX - 1) / ABX, (ny + ABY - 1) / ABY, nz), block(ABX, ABY, 1);
	adj_flux_kernel<<<grid, block, 0, st>>>(g, state, P);
	adj_mu_kernel<<<grid, block, 0, st>>>(g, state, P, lam, mu);
	adj_lface_kernel<<<grid, block, 0, st>>>(g, lam, mu, G);
	adj_lflux_kernel<<<grid, block, 0, st>>>(g, lam, mu, G, LP);
	if (reverse) adj_jacr_kernel<<<grid, block, 0, st>>>(g, state, LP, W);
	else adj_jac_kernel<<<grid, block, 0, st>>>(g, state, LP, W);
	adj_cell_kernel<<<grid, block, 0, st>>>(g, lam, mu, W, V);
	adj_face_kernel<<<grid, block, 0, st>>>(g, lam, G, V);
}

} // namespace pamhd_dev
