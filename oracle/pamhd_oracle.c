/*
FP64 CPU oracle for the PAMHD ideal-MHD finite-volume step.

TEST INFRASTRUCTURE ONLY: used by tests/, the smoke check and the CPU
baseline of bench.py as the checker; never linked into the product.

Build with -O2 -ffp-contract=off (no fused multiply-add) so that every
expression rounds exactly as the reference's x86-64 build does.  The
operation order of each expression follows the cited reference line;
where the reference writes pamhd::add(a, pamhd::neg(b)) this file writes
a - b (IEEE subtraction is defined as a + (-b)), and pamhd::mul(s, v) is
v[i] * s (multiplication commutes exactly).
*/
#include "pamhd_oracle.h"

#include <float.h>
#include <limits.h>
#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <stdlib.h>
#include <string.h>

/* std::max / std::min semantics (the first argument wins ties and NaNs) */
static inline double smax(double a, double b) { return (a < b) ? b : a; }
static inline double smin(double a, double b) { return (b < a) ? b : a; }

enum { RHO = 0, M0 = 1, M1 = 2, M2 = 3, NRJ = 4, B0 = 5, B1 = 6, B2 = 7 };

/* pamhd::dot / norm2 of common_functions.hpp:143-163: (a0b0 + a1b1) + a2b2 */
static inline double dot3(const double* a, const double* b)
{
	return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

/* ---------------------------------------------------------------------- */
/* physics helpers                                                        */
/* ---------------------------------------------------------------------- */

/* source/mhd/common.hpp:76-115 */
/* number of faces the last step solved: the third return value of
   timestep(), source/mhd/solve.hpp:253-313, :1918 */
static unsigned long long or_fc = 0;

unsigned long long or_last_flux_calcs(void)
{
	return or_fc;
}

int or_get_pressure(const double U[8], double gamma, double mu0, double* p)
{
	if (U[RHO] <= 0) return OR_ERR_DENSITY;
	const double
		kinetic = 0.5 / U[RHO] * (U[M0] * U[M0] + U[M1] * U[M1] + U[M2] * U[M2]),
		magnetic = 0.5 / mu0 * (U[B0] * U[B0] + U[B1] * U[B1] + U[B2] * U[B2]);
	*p = (U[NRJ] - kinetic - magnetic) * (gamma - 1);
	return OR_OK;
}

/* source/mhd/common.hpp:254-311 (argument checks made by callers) */
double or_total_energy(double rho, const double v[3], double p, const double B[3],
	double gamma, double mu0)
{
	const double
		kinetic = 0.5 * rho * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]),
		magnetic = 0.5 / mu0 * (B[0] * B[0] + B[1] * B[1] + B[2] * B[2]);
	return p / (gamma - 1) + kinetic + magnetic;
}

/* source/mhd/common.hpp:324-368 */
static int sound_speed(const double U[8], double gamma, double mu0, double* cs)
{
	if (!isnormal(U[RHO]) || U[RHO] < 0) return OR_ERR_DENSITY;
	double p;
	int e = or_get_pressure(U, gamma, mu0, &p);
	if (e) return e;
	if (!isnormal(p) || p < 0) return OR_ERR_PRESSURE;
	*cs = sqrt(gamma * p / U[RHO]);
	return OR_OK;
}

/* source/mhd/common.hpp:415-491 with get_alfven_speed of :378-405 */
int or_fast_speed(const double U[8], const double bg[3], double gamma, double mu0,
	double* cf)
{
	const double t[3] = {U[B0] + bg[0], U[B1] + bg[1], U[B2] + bg[2]};
	if (!isfinite(t[0])) return OR_ERR_FIELD;
	const double mag = sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
	double sound;
	int e = sound_speed(U, gamma, mu0, &sound);
	if (e) return e;
	if (mag <= 0) {
		*cf = sound;
		return OR_OK;
	}
	/* get_alfven_speed: |B + B0| / sqrt(mu0 rho); its own density check
	   cannot fire here because sound_speed made the same check */
	const double alfven = mag / sqrt(mu0 * U[RHO]);
	const double
		sound2 = sound * sound,
		alfven2 = alfven * alfven,
		speeds2 = sound2 + alfven2;
	if (!isnormal(speeds2) || speeds2 < 0) return OR_ERR_SPEED;
	const double ratio = t[0] / mag;
	const double to_sqrt
		= sound2 * sound2 + alfven2 * alfven2
		+ 2 * sound2 * alfven2 * (1 - 2 * (ratio * ratio));
	if (!isnormal(to_sqrt) || to_sqrt < 0) return OR_ERR_SPEED;
	*cf = sqrt(0.5 * (speeds2 + sqrt(to_sqrt)));
	return OR_OK;
}

/* get_velocity, source/mhd/common.hpp:118-130 */
static inline void velocity(const double U[8], double v[3])
{
	if (U[RHO] > 0) {
		const double inv = 1 / U[RHO];
		v[0] = U[M0] * inv;
		v[1] = U[M1] * inv;
		v[2] = U[M2] * inv;
	} else {
		v[0] = v[1] = v[2] = 0;
	}
}

/* ideal-MHD x-flux with background-field splitting, source/mhd/common.hpp:138-244 */
int or_physical_flux(const double U[8], const double bg[3], double gamma,
	double mu0, double F[8])
{
	if (!isnormal(U[RHO]) || U[RHO] <= 0) return OR_ERR_DENSITY;
	if (!isfinite(U[M0])) return OR_ERR_FLUX;

	const double* B = U + B0;
	const double t[3] = {B[0] + bg[0], B[1] + bg[1], B[2] + bg[2]};
	const double
		inv = 1.0 / mu0,
		pB0 = 0.5 * inv * (bg[0] * bg[0] + bg[1] * bg[1] + bg[2] * bg[2]),
		pB1 = 0.5 * inv * (B[0] * B[0] + B[1] * B[1] + B[2] * B[2]),
		pBt = 0.5 * inv * (t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
	double pth;
	int e = or_get_pressure(U, gamma, mu0, &pth);
	if (e) return e;
	double v[3];
	velocity(U, v);

	F[RHO] = U[M0];
	if (!isfinite(F[RHO])) return OR_ERR_FLUX;

	/* add(mul(m, v0), mul(-inv, add(mul(t0, t), mul(-bg0, bg)))) */
	for (int i = 0; i < 3; i++) {
		F[M0 + i] = U[M0 + i] * v[0] + (t[i] * t[0] + bg[i] * (-bg[0])) * (-inv);
	}
	F[M0] += pth + pBt - pB0;

	F[NRJ]
		= v[0] * (U[NRJ] + pth + pB1)
		- B[0] * dot3(v, B) * inv
		+ inv * (
			B[1] * (v[0] * bg[1] - v[1] * bg[0])
			+ B[2] * (v[0] * bg[2] - v[2] * bg[0]));

	/* add(mul(v0, t), mul(-t0, v)) */
	for (int i = 0; i < 3; i++) {
		F[B0 + i] = t[i] * v[0] + v[i] * (-t[0]);
	}
	F[B0] = 0;
	return OR_OK;
}

/* ---------------------------------------------------------------------- */
/* Riemann solvers                                                        */
/* ---------------------------------------------------------------------- */

/* source/mhd/rusanov.hpp:59-214 */
int or_flux_rusanov(const double L[8], const double R[8], const double bg[3],
	double gamma, double mu0, double F[8], double* s_out)
{
	if (!isnormal(L[RHO]) || L[RHO] < 0) return OR_ERR_DENSITY;
	if (!isnormal(R[RHO]) || R[RHO] < 0) return OR_ERR_DENSITY;
	double cfL, cfR;
	int e = or_fast_speed(L, bg, gamma, mu0, &cfL);
	if (e) return e;
	e = or_fast_speed(R, bg, gamma, mu0, &cfR);
	if (e) return e;
	const double
		max_fast = smax(cfL, cfR),
		s = smax(fabs(L[M0] / L[RHO]) + max_fast, fabs(R[M0] / R[RHO]) + max_fast);
	if (!isfinite(max_fast) || !isfinite(s)) return OR_ERR_SPEED;

	double FL[8], FR[8];
	e = or_physical_flux(L, bg, gamma, mu0, FL);
	if (e) return e;
	e = or_physical_flux(R, bg, gamma, mu0, FR);
	if (e) return e;
	if (!isfinite(FL[RHO]) || !isfinite(FR[RHO])) return OR_ERR_FLUX;

	F[RHO] = (FL[RHO] + FR[RHO]) / 2 - (R[RHO] - L[RHO]) * (s / 2);
	F[NRJ] = (FL[NRJ] + FR[NRJ]) / 2 - (R[NRJ] - L[NRJ]) * (s / 2);
	const double hs = -s / 2;
	for (int i = 0; i < 3; i++) {
		F[M0 + i] = (FL[M0 + i] + FR[M0 + i]) * 0.5 + (R[M0 + i] - L[M0 + i]) * hs;
		F[B0 + i] = (FL[B0 + i] + FR[B0 + i]) * 0.5 + (R[B0 + i] - L[B0 + i]) * hs;
	}
	if (!isfinite(F[RHO])) return OR_ERR_FLUX;
	*s_out = s;
	return OR_OK;
}

/* source/mhd/hll_athena.hpp:57-275 */
int or_flux_hll(const double L[8], const double R[8], const double bg[3],
	double gamma, double mu0, double F[8], double* s_out)
{
	double vL[3], vR[3];
	velocity(L, vL);
	velocity(R, vR);
	double pL, pR;
	int e = or_get_pressure(L, gamma, mu0, &pL);
	if (e) return e;
	e = or_get_pressure(R, gamma, mu0, &pR);
	if (e) return e;
	const double
		pmL = dot3(L + B0, L + B0) / (2 * mu0),
		pmR = dot3(R + B0, R + B0) / (2 * mu0);
	double cfL, cfR;
	e = or_fast_speed(L, bg, gamma, mu0, &cfL);
	if (e) return e;
	e = or_fast_speed(R, bg, gamma, mu0, &cfR);
	if (e) return e;
	const double
		ms = smax(cfL, cfR),
		msn = (vL[0] <= vR[0]) ? vL[0] - ms : vR[0] - ms,
		msp = (vL[0] <= vR[0]) ? vR[0] + ms : vL[0] + ms,
		bm = smin(msn, 0.0),
		bp = smax(msp, 0.0);

	if (!isnormal(pL) || pL < 0) return OR_ERR_PRESSURE;
	if (!isfinite(pmL) || pmL < 0) return OR_ERR_PRESSURE;
	if (!isfinite(msn)) return OR_ERR_SPEED;
	if (!isnormal(pR) || pR < 0) return OR_ERR_PRESSURE;
	if (!isfinite(pmR) || pmR < 0) return OR_ERR_PRESSURE;
	if (!isfinite(msp)) return OR_ERR_SPEED;

	if (!isnormal(bp - bm) || bp - bm < 0) {
		for (int i = 0; i < 8; i++) F[i] = 0;
		*s_out = 0;
		return OR_OK;
	}

	double FL[8], FR[8];
	e = or_physical_flux(L, bg, gamma, mu0, FL);
	if (e) return e;
	e = or_physical_flux(R, bg, gamma, mu0, FR);
	if (e) return e;

	FL[RHO] -= L[RHO] * bm;
	FL[NRJ] -= L[NRJ] * bm;
	FR[RHO] -= R[RHO] * bp;
	FR[NRJ] -= R[NRJ] * bp;
	for (int i = 0; i < 3; i++) {
		FL[M0 + i] = FL[M0 + i] + L[M0 + i] * (-bm);
		FL[B0 + i] = FL[B0 + i] + L[B0 + i] * (-bm);
		FR[M0 + i] = FR[M0 + i] + R[M0 + i] * (-bp);
		FR[B0 + i] = FR[B0 + i] + R[B0 + i] * (-bp);
	}
	FR[B0] = FL[B0] = 0;

	F[RHO] = (FL[RHO] + FR[RHO]) / 2 + (FL[RHO] - FR[RHO]) * (bp + bm) / (bp - bm) / 2.0;
	F[NRJ] = (FL[NRJ] + FR[NRJ]) / 2 + (FL[NRJ] - FR[NRJ]) * (bp + bm) / (bp - bm) / 2.0;
	const double w = (bp + bm) / (bp - bm) / 2.0;
	for (int i = 0; i < 3; i++) {
		F[M0 + i] = (FL[M0 + i] + FR[M0 + i]) * 0.5 + (FL[M0 + i] - FR[M0 + i]) * w;
		F[B0 + i] = (FL[B0 + i] + FR[B0 + i]) * 0.5 + (FL[B0 + i] - FR[B0 + i]) * w;
	}
	*s_out = smax(fabs(bp), fabs(bm));
	return OR_OK;
}

/* One side's star state of source/mhd/hlld_athena.hpp:316-401 (eqns 43-48). */
static void hlld_star(const double U[8], const double v[3], double p, double nm,
	double sfd, double smd, double sm, double ds, double Bx, double Bx2,
	double inv, double ptst, double S[8], double* vdbs)
{
	const double eps = DBL_EPSILON;
	const double tmp = U[RHO] * sfd * smd - Bx2 * inv;
	if (fabs(tmp) < eps * ptst) {
		for (int i = 0; i < 3; i++) {
			S[M0 + i] = v[i] * ds;
			S[B0 + i] = U[B0 + i];
		}
	} else {
		const double tmp1 = Bx * (sm - v[0]) * inv / tmp;
		for (int i = 0; i < 3; i++) {
			S[M0 + i] = (v[i] + U[B0 + i] * (-tmp1)) * ds;
		}
		const double tmp2 = (U[RHO] * (sfd * sfd) - Bx2 * inv) / tmp;
		for (int i = 0; i < 3; i++) {
			S[B0 + i] = U[B0 + i] * tmp2;
		}
	}
	S[RHO] = ds;
	S[M0] = ds * sm;
	S[B0] = Bx;
	*vdbs = dot3(S + M0, S + B0) / ds;
	S[NRJ]
		= (
			sfd * U[NRJ]
			- (p + nm) * v[0]
			+ ptst * sm
			+ Bx * inv * (dot3(v, U + B0) - *vdbs)
		) / smd;
}

/* source/mhd/hlld_athena.hpp:54-559 (ignores the background field, :77) */
int or_flux_hlld(const double L[8], const double R[8], const double bg_unused[3],
	double gamma, double mu0, double F[8], double* s_out)
{
	(void)bg_unused;
	const double zero[3] = {0, 0, 0};
	const double eps = DBL_EPSILON;

	if (L[RHO] <= 0) return OR_ERR_DENSITY;
	if (R[RHO] <= 0) return OR_ERR_DENSITY;
	double pL, pR;
	int e = or_get_pressure(L, gamma, mu0, &pL);
	if (e) return e;
	if (pL <= 0) return OR_ERR_PRESSURE;
	e = or_get_pressure(R, gamma, mu0, &pR);
	if (e) return e;
	if (pR <= 0) return OR_ERR_PRESSURE;

	const double
		inv = 1.0 / mu0,
		inv_sqrt = 1.0 / sqrt(mu0),
		Bx = (L[B0] + R[B0]) / 2,
		Bx2 = Bx * Bx;
	double cfL, cfR;
	e = or_fast_speed(L, zero, gamma, mu0, &cfL);
	if (e) return e;
	e = or_fast_speed(R, zero, gamma, mu0, &cfR);
	if (e) return e;
	const double ms = smax(cfL, cfR);

	double vL[3], vR[3];
	{
		const double iL = 1 / L[RHO], iR = 1 / R[RHO];
		for (int i = 0; i < 3; i++) {
			vL[i] = L[M0 + i] * iL;
			vR[i] = R[M0 + i] * iR;
		}
	}
	const double
		msn = (vL[0] <= vR[0]) ? vL[0] - ms : vR[0] - ms,
		msp = (vL[0] <= vR[0]) ? vR[0] + ms : vL[0] + ms;

	double FL[8], FR[8];
	e = or_physical_flux(L, zero, gamma, mu0, FL);
	if (e) return e;
	e = or_physical_flux(R, zero, gamma, mu0, FR);
	if (e) return e;

	if (msn >= 0.0) {
		memcpy(F, FL, sizeof(FL));
		*s_out = fabs(msp);
		return OR_OK;
	} else if (msp <= 0.0) {
		memcpy(F, FR, sizeof(FR));
		*s_out = fabs(msn);
		return OR_OK;
	}

	const double
		sfdn = msn - vL[0],
		sfdp = msp - vR[0],
		nmn = dot3(L + B0, L + B0) * (inv / 2),
		nmp = dot3(R + B0, R + B0) * (inv / 2),
		/* eqn 38 */
		sm
			= (
				sfdp * R[M0]
				- sfdn * L[M0]
				- (pR + nmp)
				+ (pL + nmn)
			) / (
				sfdp * R[RHO]
				- sfdn * L[RHO]
			),
		smdn = msn - sm,
		smdp = msp - sm,
		/* eqn 43 */
		dsn = L[RHO] * sfdn / smdn,
		dssn = sqrt(dsn),
		dsp = R[RHO] * sfdp / smdp,
		dssp = sqrt(dsp),
		/* eqn 51 */
		ssn = sm - fabs(Bx) / sqrt(dsn * mu0),
		ssp = sm + fabs(Bx) / sqrt(dsp * mu0),
		ptst = pL + nmn + L[RHO] * sfdn * (sm - vL[0]);

	double Sn[8], Sp[8], vdbsn, vdbsp;
	hlld_star(L, vL, pL, nmn, sfdn, smdn, sm, dsn, Bx, Bx2, inv, ptst, Sn, &vdbsn);
	hlld_star(R, vR, pR, nmp, sfdp, smdp, sm, dsp, Bx, Bx2, inv, ptst, Sp, &vdbsp);

	double vsn[3], vsp[3];
	{
		const double in = 1 / dsn, ip = 1 / dsp;
		for (int i = 0; i < 3; i++) {
			vsn[i] = Sn[M0 + i] * in;
			vsp[i] = Sp[M0 + i] * ip;
		}
	}

	double S2n[8], S2p[8];
	if (Bx2 / 2 < eps * ptst) {
		memcpy(S2n, Sn, sizeof(Sn));
		memcpy(S2p, Sp, sizeof(Sp));
	} else {
		const double isds = 1.0 / (dssn + dssp);
		const double sgn = (Bx > 0) ? 1.0 : -1.0;
		S2n[RHO] = Sn[RHO];
		S2p[RHO] = Sp[RHO];
		double v2[3];
		const double c1 = sgn * inv_sqrt;
		for (int i = 0; i < 3; i++) {
			v2[i] = ((vsn[i] * dssn + vsp[i] * dssp) + (Sp[B0 + i] - Sn[B0 + i]) * c1) * isds;
		}
		v2[0] = sm;
		/* eqs 59 & 60 */
		for (int i = 0; i < 3; i++) {
			S2n[M0 + i] = v2[i] * Sn[RHO];
			S2p[M0 + i] = v2[i] * Sp[RHO];
		}
		S2n[M0] = Sn[M0];
		S2p[M0] = Sp[M0];
		/* eqs 61 & 62 */
		const double c2 = sgn * dssn * dssp * sqrt(mu0);
		for (int i = 0; i < 3; i++) {
			const double b = ((Sp[B0 + i] * dssn + Sn[B0 + i] * dssp) + (vsp[i] - vsn[i]) * c2) * isds;
			S2n[B0 + i] = b;
			S2p[B0 + i] = b;
		}
		S2n[B0] = S2p[B0] = Bx;
		/* eqn 63 */
		const double vdb2 = dot3(v2, S2n + B0);
		S2n[NRJ] = Sn[NRJ] - dssn * sgn * (vdbsn - vdb2) * inv_sqrt;
		S2p[NRJ] = Sp[NRJ] + dssp * sgn * (vdbsp - vdb2) * inv_sqrt;
	}

	if (ssn >= 0) {
		F[RHO] = FL[RHO] + (Sn[RHO] - L[RHO]) * msn;
		F[NRJ] = FL[NRJ] + (Sn[NRJ] - L[NRJ]) * msn;
		for (int i = 0; i < 3; i++) {
			F[M0 + i] = FL[M0 + i] + (Sn[M0 + i] - L[M0 + i]) * msn;
			F[B0 + i] = FL[B0 + i] + (Sn[B0 + i] - L[B0 + i]) * msn;
		}
	} else if (sm >= 0) {
		const double d = ssn - msn;
		for (int q = 0; q < 8; q++) {
			F[q] = FL[q] + S2n[q] * ssn - Sn[q] * d - L[q] * msn;
		}
	} else if (ssp > 0) {
		const double d = ssp - msp;
		for (int q = 0; q < 8; q++) {
			F[q] = FR[q] + S2p[q] * ssp - Sp[q] * d - R[q] * msp;
		}
	} else {
		F[RHO] = FR[RHO] + (Sp[RHO] - R[RHO]) * msp;
		F[NRJ] = FR[NRJ] + (Sp[NRJ] - R[NRJ]) * msp;
		for (int i = 0; i < 3; i++) {
			F[M0 + i] = FR[M0 + i] + (Sp[M0 + i] - R[M0 + i]) * msp;
			F[B0 + i] = FR[B0 + i] + (Sp[B0 + i] - R[B0 + i]) * msp;
		}
	}
	F[B0] = 0;
	*s_out = smax(fabs(msn), fabs(msp));
	return OR_OK;
}

/* Roe eigensystem, source/mhd/roe_athena.hpp:56-296 (eqs. B17-B29 of the
   Athena method paper); rem holds right eigenvectors as columns, lem left
   eigenvectors as rows. */
void or_roe_eigensystem(double d, double v1, double v2, double v3, double h,
	double b1, double b2, double b3, double x, double y, double gamma,
	double ev[7], double rem[7][7], double lem[7][7])
{
	const double tiny = DBL_EPSILON;
	const double g1 = gamma - 1, g2 = gamma - 2;

	const double di = 1.0 / d;
	const double vsq = v1 * v1 + v2 * v2 + v3 * v3;
	const double btsq = b2 * b2 + b3 * b3;
	const double bt_starsq = (g1 - g2 * y) * btsq;
	const double vaxsq = b1 * b1 * di;
	const double hp = h - (vaxsq + btsq * di);
	const double twid_asq = smax((g1 * (hp - 0.5 * vsq) - g2 * x), tiny);

	/* fast and slow magnetosonic speeds, eq. B18 */
	const double ct2 = bt_starsq * di;
	const double tsum = vaxsq + ct2 + twid_asq;
	const double tdif = vaxsq + ct2 - twid_asq;
	const double cf2_cs2 = sqrt(tdif * tdif + 4.0 * twid_asq * ct2);
	const double cfsq = 0.5 * (tsum + cf2_cs2);
	const double cf = sqrt(cfsq);
	const double cssq = twid_asq * vaxsq / cfsq;
	const double cs = sqrt(cssq);

	/* beta(s), eqs. A17, B20, B28 */
	const double bt = sqrt(btsq);
	const double bt_star = sqrt(bt_starsq);
	double bet2, bet3;
	if (bt == 0.0) {
		bet2 = 1.0;
		bet3 = 0.0;
	} else {
		bet2 = b2 / bt;
		bet3 = b3 / bt;
	}
	const double bet2_star = bet2 / sqrt(g1 - g2 * y);
	const double bet3_star = bet3 / sqrt(g1 - g2 * y);
	const double bet_starsq = bet2_star * bet2_star + bet3_star * bet3_star;
	const double vbet = v2 * bet2_star + v3 * bet3_star;

	/* alpha(s), eq. A16 */
	double alpha_f, alpha_s;
	if ((cfsq - cssq) == 0.0) {
		alpha_f = 1.0;
		alpha_s = 0.0;
	} else if ((twid_asq - cssq) <= 0.0) {
		alpha_f = 0.0;
		alpha_s = 1.0;
	} else if ((cfsq - twid_asq) <= 0.0) {
		alpha_f = 1.0;
		alpha_s = 0.0;
	} else {
		alpha_f = sqrt((twid_asq - cssq) / (cfsq - cssq));
		alpha_s = sqrt((cfsq - twid_asq) / (cfsq - cssq));
	}

	/* Q(s) and A(s), eqs. A14-15 */
	const double sqrtd = sqrt(d);
	const double isqrtd = 1.0 / sqrtd;
	const double s = (b1 > 0) ? 1 : -1;
	const double twid_a = sqrt(twid_asq);
	double qf = cf * alpha_f * s;
	double qs = cs * alpha_s * s;
	const double af_prime = twid_a * alpha_f * isqrtd;
	const double as_prime = twid_a * alpha_s * isqrtd;
	const double afpbb = af_prime * bt_star * bet_starsq;
	const double aspbb = as_prime * bt_star * bet_starsq;

	/* eigenvalues, eq. B17 */
	const double vax = sqrt(vaxsq);
	ev[0] = v1 - cf;
	ev[1] = v1 - vax;
	ev[2] = v1 - cs;
	ev[3] = v1;
	ev[4] = v1 + cs;
	ev[5] = v1 + vax;
	ev[6] = v1 + cf;

	/* right eigenvectors as columns, eq. B21 */
	rem[0][0] = alpha_f;
	rem[0][1] = 0.0;
	rem[0][2] = alpha_s;
	rem[0][3] = 1.0;
	rem[0][4] = alpha_s;
	rem[0][5] = 0.0;
	rem[0][6] = alpha_f;

	rem[1][0] = alpha_f * ev[0];
	rem[1][1] = 0.0;
	rem[1][2] = alpha_s * ev[2];
	rem[1][3] = v1;
	rem[1][4] = alpha_s * ev[4];
	rem[1][5] = 0.0;
	rem[1][6] = alpha_f * ev[6];

	double qa = alpha_f * v2, qb = alpha_s * v2, qc = qs * bet2_star, qd = qf * bet2_star;
	rem[2][0] = qa + qc;
	rem[2][1] = -bet3;
	rem[2][2] = qb - qd;
	rem[2][3] = v2;
	rem[2][4] = qb + qd;
	rem[2][5] = bet3;
	rem[2][6] = qa - qc;

	qa = alpha_f * v3;
	qb = alpha_s * v3;
	qc = qs * bet3_star;
	qd = qf * bet3_star;
	rem[3][0] = qa + qc;
	rem[3][1] = bet2;
	rem[3][2] = qb - qd;
	rem[3][3] = v3;
	rem[3][4] = qb + qd;
	rem[3][5] = -bet2;
	rem[3][6] = qa - qc;

	rem[4][0] = alpha_f * (hp - v1 * cf) + qs * vbet + aspbb;
	rem[4][1] = -(v2 * bet3 - v3 * bet2);
	rem[4][2] = alpha_s * (hp - v1 * cs) - qf * vbet - afpbb;
	rem[4][3] = 0.5 * vsq + g2 * x / g1;
	rem[4][4] = alpha_s * (hp + v1 * cs) + qf * vbet - afpbb;
	rem[4][5] = -rem[4][1];
	rem[4][6] = alpha_f * (hp + v1 * cf) - qs * vbet + aspbb;

	rem[5][0] = as_prime * bet2_star;
	rem[5][1] = -bet3 * s * isqrtd;
	rem[5][2] = -af_prime * bet2_star;
	rem[5][3] = 0.0;
	rem[5][4] = rem[5][2];
	rem[5][5] = rem[5][1];
	rem[5][6] = rem[5][0];

	rem[6][0] = as_prime * bet3_star;
	rem[6][1] = bet2 * s * isqrtd;
	rem[6][2] = -af_prime * bet3_star;
	rem[6][3] = 0.0;
	rem[6][4] = rem[6][2];
	rem[6][5] = rem[6][1];
	rem[6][6] = rem[6][0];

	/* left eigenvectors as rows, eq. B29; normalised by 1/2a^2 */
	double norm = 0.5 / twid_asq;
	const double cff = norm * alpha_f * cf;
	const double css = norm * alpha_s * cs;
	qf *= norm;
	qs *= norm;
	const double af = norm * af_prime * d;
	const double as = norm * as_prime * d;
	const double afpb = norm * af_prime * bt_star;
	const double aspb = norm * as_prime * bt_star;

	/* normalised by (gamma-1)/2a^2 */
	norm *= g1;
	alpha_f *= norm;
	alpha_s *= norm;
	const double q2_star = bet2_star / bet_starsq;
	const double q3_star = bet3_star / bet_starsq;
	const double vqstr = (v2 * q2_star + v3 * q3_star);
	norm *= 2.0;

	lem[0][0] = alpha_f * (vsq - hp) + cff * (cf + v1) - qs * vqstr - aspb;
	lem[0][1] = -alpha_f * v1 - cff;
	lem[0][2] = -alpha_f * v2 + qs * q2_star;
	lem[0][3] = -alpha_f * v3 + qs * q3_star;
	lem[0][4] = alpha_f;
	lem[0][5] = as * q2_star - alpha_f * b2;
	lem[0][6] = as * q3_star - alpha_f * b3;

	lem[1][0] = 0.5 * (v2 * bet3 - v3 * bet2);
	lem[1][1] = 0.0;
	lem[1][2] = -0.5 * bet3;
	lem[1][3] = 0.5 * bet2;
	lem[1][4] = 0.0;
	lem[1][5] = -0.5 * sqrtd * bet3 * s;
	lem[1][6] = 0.5 * sqrtd * bet2 * s;

	lem[2][0] = alpha_s * (vsq - hp) + css * (cs + v1) + qf * vqstr + afpb;
	lem[2][1] = -alpha_s * v1 - css;
	lem[2][2] = -alpha_s * v2 - qf * q2_star;
	lem[2][3] = -alpha_s * v3 - qf * q3_star;
	lem[2][4] = alpha_s;
	lem[2][5] = -af * q2_star - alpha_s * b2;
	lem[2][6] = -af * q3_star - alpha_s * b3;

	lem[3][0] = 1.0 - norm * (0.5 * vsq - g2 * x / g1);
	lem[3][1] = norm * v1;
	lem[3][2] = norm * v2;
	lem[3][3] = norm * v3;
	lem[3][4] = -norm;
	lem[3][5] = norm * b2;
	lem[3][6] = norm * b3;

	lem[4][0] = alpha_s * (vsq - hp) + css * (cs - v1) - qf * vqstr + afpb;
	lem[4][1] = -alpha_s * v1 + css;
	lem[4][2] = -alpha_s * v2 + qf * q2_star;
	lem[4][3] = -alpha_s * v3 + qf * q3_star;
	lem[4][4] = alpha_s;
	lem[4][5] = lem[2][5];
	lem[4][6] = lem[2][6];

	lem[5][0] = -lem[1][0];
	lem[5][1] = 0.0;
	lem[5][2] = -lem[1][2];
	lem[5][3] = -lem[1][3];
	lem[5][4] = 0.0;
	lem[5][5] = lem[1][5];
	lem[5][6] = lem[1][6];

	lem[6][0] = alpha_f * (vsq - hp) + cff * (cf - v1) + qs * vqstr - aspb;
	lem[6][1] = -alpha_f * v1 + cff;
	lem[6][2] = -alpha_f * v2 - qs * q2_star;
	lem[6][3] = -alpha_f * v3 - qs * q3_star;
	lem[6][4] = alpha_f;
	lem[6][5] = lem[0][5];
	lem[6][6] = lem[0][6];
}

/* 1-D Roe flux in scaled units, source/mhd/roe_athena.hpp:341-559.
   Ul/Ur = {d, Mx, My, Mz, E, By, Bz}, Wl/Wr = {d, Vx, Vy, Vz, P, By, Bz}. */
static int roe_fluxes(const double Ul[7], const double Ur[7], const double Wl[7],
	const double Wr[7], double Bxi, double gamma, double flux[7])
{
	enum { D = 0, X = 1, Y = 2, Z = 3, E = 4, BY = 5, BZ = 6 };
	double ev[7], rem[7][7], lem[7][7], a[7], dU[7], u[7];

	const double sqrtdl = sqrt(Wl[D]);
	const double sqrtdr = sqrt(Wr[D]);
	const double isdlpdr = 1.0 / (sqrtdl + sqrtdr);
	const double droe = sqrtdl * sqrtdr;
	const double v1roe = (sqrtdl * Wl[X] + sqrtdr * Wr[X]) * isdlpdr;
	const double v2roe = (sqrtdl * Wl[Y] + sqrtdr * Wr[Y]) * isdlpdr;
	const double v3roe = (sqrtdl * Wl[Z] + sqrtdr * Wr[Z]) * isdlpdr;
	/* the Roe average of the magnetic field is defined differently */
	const double b2roe = (sqrtdr * Wl[BY] + sqrtdl * Wr[BY]) * isdlpdr;
	const double b3roe = (sqrtdr * Wl[BZ] + sqrtdl * Wr[BZ]) * isdlpdr;
	const double dby = Wl[BY] - Wr[BY], dbz = Wl[BZ] - Wr[BZ], sd = sqrtdl + sqrtdr;
	const double x = 0.5 * (dby * dby + dbz * dbz) / (sd * sd);
	const double y = 0.5 * (Wl[D] + Wr[D]) / droe;
	const double pbl = 0.5 * (Bxi * Bxi + Wl[BY] * Wl[BY] + Wl[BZ] * Wl[BZ]);
	const double pbr = 0.5 * (Bxi * Bxi + Wr[BY] * Wr[BY] + Wr[BZ] * Wr[BZ]);
	/* enthalpy average, Roe (1981) */
	const double hroe = ((Ul[E] + Wl[E] + pbl) / sqrtdl + (Ur[E] + Wr[E] + pbr) / sqrtdr) * isdlpdr;

	or_roe_eigensystem(droe, v1roe, v2roe, v3roe, hroe, Bxi, b2roe, b3roe, x, y,
		gamma, ev, rem, lem);

	double Fl[7], Fr[7];
	Fl[D] = Ul[X];
	Fr[D] = Ur[X];
	Fl[X] = Ul[X] * Wl[X];
	Fr[X] = Ur[X] * Wr[X];
	Fl[Y] = Ul[X] * Wl[Y];
	Fr[Y] = Ur[X] * Wr[Y];
	Fl[Z] = Ul[X] * Wl[Z];
	Fr[Z] = Ur[X] * Wr[Z];
	Fl[X] += Wl[E];
	Fr[X] += Wr[E];
	Fl[E] = (Ul[E] + Wl[E]) * Wl[X];
	Fr[E] = (Ur[E] + Wr[E]) * Wr[X];
	Fl[X] -= 0.5 * (Bxi * Bxi - Wl[BY] * Wl[BY] - Wl[BZ] * Wl[BZ]);
	Fr[X] -= 0.5 * (Bxi * Bxi - Wr[BY] * Wr[BY] - Wr[BZ] * Wr[BZ]);
	Fl[Y] -= Bxi * Wl[BY];
	Fr[Y] -= Bxi * Wr[BY];
	Fl[Z] -= Bxi * Wl[BZ];
	Fr[Z] -= Bxi * Wr[BZ];
	Fl[E] += pbl * Wl[X] - Bxi * (Bxi * Wl[X] + Wl[BY] * Wl[Y] + Wl[BZ] * Wl[Z]);
	Fr[E] += pbr * Wr[X] - Bxi * (Bxi * Wr[X] + Wr[BY] * Wr[Y] + Wr[BZ] * Wr[Z]);
	Fl[BY] = Wl[BY] * Wl[X] - Bxi * Wl[Y];
	Fr[BY] = Wr[BY] * Wr[X] - Bxi * Wr[Y];
	Fl[BZ] = Wl[BZ] * Wl[X] - Bxi * Wl[Z];
	Fr[BZ] = Wr[BZ] * Wr[X] - Bxi * Wr[Z];

	/* upwind flux if the flow is supersonic */
	if (ev[0] >= 0.0) {
		memcpy(flux, Fl, sizeof(Fl));
		return OR_OK;
	}
	if (ev[6] <= 0.0) {
		memcpy(flux, Fr, sizeof(Fr));
		return OR_OK;
	}

	/* projection of dU onto the left eigenvectors */
	for (int m = 0; m < 7; m++) dU[m] = Ur[m] - Ul[m];
	for (int n = 0; n < 7; n++) {
		a[n] = 0.0;
		for (int m = 0; m < 7; m++) a[n] += lem[n][m] * dU[m];
	}

	/* positivity of the intermediate states */
	for (int m = 0; m < 7; m++) u[m] = Ul[m];
	for (int n = 0; n < 6; n++) {
		for (int m = 0; m < 7; m++) u[m] += a[n] * rem[m][n];
		if (ev[n + 1] > ev[n]) {
			if (u[0] <= 0.0) return OR_ERR_ROE_DENSITY;
			double p = u[4] - 0.5 * (u[1] * u[1] + u[2] * u[2] + u[3] * u[3]) / u[0];
			p -= 0.5 * (u[5] * u[5] + u[6] * u[6] + Bxi * Bxi);
			if (p < 0.0) return OR_ERR_ROE_PRESSURE;
		}
	}

	double coeff[7];
	for (int m = 0; m < 7; m++) coeff[m] = 0.5 * a[m] * smax(fabs(ev[m]), 0.0);
	for (int q = 0; q < 7; q++) flux[q] = 0.5 * (Fl[q] + Fr[q]);
	for (int m = 0; m < 7; m++) {
		for (int q = 0; q < 7; q++) flux[q] -= coeff[m] * rem[q][m];
	}
	return OR_OK;
}

/* source/mhd/roe_athena.hpp:567-818 (ignores the background field, :584) */
int or_flux_roe(const double L[8], const double R[8], const double bg_unused[3],
	double gamma, double mu0, double F[8], double* s_out)
{
	(void)bg_unused;
	const double zero[3] = {0, 0, 0};
	if (L[RHO] <= 0) return OR_ERR_DENSITY;
	if (R[RHO] <= 0) return OR_ERR_DENSITY;
	double pL, pR;
	int e = or_get_pressure(L, gamma, mu0, &pL);
	if (e) return e;
	if (pL <= 0) return OR_ERR_PRESSURE;
	e = or_get_pressure(R, gamma, mu0, &pR);
	if (e) return e;
	if (pR <= 0) return OR_ERR_PRESSURE;

	const double Bxi = (L[B0] + R[B0]) / (2.0 * sqrt(mu0));
	const double Ul[7] = {L[RHO], L[M0], L[M1], L[M2], L[NRJ],
		L[B1] / sqrt(mu0), L[B2] / sqrt(mu0)};
	const double Ur[7] = {R[RHO], R[M0], R[M1], R[M2], R[NRJ],
		R[B1] / sqrt(mu0), R[B2] / sqrt(mu0)};

	/* get_primitive with vacuum permeability 1 on the unscaled field
	   (roe_athena.hpp:705-721) */
	double pl1, pr1, vL[3], vR[3];
	e = or_get_pressure(L, gamma, 1.0, &pl1);
	if (e) return e;
	e = or_get_pressure(R, gamma, 1.0, &pr1);
	if (e) return e;
	velocity(L, vL);
	velocity(R, vR);
	const double Wl[7] = {L[RHO], vL[0], vL[1], vL[2], pl1, L[B1] / sqrt(mu0), L[B2] / sqrt(mu0)};
	const double Wr[7] = {R[RHO], vR[0], vR[1], vR[2], pr1, R[B1] / sqrt(mu0), R[B2] / sqrt(mu0)};

	double f7[7];
	e = roe_fluxes(Ul, Ur, Wl, Wr, Bxi, gamma, f7);
	if (e) return e;

	F[RHO] = f7[0];
	F[M0] = f7[1];
	F[M1] = f7[2];
	F[M2] = f7[3];
	F[NRJ] = f7[4];
	F[B0] = 0;
	F[B1] = f7[5] * sqrt(mu0);
	F[B2] = f7[6] * sqrt(mu0);

	double cfL, cfR;
	e = or_fast_speed(L, zero, gamma, mu0, &cfL);
	if (e) return e;
	e = or_fast_speed(R, zero, gamma, mu0, &cfR);
	if (e) return e;
	const double ms = smax(cfL, cfR);
	const double iL = 1 / L[RHO], iR = 1 / R[RHO];
	const double vL0 = L[M0] * iL, vR0 = R[M0] * iR;
	const double
		msn = (vL0 <= vR0) ? vL0 - ms : vR0 - ms,
		msp = (vL0 <= vR0) ? vR0 + ms : vL0 + ms;
	if (msn >= 0.0) {
		*s_out = fabs(msp);
	} else if (msp <= 0.0) {
		*s_out = fabs(msn);
	} else {
		*s_out = smax(fabs(msn), fabs(msp));
	}
	return OR_OK;
}

int or_flux(int solver, const double L[8], const double R[8], const double bg[3],
	double gamma, double mu0, double F[8], double* s)
{
	switch (solver) {
	case OR_RUSANOV: return or_flux_rusanov(L, R, bg, gamma, mu0, F, s);
	case OR_HLL_ATHENA: return or_flux_hll(L, R, bg, gamma, mu0, F, s);
	case OR_HLLD_ATHENA: return or_flux_hlld(L, R, bg, gamma, mu0, F, s);
	case OR_ROE_ATHENA: return or_flux_roe(L, R, bg, gamma, mu0, F, s);
	default: return OR_ERR_SOLVER;
	}
}

size_t or_face_flux_batch(int solver, size_t n, const double* L, const double* R,
	const double* bg, double gamma, double mu0, double* F, double* smax_out,
	int* status)
{
	size_t bad = 0;
	for (size_t f = 0; f < n; f++) {
		double l[8], r[8], b[3] = {0, 0, 0}, out[8] = {0, 0, 0, 0, 0, 0, 0, 0}, s = 0;
		for (int q = 0; q < 8; q++) {
			l[q] = L[q * n + f];
			r[q] = R[q * n + f];
		}
		if (bg) for (int q = 0; q < 3; q++) b[q] = bg[q * n + f];
		const int e = or_flux(solver, l, r, b, gamma, mu0, out, &s);
		if (e) {
			bad++;
			for (int q = 0; q < 8; q++) out[q] = 0;
			s = 0;
		}
		for (int q = 0; q < 8; q++) F[q * n + f] = out[q];
		smax_out[f] = s;
		if (status) status[f] = e;
	}
	return bad;
}

/* ---------------------------------------------------------------------- */
/* uniform periodic staggered grid                                        */
/* ---------------------------------------------------------------------- */

/*
get_rotated_vector, source/mhd/common.hpp:612-648: a d-face (d = 0, 1, 2 for
x, y, z) sees the global vector v as
    d = 0: (v0, v1, v2),  d = 1: (v1, v2, v0),  d = 2: (v2, v0, v1)
*/
static inline void rot_to_face(int d, const double* v, double* r)
{
	if (d == 0) { r[0] = v[0]; r[1] = v[1]; r[2] = v[2]; }
	else if (d == 1) { r[0] = v[1]; r[1] = v[2]; r[2] = v[0]; }
	else { r[0] = v[2]; r[1] = v[0]; r[2] = v[1]; }
}
static inline void rot_to_global(int d, const double* r, double* v)
{
	if (d == 0) { v[0] = r[0]; v[1] = r[1]; v[2] = r[2]; }
	else if (d == 1) { v[1] = r[0]; v[2] = r[1]; v[0] = r[2]; }
	else { v[2] = r[0]; v[0] = r[1]; v[1] = r[2]; }
}

typedef struct {
	long nx, ny, nz, n;
} dims_t;

static inline long wrap(long i, long n) { return ((i % n) + n) % n; }

static inline long cid(const dims_t* D, long i, long j, long k)
{
	return wrap(i, D->nx) + D->nx * (wrap(j, D->ny) + D->ny * wrap(k, D->nz));
}

/* cell-centred B = average of the two faces, source/mhd/solve.hpp:1661-1663 */
static inline void vol_b(const dims_t* D, const double* s, long i, long j, long k,
	double B[3])
{
	const long n = D->n, c = cid(D, i, j, k);
	B[0] = 0.5 * (s[5 * n + c] + s[5 * n + cid(D, i + 1, j, k)]);
	B[1] = 0.5 * (s[6 * n + c] + s[6 * n + cid(D, i, j + 1, k)]);
	B[2] = 0.5 * (s[7 * n + c] + s[7 * n + cid(D, i, j, k + 1)]);
}

/* cell state rotated into the frame of a d-face, source/mhd/solve.hpp:106-118 */
static inline void face_state(const dims_t* D, const double* s, int d, long i,
	long j, long k, double U[8])
{
	const long n = D->n, c = cid(D, i, j, k);
	const double m[3] = {s[1 * n + c], s[2 * n + c], s[3 * n + c]};
	double B[3];
	vol_b(D, s, i, j, k, B);
	U[RHO] = s[c];
	U[NRJ] = s[4 * n + c];
	rot_to_face(d, m, U + M0);
	rot_to_face(d, B, U + B0);
}

/*
Constant-thermal-pressure energy refit of update_vol_B,
source/mhd/solve.hpp:1646-1672: pressure from (rho, m, E, vb_old), then
E = p/(gamma-1) + rho v^2/2 + vb_new^2/(2 mu0).  Cells with rho <= 0 are
skipped as in the reference.
*/
static void refit_energy(double rho, const double m[3], double* E,
	const double vb_old[3], const double vb_new[3], double gamma, double mu0)
{
	if (rho <= 0) return;
	const double U[8] = {rho, m[0], m[1], m[2], *E, vb_old[0], vb_old[1], vb_old[2]};
	double p;
	or_get_pressure(U, gamma, mu0, &p);
	const double inv = 1 / rho;
	const double v[3] = {m[0] * inv, m[1] * inv, m[2] * inv};
	*E = or_total_energy(rho, v, p, vb_new, gamma, mu0);
}

/* the periodic step's flux buffer, grown as needed and kept between calls
   (one caller at a time) */
static double* step_flux_buffer(size_t count)
{
	static double* buf = NULL;
	static size_t cap = 0;
	if (count > cap) {
		free(buf);
		buf = (double*)malloc(sizeof(double) * count);
		cap = buf ? count : 0;
	}
	return buf;
}

int or_step_periodic(const or_grid* g, int solver, double dt, int refits,
	const double* in, double* out, double smax_dim[3], long zero_dim[3],
	long* bad_face)
{
	const dims_t D = {g->nx, g->ny, g->nz, (long)g->nx * g->ny * g->nz};
	const long n = D.n;
	const double len[3] = {g->dx, g->dy, g->dz};
	/* P[(d*8 + q)*n + c] = dt * F_q (global frame) of the d-face owned by
	   cell c, i.e. the face on the cell's negative d side; kept between
	   calls (the reference keeps its flux slots in the cells) */
	double* P = step_flux_buffer(24 * n);
	if (!P) return OR_ERR_SOLVER;
	int err = OR_OK;
	double sm[3] = {0, 0, 0};
	long zc[3] = {0, 0, 0};

	/* flux pass of get_fluxes, source/mhd/solve.hpp:304-393 with get_flux :68-188 */
	long bad = -1;
	or_fc = 3ull * (unsigned long long)n;
	for (int d = 0; d < 3 && !err; d++) {
		double smd = 0;
		long zcd = 0;
		#pragma omp parallel for collapse(2) schedule(static) reduction(max: smd) reduction(+: zcd)
		for (long k = 0; k < D.nz; k++)
		for (long j = 0; j < D.ny; j++)
		for (long i = 0; i < D.nx; i++) {
			const long c = cid(&D, i, j, k);
			double L[8], R[8], bgr[3] = {0, 0, 0}, F[8], s = 0;
			face_state(&D, in, d, i - (d == 0), j - (d == 1), k - (d == 2), L);
			face_state(&D, in, d, i, j, k, R);
			if (g->bg) {
				/* Bg_B(L)(+d): the face is the positive d-face of L */
				const long cl = cid(&D, i - (d == 0), j - (d == 1), k - (d == 2));
				const double b[3] = {
					g->bg[(3 * d + 0) * n + cl],
					g->bg[(3 * d + 1) * n + cl],
					g->bg[(3 * d + 2) * n + cl]};
				rot_to_face(d, b, bgr);
			}
			const int e = or_flux(solver, L, R, bgr, g->gamma, g->mu0, F, &s);
			if (e) {
				#pragma omp critical
				{
					if (!err || (long)d * n + c < bad) {
						err = e;
						bad = (long)d * n + c;
					}
				}
				continue;
			}
			double Fg[8];
			Fg[RHO] = F[RHO];
			Fg[NRJ] = F[NRJ];
			rot_to_global(d, F + M0, Fg + M0);
			rot_to_global(d, F + B0, Fg + B0);
			for (int q = 0; q < 8; q++) P[(d * 8 + q) * n + c] = dt * Fg[q];
			if (s <= 0) zcd++;
			if (s > smd) smd = s;
		}
		sm[d] = smd;
		zc[d] = zcd;
	}
	if (err && bad_face) *bad_face = bad;
	if (smax_dim) for (int d = 0; d < 3; d++) smax_dim[d] = sm[d];
	if (zero_dim) for (int d = 0; d < 3; d++) zero_dim[d] = zc[d];
	if (err) return err;

	const double dx = g->dx, dy = g->dy, dz = g->dz;
#define PF(d, q, i, j, k) P[((d) * 8 + (q)) * n + cid(&D, (i), (j), (k))]
	/* edge terms of assign_face_dBs_f{x,y,z}, source/mhd/solve.hpp:610-613,
	   :926-928, :1223-1225, written from dt*F of the face owned by (i,j,k):
	     x-face: Ez = +F_By dt dz/4, Ey = -F_Bz dt dy/4
	     y-face: Ex = +F_Bz dt dx/4, Ez = -F_Bx dt dz/4
	     z-face: Ex = -F_By dt dx/4, Ey = +F_Bx dt dy/4 */
#define EZX(i, j, k) (+(PF(0, 6, i, j, k) * dz / 4))
#define EYX(i, j, k) (-(PF(0, 7, i, j, k) * dy / 4))
#define EXY(i, j, k) (+(PF(1, 7, i, j, k) * dx / 4))
#define EZY(i, j, k) (-(PF(1, 5, i, j, k) * dz / 4))
#define EXZ(i, j, k) (-(PF(2, 6, i, j, k) * dx / 4))
#define EYZ(i, j, k) (+(PF(2, 5, i, j, k) * dy / 4))

	#pragma omp parallel for collapse(2) schedule(static)
	for (long k = 0; k < D.nz; k++)
	for (long j = 0; j < D.ny; j++)
	for (long i = 0; i < D.nx; i++) {
		const long c = cid(&D, i, j, k);

		/* face B += Face_dB / area, source/mhd/solve.hpp:1589-1593; each
		   face sums the 12 edge terms of the faces that share its edges */
		double dBx = 0, dBy = 0, dBz = 0;
		dBx += EZX(i, j + 1, k);
		dBx -= EZX(i, j - 1, k);
		dBx -= EYX(i, j, k + 1);
		dBx += EYX(i, j, k - 1);
		dBx += EZY(i, j + 1, k);
		dBx += EZY(i - 1, j + 1, k);
		dBx -= EZY(i, j, k);
		dBx -= EZY(i - 1, j, k);
		dBx -= EYZ(i, j, k + 1);
		dBx -= EYZ(i - 1, j, k + 1);
		dBx += EYZ(i, j, k);
		dBx += EYZ(i - 1, j, k);

		dBy += EXY(i, j, k + 1);
		dBy -= EXY(i, j, k - 1);
		dBy -= EZY(i + 1, j, k);
		dBy += EZY(i - 1, j, k);
		dBy += EXZ(i, j, k + 1);
		dBy += EXZ(i, j - 1, k + 1);
		dBy -= EXZ(i, j, k);
		dBy -= EXZ(i, j - 1, k);
		dBy -= EZX(i + 1, j, k);
		dBy += EZX(i, j, k);
		dBy -= EZX(i + 1, j - 1, k);
		dBy += EZX(i, j - 1, k);

		dBz += EYZ(i + 1, j, k);
		dBz -= EYZ(i - 1, j, k);
		dBz -= EXZ(i, j + 1, k);
		dBz += EXZ(i, j - 1, k);
		dBz += EYX(i + 1, j, k);
		dBz += EYX(i + 1, j, k - 1);
		dBz -= EYX(i, j, k);
		dBz -= EYX(i, j, k - 1);
		dBz -= EXY(i, j + 1, k);
		dBz -= EXY(i, j + 1, k - 1);
		dBz += EXY(i, j, k);
		dBz += EXY(i, j, k - 1);

		out[5 * n + c] = in[5 * n + c] + dBx / (dy * dz);
		out[6 * n + c] = in[6 * n + c] + dBy / (dx * dz);
		out[7 * n + c] = in[7 * n + c] + dBz / (dx * dy);
	}

	/* conserved update of update_mhd_state, source/mhd/solve.hpp:1551-1587;
	   flux slots hold 0 + dt*F (:375-392) */
	#pragma omp parallel for collapse(2) schedule(static)
	for (long k = 0; k < D.nz; k++)
	for (long j = 0; j < D.ny; j++)
	for (long i = 0; i < D.nx; i++) {
		const long c = cid(&D, i, j, k);
		double lo[3][8], hi[3][8];
		for (int q = 0; q < 8; q++) {
			lo[0][q] = 0.0 + PF(0, q, i, j, k);
			hi[0][q] = 0.0 + PF(0, q, i + 1, j, k);
			lo[1][q] = 0.0 + PF(1, q, i, j, k);
			hi[1][q] = 0.0 + PF(1, q, i, j + 1, k);
			lo[2][q] = 0.0 + PF(2, q, i, j, k);
			hi[2][q] = 0.0 + PF(2, q, i, j, k + 1);
		}
		double rho = in[c];
		double m[3] = {in[n + c], in[2 * n + c], in[3 * n + c]};
		double E = in[4 * n + c], vb[3];
		vol_b(&D, in, i, j, k, vb);
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

		/* energy refits: the first uses the flux-updated volume field
		   (timestep, :1904-1908), the later ones the new face average
		   (:1913-1916 and apply_boundaries, mhd/boundaries.hpp:1289-1292) */
		double vb_new[3];
		vol_b(&D, out, i, j, k, vb_new);
		for (int r = 0; r < refits; r++) {
			refit_energy(rho, m, &E, r == 0 ? vb : vb_new, vb_new, g->gamma, g->mu0);
		}
		out[c] = rho;
		out[n + c] = m[0];
		out[2 * n + c] = m[1];
		out[3 * n + c] = m[2];
		out[4 * n + c] = E;
	}
#undef PF
#undef EZX
#undef EYX
#undef EXY
#undef EZY
#undef EXZ
#undef EYZ
	return OR_OK;
}

/* minimize_timestep + set_minmax_substepping_period with one substep,
   source/mhd/solve.hpp:1961-2006, substepping.hpp:303-347 */
double or_cfl_dt(const or_grid* g, double cfl, const double smax_dim[3],
	const long zero_dim[3])
{
	const double len[3] = {g->dx, g->dy, g->dz};
	double dt = DBL_MAX;
	for (int d = 0; d < 3; d++) {
		if (zero_dim[d] > 0) return 0;
		dt = smin(dt, cfl * len[d] / smax_dim[d]);
	}
	return dt;
}

/* source/math/nabla.hpp:58-113 */
double or_div_face2volume(const or_grid* g, const double* s)
{
	const dims_t D = {g->nx, g->ny, g->nz, (long)g->nx * g->ny * g->nz};
	const long n = D.n;
	double total = 0;
	for (long k = 0; k < D.nz; k++)
	for (long j = 0; j < D.ny; j++)
	for (long i = 0; i < D.nx; i++) {
		const long c = cid(&D, i, j, k);
		const double div
			= (s[5 * n + cid(&D, i + 1, j, k)] - s[5 * n + c]) / g->dx
			+ (s[6 * n + cid(&D, i, j + 1, k)] - s[6 * n + c]) / g->dy
			+ (s[7 * n + cid(&D, i, j, k + 1)] - s[7 * n + c]) / g->dz;
		total += fabs(div);
	}
	return total / (double)n;
}

int or_run_periodic(const or_grid* g, int solver, double cfl, double t_start,
	double t_end, long max_steps, double* state, long* steps_taken,
	double* t_final)
{
	const long n = (long)g->nx * g->ny * g->nz;
	const dims_t D = {g->nx, g->ny, g->nz, n};
	double* tmp = (double*)malloc(sizeof(double) * 8 * n);
	if (!tmp) return OR_ERR_SOLVER;

	/* apply_boundaries before the final init, tests/mhd/test.cpp:446-453:
	   one refit with the face-average field */
	for (long k = 0; k < D.nz; k++)
	for (long j = 0; j < D.ny; j++)
	for (long i = 0; i < D.nx; i++) {
		const long c = cid(&D, i, j, k);
		double vb[3];
		vol_b(&D, state, i, j, k, vb);
		const double m[3] = {state[n + c], state[2 * n + c], state[3 * n + c]};
		refit_energy(state[c], m, &state[4 * n + c], vb, vb, g->gamma, g->mu0);
	}

	/* final init with a zero time step, test.cpp:462-470 */
	double sm[3];
	long zc[3];
	int e = or_step_periodic(g, solver, 0.0, 2, state, tmp, sm, zc, NULL);
	if (e) { free(tmp); return e; }
	memcpy(state, tmp, sizeof(double) * 8 * n);

	double t = t_start;
	long steps = 0;
	while (t < t_end && steps < max_steps) {
		const double until_end = t_end - t;
		const double cfl_dt = or_cfl_dt(g, cfl, sm, zc);
		/* substepping.hpp:331-338: max_dt <= smallest -> max_dt */
		const double dt = (until_end <= cfl_dt) ? until_end : cfl_dt;
		e = or_step_periodic(g, solver, dt, 3, state, tmp, sm, zc, NULL);
		if (e) { free(tmp); return e; }
		memcpy(state, tmp, sizeof(double) * 8 * n);
		t += dt;
		steps++;
	}
	free(tmp);
	if (steps_taken) *steps_taken = steps;
	if (t_final) *t_final = t;
	return OR_OK;
}

/* ---------------------------------------------------------------------- */
/* per-cell 2^N substepping on a uniform periodic grid                    */
/* ---------------------------------------------------------------------- */

/*
The periodic step with the reference's per-cell substepping,
timestep(), source/mhd/solve.hpp:1837-1924, on a grid of normal cells:

  set_minmax_substepping_period (options), substepping.hpp:194-261
  minimize_timestep, solve.hpp:1927-2014
  set_minmax_substepping_period (dt), substepping.hpp:264-399
  restrict_substepping_period, substepping.hpp:99-191
  update_substeps (2^N periods), substepping.hpp:54-96
  for substep s = 1 .. max period:
      get_fluxes, solve.hpp:240-569: the +d faces of cell c are solved when
          s is a multiple of the smallest period among c and its face and
          edge neighbours (m), with weight dt_sub * m into the flux slots of
          both cells and into the edge terms of the face dB
      update_mhd_state, solve.hpp:1503-1603: cells whose period divides s
          apply and clear their slots and face dB, the others keep them
      update_vol_B(s), solve.hpp:1618-1683: constant-pressure refit of the
          updated cells from their own face copies
  sync_magnetic_field, common_functions.hpp:184-269 (two visits per face,
      the cell with the smaller index first)
  update_vol_B(0) of all cells.

Working layout: S 14 x n (rho, m, E, vol B, the cell's copies of its -x, +x,
-y, +y, -z, +z faces), slots 48 x n ((2d + side) * 8 + q), face dB 6 x n.
The state handed in and out is the face-once layout of or_step_periodic
(copies are equal at step boundaries).  maxv (3 x n) holds the signal speed
of each face (Max_v of both cells, which the reference sets together),
negative before the first solution.

Deviation: floor(log2(x)) of substepping.hpp:385 is computed exactly
(ilogb), and a non-finite ratio (zero smallest dt) gives 0, the value the
reference's int conversion produces on x86.
*/

#define SC(q, c) S[(long)(q) * n + (c)]
#define SLOT(dir, q, c) slot[((long)(dir) * 8 + (q)) * n + (c)]

static void sub_face_state(const double* S, long n, int d, long c, double U[8])
{
	const double m[3] = {SC(1, c), SC(2, c), SC(3, c)};
	const double B[3] = {SC(5, c), SC(6, c), SC(7, c)};
	U[RHO] = SC(0, c);
	U[NRJ] = SC(4, c);
	rot_to_face(d, m, U + M0);
	rot_to_face(d, B, U + B0);
}

int or_substep_periods(const or_grid* g, double time_step_factor, double max_dt,
	const int* substep_min, const int* substep_max, const double* maxv, int* period, int* minsub,
	double* sub_dt_out, int* max_period_out)
{
	const dims_t D = {g->nx, g->ny, g->nz, (long)g->nx * g->ny * g->nz};
	const long n = D.n;
	const double len[3] = {g->dx, g->dy, g->dz};
	double* ts = (double*)malloc(sizeof(double) * n);
	int* lomin = (int*)malloc(sizeof(int) * n);
	int* smaxv = (int*)malloc(sizeof(int) * n);
	if (!ts || !lomin || !smaxv) { free(ts); free(lomin); free(smaxv); return OR_ERR_SOLVER; }

	/* options of each cell; the smallest minimum is shifted to 0 */
	int lo = INT_MAX;
	for (long c = 0; c < n; c++) {
		lomin[c] = substep_min[c];
		smaxv[c] = substep_max[c];
		if (lomin[c] < lo) lo = lomin[c];
	}
	if (lo > 0) for (long c = 0; c < n; c++) lomin[c] -= lo;

	/* minimize_timestep: each cell's limit over its six faces */
	for (long k = 0; k < D.nz; k++)
	for (long j = 0; j < D.ny; j++)
	for (long i = 0; i < D.nx; i++) {
		const long c = cid(&D, i, j, k);
		double t = DBL_MAX;
		for (int d = 0; d < 3; d++) {
			const long fm = c, fp = cid(&D, i + (d == 0), j + (d == 1), k + (d == 2));
			const double v[2] = {maxv[(long)d * n + fm], maxv[(long)d * n + fp]};
			for (int sd = 0; sd < 2; sd++) {
				if (v[sd] <= 0) t = 0;
				else t = smin(time_step_factor * len[d] / v[sd], t);
			}
		}
		ts[c] = t;
	}

	double smallest = DBL_MAX;
	for (long c = 0; c < n; c++) smallest = smin(ts[c], smallest);
	double sub_dt;
	if (max_dt <= smallest) {
		for (long c = 0; c < n; c++) lomin[c] = smaxv[c] = 0;
		sub_dt = max_dt;
	} else {
		double ret = DBL_MAX;
		for (long c = 0; c < n; c++) ret = smin(ret, ts[c] / (double)(1 << lomin[c]));
		for (long c = 0; c < n; c++) {
			const double dt = ts[c];
			if (dt > 0) {
				const double x = smin(dt, max_dt) / ret;
				int step = (isfinite(x) && x > 0) ? ilogb(x) : 0;
				if (step < 0) step = 0;
				if (smaxv[c] > step) smaxv[c] = step;
				/* the reference aborts here; a level above 30 would
				   overflow its int period 1 << level */
				if (lomin[c] > smaxv[c] || smaxv[c] > 30) {
					free(ts); free(lomin); free(smaxv);
					return OR_ERR_SUBSTEP;
				}
			} else {
				lomin[c] = smaxv[c] = 0;
			}
		}
		sub_dt = ret;
	}

	/* restrict_substepping_period: at most one level above any face or
	   edge neighbour (the fixed point is independent of the sweep order) */
	int* e = period;
	for (long c = 0; c < n; c++) e[c] = smaxv[c];
	for (;;) {
		long modified = 0;
		for (long k = 0; k < D.nz; k++)
		for (long j = 0; j < D.ny; j++)
		for (long i = 0; i < D.nx; i++) {
			const long c = cid(&D, i, j, k);
			for (int dk = -1; dk <= 1; dk++)
			for (int dj = -1; dj <= 1; dj++)
			for (int di = -1; di <= 1; di++) {
				const int nz_ = (di != 0) + (dj != 0) + (dk != 0);
				if (nz_ == 0 || nz_ == 3) continue;   /* face and edge neighbours */
				const long nb = cid(&D, i + di, j + dj, k + dk);
				if (e[c] > e[nb] + 1) { e[c] = e[nb] + 1; modified++; }
			}
		}
		if (!modified) break;
	}
	/* update_substeps: 2^N periods */
	int maxp = -1;
	for (long c = 0; c < n; c++) { e[c] = 1 << e[c]; if (e[c] > maxp) maxp = e[c]; }
	/* smallest period among each cell and its face and edge neighbours */
	for (long k = 0; k < D.nz; k++)
	for (long j = 0; j < D.ny; j++)
	for (long i = 0; i < D.nx; i++) {
		const long c = cid(&D, i, j, k);
		int m = e[c];
		for (int dk = -1; dk <= 1; dk++)
		for (int dj = -1; dj <= 1; dj++)
		for (int di = -1; di <= 1; di++) {
			const int nz_ = (di != 0) + (dj != 0) + (dk != 0);
			if (nz_ == 0 || nz_ == 3) continue;
			const int pn = e[cid(&D, i + di, j + dj, k + dk)];
			if (pn < m) m = pn;
		}
		minsub[c] = m;
	}
	free(ts); free(lomin); free(smaxv);
	*sub_dt_out = sub_dt;
	*max_period_out = maxp;
	return OR_OK;
}

int or_step_substep(const or_grid* g, int solver, double time_step_factor, double max_dt,
	const int* substep_min, const int* substep_max, double* state, double* maxv, double* dt_out,
	int* max_period_out, long* bad_face)
{
	const dims_t D = {g->nx, g->ny, g->nz, (long)g->nx * g->ny * g->nz};
	const long n = D.n;
	const double len[3] = {g->dx, g->dy, g->dz};
	const double area[3] = {g->dy * g->dz, g->dx * g->dz, g->dx * g->dy};
	int* period = (int*)malloc(sizeof(int) * n);
	int* minsub = (int*)malloc(sizeof(int) * n);
	double* S = (double*)malloc(sizeof(double) * 14 * n);
	double* slot = (double*)calloc((size_t)48 * n, sizeof(double));
	double* fdb = (double*)calloc((size_t)6 * n, sizeof(double));
	double* P = (double*)malloc(sizeof(double) * 24 * n);
	int err = OR_OK;
	if (!period || !minsub || !S || !slot || !fdb || !P) { err = OR_ERR_SOLVER; goto done; }

	double sub_dt;
	int maxp;
	err = or_substep_periods(g, time_step_factor, max_dt, substep_min, substep_max, maxv,
		period, minsub, &sub_dt, &maxp);
	if (err) goto done;

	/* expand: vol B = face average, both copies of each face equal */
	for (long k = 0; k < D.nz; k++)
	for (long j = 0; j < D.ny; j++)
	for (long i = 0; i < D.nx; i++) {
		const long c = cid(&D, i, j, k);
		for (int q = 0; q < 5; q++) SC(q, c) = state[q * n + c];
		double B[3];
		vol_b(&D, state, i, j, k, B);
		for (int d = 0; d < 3; d++) {
			SC(5 + d, c) = B[d];
			SC(8 + 2 * d, c) = state[(5 + d) * n + c];
			SC(9 + 2 * d, c) = state[(5 + d) * n + cid(&D, i + (d == 0), j + (d == 1), k + (d == 2))];
		}
	}

	double total_dt = 0;
	const double dx = g->dx, dy = g->dy, dz = g->dz;
	or_fc = 0;
	for (int sub = 1; sub <= maxp; sub++) {
		total_dt += sub_dt;
		/* fluxes of the faces due at this substep; P = 0 for the others */
		long bad = -1;
		for (int d = 0; d < 3 && !err; d++)
		for (long k = 0; k < D.nz; k++)
		for (long j = 0; j < D.ny; j++)
		for (long i = 0; i < D.nx; i++) {
			const long c = cid(&D, i, j, k);
			const long o = cid(&D, i - (d == 0), j - (d == 1), k - (d == 2));  /* owner: -side cell */
			const int m = minsub[o];
			if (sub % m != 0) {
				for (int q = 0; q < 8; q++) P[(d * 8 + q) * n + c] = 0;
				continue;
			}
			or_fc++;
			double L[8], R[8], bgr[3] = {0, 0, 0}, F[8], sp = 0;
			sub_face_state(S, n, d, o, L);
			sub_face_state(S, n, d, c, R);
			if (g->bg) {
				const double b[3] = {g->bg[(3 * d + 0) * n + o], g->bg[(3 * d + 1) * n + o],
					g->bg[(3 * d + 2) * n + o]};
				rot_to_face(d, b, bgr);
			}
			const int e = or_flux(solver, L, R, bgr, g->gamma, g->mu0, F, &sp);
			if (e) {
				if (!err || (long)d * n + c < bad) { err = e; bad = (long)d * n + c; }
				continue;
			}
			maxv[(long)d * n + c] = sp;
			double Fg[8];
			Fg[RHO] = F[RHO];
			Fg[NRJ] = F[NRJ];
			rot_to_global(d, F + M0, Fg + M0);
			rot_to_global(d, F + B0, Fg + B0);
			const double min_dt = sub_dt * m;
			for (int q = 0; q < 8; q++) P[(d * 8 + q) * n + c] = min_dt * Fg[q];
		}
		if (err) { if (bad_face) *bad_face = bad; goto done; }

#define PF(d, q, i, j, k) P[((d) * 8 + (q)) * n + cid(&D, (i), (j), (k))]
#define EZX(i, j, k) (+(PF(0, 6, i, j, k) * dz / 4))
#define EYX(i, j, k) (-(PF(0, 7, i, j, k) * dy / 4))
#define EXY(i, j, k) (+(PF(1, 7, i, j, k) * dx / 4))
#define EZY(i, j, k) (-(PF(1, 5, i, j, k) * dz / 4))
#define EXZ(i, j, k) (-(PF(2, 6, i, j, k) * dx / 4))
#define EYZ(i, j, k) (+(PF(2, 5, i, j, k) * dy / 4))
		for (long k = 0; k < D.nz; k++)
		for (long j = 0; j < D.ny; j++)
		for (long i = 0; i < D.nx; i++) {
			const long c = cid(&D, i, j, k);
			/* face dB of the cell's six copies: the 12 edge terms of each
			   face in the order of or_step_periodic */
			double db[6];
			for (int side = 0; side < 2; side++) {
				const long a = i + side;    /* x-face at a, y-face at j + side, z-face at k + side */
				double r = 0;
				r += EZX(a, j + 1, k); r -= EZX(a, j - 1, k); r -= EYX(a, j, k + 1); r += EYX(a, j, k - 1);
				r += EZY(a, j + 1, k); r += EZY(a - 1, j + 1, k); r -= EZY(a, j, k); r -= EZY(a - 1, j, k);
				r -= EYZ(a, j, k + 1); r -= EYZ(a - 1, j, k + 1); r += EYZ(a, j, k); r += EYZ(a - 1, j, k);
				db[0 + side] = r;
				const long b = j + side;
				r = 0;
				r += EXY(i, b, k + 1); r -= EXY(i, b, k - 1); r -= EZY(i + 1, b, k); r += EZY(i - 1, b, k);
				r += EXZ(i, b, k + 1); r += EXZ(i, b - 1, k + 1); r -= EXZ(i, b, k); r -= EXZ(i, b - 1, k);
				r -= EZX(i + 1, b, k); r += EZX(i, b, k); r -= EZX(i + 1, b - 1, k); r += EZX(i, b - 1, k);
				db[2 + side] = r;
				const long z = k + side;
				r = 0;
				r += EYZ(i + 1, j, z); r -= EYZ(i - 1, j, z); r -= EXZ(i, j + 1, z); r += EXZ(i, j - 1, z);
				r += EYX(i + 1, j, z); r += EYX(i + 1, j, z - 1); r -= EYX(i, j, z); r -= EYX(i, j, z - 1);
				r -= EXY(i, j + 1, z); r -= EXY(i, j + 1, z - 1); r += EXY(i, j, z); r += EXY(i, j, z - 1);
				db[4 + side] = r;
			}
			for (int f = 0; f < 6; f++) fdb[(long)f * n + c] += db[f];
			/* flux slots: -d slot from the cell's own d-face, +d from the next */
			for (int d = 0; d < 3; d++) {
				const long cp = cid(&D, i + (d == 0), j + (d == 1), k + (d == 2));
				for (int q = 0; q < 8; q++) {
					SLOT(2 * d, q, c) += P[(d * 8 + q) * n + c];
					SLOT(2 * d + 1, q, c) += P[(d * 8 + q) * n + cp];
				}
			}
			if (sub % period[c] != 0) continue;

			/* update_mhd_state of a cell due at this substep */
			double rho = SC(0, c), m[3] = {SC(1, c), SC(2, c), SC(3, c)}, E = SC(4, c);
			double vb[3] = {SC(5, c), SC(6, c), SC(7, c)};
			for (int d = 0; d < 3; d++) rho += (SLOT(2 * d, RHO, c) - SLOT(2 * d + 1, RHO, c)) / len[d];
			for (int d = 0; d < 3; d++) {
				const double il = 1 / len[d];
				for (int q = 0; q < 3; q++)
					m[q] = m[q] + (SLOT(2 * d, M0 + q, c) - SLOT(2 * d + 1, M0 + q, c)) * il;
			}
			for (int d = 0; d < 3; d++) E += (SLOT(2 * d, NRJ, c) - SLOT(2 * d + 1, NRJ, c)) / len[d];
			for (int d = 0; d < 3; d++) {
				const double il = 1 / len[d];
				for (int q = 0; q < 3; q++)
					vb[q] = vb[q] + (SLOT(2 * d, B0 + q, c) - SLOT(2 * d + 1, B0 + q, c)) * il;
			}
			for (int d = 0; d < 3; d++) {
				SC(8 + 2 * d, c) += fdb[(long)(2 * d) * n + c] / area[d];
				SC(9 + 2 * d, c) += fdb[(long)(2 * d + 1) * n + c] / area[d];
			}
			for (int dir = 0; dir < 6; dir++) {
				for (int q = 0; q < 8; q++) SLOT(dir, q, c) = 0;
				fdb[(long)dir * n + c] = 0;
			}
			/* update_vol_B(s): constant-pressure refit from the own copies */
			SC(0, c) = rho;
			for (int q = 0; q < 3; q++) SC(1 + q, c) = m[q];
			if (rho > 0) {
				const double vb_new[3] = {0.5 * (SC(8, c) + SC(9, c)), 0.5 * (SC(10, c) + SC(11, c)),
					0.5 * (SC(12, c) + SC(13, c))};
				refit_energy(rho, m, &E, vb, vb_new, g->gamma, g->mu0);
				for (int q = 0; q < 3; q++) SC(5 + q, c) = vb_new[q];
			} else {
				for (int q = 0; q < 3; q++) SC(5 + q, c) = vb[q];
			}
			SC(4, c) = E;
		}
#undef PF
#undef EZX
#undef EYX
#undef EXY
#undef EZY
#undef EXZ
#undef EYZ
	}

	/* sync_magnetic_field: per face, the cell with the smaller index visits
	   first; each visit moves both copies by half their difference */
	for (int d = 0; d < 3; d++)
	for (long k = 0; k < D.nz; k++)
	for (long j = 0; j < D.ny; j++)
	for (long i = 0; i < D.nx; i++) {
		const long b = cid(&D, i, j, k);                                   /* -d copy */
		const long a = cid(&D, i - (d == 0), j - (d == 1), k - (d == 2));   /* +d copy */
		double* pa = &SC(9 + 2 * d, a);
		double* pb = &SC(8 + 2 * d, b);
		if (a < b) {
			double diff = *pa - *pb; *pa -= diff / 2; *pb += diff / 2;
			diff = *pb - *pa; *pb -= diff / 2; *pa += diff / 2;
		} else {
			double diff = *pb - *pa; *pb -= diff / 2; *pa += diff / 2;
			diff = *pa - *pb; *pa -= diff / 2; *pb += diff / 2;
		}
	}
	/* update_vol_B(0) of all cells, then back to the face-once layout */
	for (long c = 0; c < n; c++) {
		double rho = SC(0, c), m[3] = {SC(1, c), SC(2, c), SC(3, c)}, E = SC(4, c);
		const double vb[3] = {SC(5, c), SC(6, c), SC(7, c)};
		if (rho > 0) {
			const double vb_new[3] = {0.5 * (SC(8, c) + SC(9, c)), 0.5 * (SC(10, c) + SC(11, c)),
				0.5 * (SC(12, c) + SC(13, c))};
			refit_energy(rho, m, &E, vb, vb_new, g->gamma, g->mu0);
		}
		for (int q = 0; q < 4; q++) state[q * n + c] = q == 0 ? rho : m[q - 1];
		state[4 * n + c] = E;
		for (int d = 0; d < 3; d++) state[(5 + d) * n + c] = SC(8 + 2 * d, c);
	}
	*dt_out = total_dt;
	*max_period_out = maxp;
done:
	free(period); free(minsub); free(S); free(slot); free(fdb); free(P);
	return err;
}
#undef SC
#undef SLOT

/* apply_boundaries of a periodic grid: one refit with the face average */
void or_refit_periodic(const or_grid* g, double* state)
{
	const dims_t D = {g->nx, g->ny, g->nz, (long)g->nx * g->ny * g->nz};
	const long n = D.n;
	for (long k = 0; k < D.nz; k++)
	for (long j = 0; j < D.ny; j++)
	for (long i = 0; i < D.nx; i++) {
		const long c = cid(&D, i, j, k);
		double vb[3];
		vol_b(&D, state, i, j, k, vb);
		const double m[3] = {state[n + c], state[2 * n + c], state[3 * n + c]};
		refit_energy(state[c], m, &state[4 * n + c], vb, vb, g->gamma, g->mu0);
	}
}

int or_run_periodic_substep(const or_grid* g, int solver, double cfl, int substep_min_all,
	int substep_max_all, double t_start, double t_end, long max_steps, double* state,
	long* steps_taken, double* t_final, int* last_max_period)
{
	const long n = (long)g->nx * g->ny * g->nz;
	double* maxv = (double*)malloc(sizeof(double) * 3 * n);
	int* substep_min = (int*)malloc(sizeof(int) * n);
	int* substep_max = (int*)malloc(sizeof(int) * n);
	if (!maxv || !substep_min || !substep_max) { free(maxv); free(substep_min); free(substep_max); return OR_ERR_SOLVER; }
	for (long c = 0; c < n; c++) { substep_min[c] = substep_min_all; substep_max[c] = substep_max_all; }
	for (long f = 0; f < 3 * n; f++) maxv[f] = -1;
	or_refit_periodic(g, state);
	/* final init: the speeds are unknown, so dt = 0 and one substep */
	double dt;
	int mp = 0;
	int e = or_step_substep(g, solver, cfl, 0.0, substep_min, substep_max, state, maxv, &dt, &mp, NULL);
	double t = t_start;
	long steps = 0;
	while (!e && t < t_end && steps < max_steps) {
		e = or_step_substep(g, solver, cfl, t_end - t, substep_min, substep_max, state, maxv,
			&dt, &mp, NULL);
		if (e) break;
		or_refit_periodic(g, state);
		t += dt;
		steps++;
	}
	free(maxv);
	free(substep_min);
	free(substep_max);
	if (e) return e;
	if (steps_taken) *steps_taken = steps;
	if (t_final) *t_final = t;
	if (last_max_period) *last_max_period = mp;
	return OR_OK;
}

/* ---------------------------------------------------------------------- */
/* 1-D plain method of tests/mhd/reference.cpp                            */
/* ---------------------------------------------------------------------- */

/* solve_mhd, tests/mhd/reference.cpp:284-440; returns the next max dt */
static int plain1d_step(const or_plain1d* p, int solver, double dt, double* s,
	double* max_dt_out)
{
	const size_t N = p->ncells;
	const double zero[3] = {0, 0, 0};
	double* fl = (double*)calloc(8 * N, sizeof(double)); /* Mas_f, Mom_f, Nrj_f, Mag_f */
	double* ele = (double*)calloc(3 * N, sizeof(double));
	if (!fl || !ele) { free(fl); free(ele); return OR_ERR_SOLVER; }
	double max_dt = DBL_MAX;
#define S(q, c) s[(q) * N + (c)]
#define FL(q, c) fl[(q) * N + (c)]
#define EL(q, c) ele[(q) * N + (c)]
	for (size_t c = 0; c + 1 < N; c++) {
		double L[8], R[8], F[8], vel = -1;
		for (int q = 0; q < 8; q++) {
			L[q] = S(q, c);
			R[q] = S(q, c + 1);
		}
		const int e = or_flux(solver, L, R, zero, p->gamma, p->mu0, F, &vel);
		if (e) { free(fl); free(ele); return e; }
		const double fa = 1.0 * dt; /* face_area * dt */
		F[RHO] *= fa;
		for (int q = 0; q < 3; q++) F[M0 + q] = F[M0 + q] * fa;
		F[NRJ] *= fa;
		max_dt = smin(max_dt, p->cell_size / vel);
		if (c > 0) {
			FL(RHO, c) -= F[RHO];
			for (int q = 0; q < 3; q++) FL(M0 + q, c) = FL(M0 + q, c) - F[M0 + q];
			FL(NRJ, c) -= F[NRJ];
		}
		if (c < N - 2) {
			FL(RHO, c + 1) += F[RHO];
			for (int q = 0; q < 3; q++) FL(M0 + q, c + 1) = FL(M0 + q, c + 1) + F[M0 + q];
			FL(NRJ, c + 1) += F[NRJ];
		}
		/* upwind E = -v x B on the cell's +x edge, :393-404 */
		const double ic = 1 / S(RHO, c), in = 1 / S(RHO, c + 1);
		const double vc[3] = {S(M0, c) * ic, S(M1, c) * ic, S(M2, c) * ic};
		const double vn[3] = {S(M0, c + 1) * in, S(M1, c + 1) * in, S(M2, c + 1) * in};
		EL(0, c) = 0;
		if (vc[0] + vn[0] > 0) {
			EL(1, c) = vc[0] * S(B2, c) - vc[2] * S(B0, c);
			EL(2, c) = vc[1] * S(B0, c) - vc[0] * S(B1, c);
		} else {
			EL(1, c) = vn[0] * S(B2, c + 1) - vn[2] * S(B0, c);
			EL(2, c) = vn[1] * S(B0, c) - vn[0] * S(B1, c + 1);
		}
		FL(B0, c) = 0;
		FL(B1, c) = EL(2, c);
		FL(B2, c) = -EL(1, c);
		if (c > 0) {
			FL(B1, c) -= EL(2, c - 1);
			FL(B2, c) += EL(1, c - 1);
		}
		for (int q = 0; q < 3; q++) FL(B0 + q, c) = FL(B0 + q, c) * dt;
	}
	/* apply_fluxes, source/mhd/common.hpp:707-743 */
	const double inverse_volume = 1.0 / (1.0 * p->cell_size);
	int err = OR_OK;
	for (size_t c = 0; c < N; c++) {
		S(RHO, c) += FL(RHO, c) * inverse_volume;
		for (int q = 0; q < 3; q++) S(M0 + q, c) = S(M0 + q, c) + FL(M0 + q, c) * inverse_volume;
		S(NRJ, c) += FL(NRJ, c) * inverse_volume;
		for (int q = 0; q < 3; q++) S(B0 + q, c) = S(B0 + q, c) + FL(B0 + q, c) * inverse_volume;
		if (S(RHO, c) <= 0) { err = OR_ERR_DENSITY; break; }
		double U[8], pr;
		for (int q = 0; q < 8; q++) U[q] = S(q, c);
		or_get_pressure(U, p->gamma, p->mu0, &pr);
		if (pr <= 0) { err = OR_ERR_PRESSURE; break; }
	}
#undef S
#undef FL
#undef EL
	free(fl);
	free(ele);
	*max_dt_out = max_dt;
	return err;
}

/* main loop of tests/mhd/reference.cpp:830-877 */
int or_plain1d_run(const or_plain1d* p, int solver, double t_end, double* state,
	long* steps_taken)
{
	double max_dt = 0, t = 0;
	long steps = 0;
	while (t < t_end) {
		const double until_end = t_end - t;
		const double time_step = smin(p->cfl * max_dt, until_end);
		double next;
		const int e = plain1d_step(p, solver, time_step, state, &next);
		if (e) return e;
		max_dt = smin(DBL_MAX, next);
		t += time_step;
		steps++;
	}
	if (steps_taken) *steps_taken = steps;
	return OR_OK;
}

/* 7-point Gauss-Legendre rule (textbook nodes and weights), mapped to
   [a, b] and divided by b - a */
double or_gl7_average(double (*f)(double, void*), void* ctx, double a, double b)
{
	static const double x[4] = {
		0.0,
		0.405845151377397166906606412076961,
		0.741531185599394439863864773280788,
		0.949107912342758524526189684047851};
	static const double w[4] = {
		0.417959183673469387755102040816327,
		0.381830050505118944950369775488975,
		0.279705391489276667901467771423780,
		0.129484966168869693270611432679082};
	const double avg = (a + b) * 0.5, half = (b - a) * 0.5;
	double r = f(avg, ctx) * w[0];
	for (int i = 1; i < 4; i++) {
		const double fp = f(avg + half * x[i], ctx);
		const double fm = f(avg - half * x[i], ctx);
		r += (fp + fm) * w[i];
	}
	return r * half / (b - a);
}

int or_set_threads(int n)
{
#ifdef _OPENMP
	if (n > 0) omp_set_num_threads(n);
	return omp_get_max_threads();
#else
	(void)n;
	return 1;
#endif
}

/* ---------------------------------------------------------------------- */
/* divergence cleaning: periodic Poisson projection of the face field     */
/* ---------------------------------------------------------------------- */

/*
Replaces divergence::remove (source/divergence/remove.hpp:609-769, whose
Poisson solve is dccrg's, not present in the reference checkout) for the
staggered field: rhs = face divergence per cell, solve the periodic
7-point problem L phi = rhs by geometric multigrid, then
B_face -= face gradient of phi, so the new face divergence is the solver's
residual.  Levels halve every dimension that is even and >= 4; smoothing is
red-black Gauss-Seidel when every dimension of the level is even (colours
are then consistent across the periodic wrap), damped Jacobi (omega 0.8)
otherwise; restriction averages the children, prolongation is trilinear
in the coarsened dimensions.  Only dimensions whose spacing is within 1.5x
of the level's finest are coarsened (semi-coarsening for anisotropic cells).
*/
typedef struct {
	long n[3];
	double h[3];
	long ncell;
	double* phi;
	double* f;
	double* r;
} mg_level;

static inline long mg_idx(const mg_level* L, long i, long j, long k)
{
	return wrap(i, L->n[0]) + L->n[0] * (wrap(j, L->n[1]) + L->n[1] * wrap(k, L->n[2]));
}

/* (L phi)(i,j,k); dimensions of length 1 contribute nothing */
static double mg_apply(const mg_level* L, const double* phi, long i, long j, long k)
{
	const double c = phi[mg_idx(L, i, j, k)];
	double s = 0;
	if (L->n[0] > 1) s += (phi[mg_idx(L, i + 1, j, k)] - 2 * c + phi[mg_idx(L, i - 1, j, k)]) / (L->h[0] * L->h[0]);
	if (L->n[1] > 1) s += (phi[mg_idx(L, i, j + 1, k)] - 2 * c + phi[mg_idx(L, i, j - 1, k)]) / (L->h[1] * L->h[1]);
	if (L->n[2] > 1) s += (phi[mg_idx(L, i, j, k + 1)] - 2 * c + phi[mg_idx(L, i, j, k - 1)]) / (L->h[2] * L->h[2]);
	return s;
}

/* new value of phi at (i,j,k) from its neighbours: (sum nb/h^2 - f) / diag */
static double mg_relax_value(const mg_level* L, const double* phi, long i, long j, long k)
{
	double s = 0, diag = 0;
	if (L->n[0] > 1) {
		s += (phi[mg_idx(L, i + 1, j, k)] + phi[mg_idx(L, i - 1, j, k)]) / (L->h[0] * L->h[0]);
		diag += 2 / (L->h[0] * L->h[0]);
	}
	if (L->n[1] > 1) {
		s += (phi[mg_idx(L, i, j + 1, k)] + phi[mg_idx(L, i, j - 1, k)]) / (L->h[1] * L->h[1]);
		diag += 2 / (L->h[1] * L->h[1]);
	}
	if (L->n[2] > 1) {
		s += (phi[mg_idx(L, i, j, k + 1)] + phi[mg_idx(L, i, j, k - 1)]) / (L->h[2] * L->h[2]);
		diag += 2 / (L->h[2] * L->h[2]);
	}
	if (diag == 0) return phi[mg_idx(L, i, j, k)];
	return (s - L->f[mg_idx(L, i, j, k)]) / diag;
}

static int mg_rb_ok(const mg_level* L)
{
	for (int d = 0; d < 3; d++) if (L->n[d] > 1 && L->n[d] % 2) return 0;
	return 1;
}

static void mg_smooth(mg_level* L, int sweeps)
{
	if (mg_rb_ok(L)) {
		for (int s = 0; s < sweeps; s++) {
			for (int color = 0; color < 2; color++) {
				#pragma omp parallel for collapse(2) schedule(static)
				for (long k = 0; k < L->n[2]; k++)
				for (long j = 0; j < L->n[1]; j++)
				for (long i = 0; i < L->n[0]; i++) {
					if (((i + j + k) & 1) != color) continue;
					L->phi[mg_idx(L, i, j, k)] = mg_relax_value(L, L->phi, i, j, k);
				}
			}
		}
	} else {
		for (int s = 0; s < sweeps; s++) {
			#pragma omp parallel for collapse(2) schedule(static)
			for (long k = 0; k < L->n[2]; k++)
			for (long j = 0; j < L->n[1]; j++)
			for (long i = 0; i < L->n[0]; i++) {
				const long c = mg_idx(L, i, j, k);
				L->r[c] = L->phi[c] + 0.8 * (mg_relax_value(L, L->phi, i, j, k) - L->phi[c]);
			}
			memcpy(L->phi, L->r, sizeof(double) * L->ncell);
		}
	}
}

static void mg_residual(mg_level* L)
{
	#pragma omp parallel for collapse(2) schedule(static)
	for (long k = 0; k < L->n[2]; k++)
	for (long j = 0; j < L->n[1]; j++)
	for (long i = 0; i < L->n[0]; i++) {
		const long c = mg_idx(L, i, j, k);
		L->r[c] = L->f[c] - mg_apply(L, L->phi, i, j, k);
	}
}

/* coarse f = average of the fine residual over the children */
static void mg_restrict(const mg_level* F, mg_level* C)
{
	long ratio[3];
	for (int d = 0; d < 3; d++) ratio[d] = F->n[d] / C->n[d];
	const double w = 1.0 / (double)(ratio[0] * ratio[1] * ratio[2]);
	#pragma omp parallel for collapse(2) schedule(static)
	for (long k = 0; k < C->n[2]; k++)
	for (long j = 0; j < C->n[1]; j++)
	for (long i = 0; i < C->n[0]; i++) {
		double s = 0;
		for (long c = 0; c < ratio[2]; c++)
		for (long b = 0; b < ratio[1]; b++)
		for (long a = 0; a < ratio[0]; a++) {
			s += F->r[mg_idx(F, i * ratio[0] + a, j * ratio[1] + b, k * ratio[2] + c)];
		}
		C->f[mg_idx(C, i, j, k)] = s * w;
		C->phi[mg_idx(C, i, j, k)] = 0;
	}
}

/* fine phi += trilinear interpolation of the coarse phi; in a coarsened
   dimension a child at fine index 2I + a takes 3/4 of coarse cell I and 1/4
   of its neighbour on the child's side */
static void mg_prolong_add(mg_level* F, const mg_level* C)
{
	long ratio[3];
	for (int d = 0; d < 3; d++) ratio[d] = F->n[d] / C->n[d];
	#pragma omp parallel for collapse(2) schedule(static)
	for (long k = 0; k < F->n[2]; k++)
	for (long j = 0; j < F->n[1]; j++)
	for (long i = 0; i < F->n[0]; i++) {
		const long idx[3] = {i, j, k};
		long ci[3][2];
		double cw[3][2];
		int cnt[3];
		for (int d = 0; d < 3; d++) {
			if (ratio[d] == 2) {
				const long I = idx[d] / 2, side = (idx[d] % 2) ? 1 : -1;
				ci[d][0] = I;
				cw[d][0] = 0.75;
				ci[d][1] = I + side;
				cw[d][1] = 0.25;
				cnt[d] = 2;
			} else {
				ci[d][0] = idx[d];
				cw[d][0] = 1.0;
				cnt[d] = 1;
			}
		}
		double s = 0;
		for (int c = 0; c < cnt[2]; c++)
		for (int b = 0; b < cnt[1]; b++)
		for (int a = 0; a < cnt[0]; a++) {
			s += cw[0][a] * cw[1][b] * cw[2][c] * C->phi[mg_idx(C, ci[0][a], ci[1][b], ci[2][c])];
		}
		F->phi[mg_idx(F, i, j, k)] += s;
	}
}

static void mg_vcycle(mg_level* lv, int l, int nlev)
{
	if (l == nlev - 1) {
		mg_smooth(&lv[l], 60);
		return;
	}
	mg_smooth(&lv[l], 2);
	mg_residual(&lv[l]);
	mg_restrict(&lv[l], &lv[l + 1]);
	mg_vcycle(lv, l + 1, nlev);
	mg_prolong_add(&lv[l], &lv[l + 1]);
	mg_smooth(&lv[l], 2);
}

/* 2-norm of a - mean(a): the mean lies in the periodic Laplacian's null
   space, so it is excluded from the convergence test */
static double l2_centred(const double* a, long n)
{
	double m = 0;
	for (long c = 0; c < n; c++) m += a[c];
	m /= (double)n;
	double s = 0;
	for (long c = 0; c < n; c++) s += (a[c] - m) * (a[c] - m);
	return sqrt(s);
}

int or_poisson_periodic(const int dims[3], const double h[3], const double* rhs,
	double* phi, double tol, double atol, int max_cycles, double* rel_residual,
	int* cycles)
{
	mg_level lv[32];
	int nlev = 0;
	long n[3] = {dims[0], dims[1], dims[2]};
	double hh[3] = {h[0], h[1], h[2]};
	for (;;) {
		mg_level* L = &lv[nlev++];
		for (int d = 0; d < 3; d++) {
			L->n[d] = n[d];
			L->h[d] = hh[d];
		}
		L->ncell = n[0] * n[1] * n[2];
		L->phi = (double*)calloc(L->ncell, sizeof(double));
		L->f = (double*)calloc(L->ncell, sizeof(double));
		L->r = (double*)calloc(L->ncell, sizeof(double));
		/* semi-coarsening: halve the even dimensions whose spacing is within
		   1.5x of the finest active spacing, so coarse levels tend to
		   isotropy and point smoothing keeps working */
		double hmin = 1e300;
		for (int d = 0; d < 3; d++) if (n[d] > 1 && hh[d] < hmin) hmin = hh[d];
		int coarsened = 0;
		for (int d = 0; d < 3; d++) {
			if (n[d] % 2 == 0 && n[d] >= 4 && hh[d] <= 1.5 * hmin) {
				n[d] /= 2;
				hh[d] *= 2;
				coarsened = 1;
			}
		}
		if (!coarsened || L->ncell <= 64 || nlev == 32) break;
	}
	mg_level* F = &lv[0];
	/* the face divergence of a periodic field sums to zero by telescoping;
	   its roundoff-sized mean is left in the data (it only drifts phi by a
	   constant) and excluded from the norms of the convergence test */
	for (long c = 0; c < F->ncell; c++) F->f[c] = rhs[c];
	const double fnorm = l2_centred(F->f, F->ncell);
	const double target = tol * fnorm > atol ? tol * fnorm : atol;
	double rel = 0, rnorm = fnorm;
	int it = 0;
	if (fnorm > target) {
		for (it = 1; it <= max_cycles; it++) {
			mg_vcycle(lv, 0, nlev);
			mg_residual(F);
			rnorm = l2_centred(F->r, F->ncell);
			if (rnorm <= target) break;
		}
		if (it > max_cycles) it = max_cycles;
	}
	rel = fnorm > 0 ? rnorm / fnorm : 0;
	memcpy(phi, F->phi, sizeof(double) * F->ncell);
	for (int l = 0; l < nlev; l++) {
		free(lv[l].phi);
		free(lv[l].f);
		free(lv[l].r);
	}
	if (rel_residual) *rel_residual = rel;
	if (cycles) *cycles = it;
	return rnorm > target ? OR_ERR_SOLVER : OR_OK;
}

int or_clean_divergence(const or_grid* g, double* s, double tol, int max_cycles,
	double* rel_residual, int* cycles)
{
	const dims_t D = {g->nx, g->ny, g->nz, (long)g->nx * g->ny * g->nz};
	const long n = D.n;
	double* rhs = (double*)calloc(n, sizeof(double));
	double* phi = (double*)calloc(n, sizeof(double));
	double* vb_old = (double*)calloc(3 * n, sizeof(double));
	for (long k = 0; k < D.nz; k++)
	for (long j = 0; j < D.ny; j++)
	for (long i = 0; i < D.nx; i++) {
		const long c = cid(&D, i, j, k);
		rhs[c] = (s[5 * n + cid(&D, i + 1, j, k)] - s[5 * n + c]) / g->dx
			+ (s[6 * n + cid(&D, i, j + 1, k)] - s[6 * n + c]) / g->dy
			+ (s[7 * n + cid(&D, i, j, k + 1)] - s[7 * n + c]) / g->dz;
		double vb[3];
		vol_b(&D, s, i, j, k, vb);
		for (int q = 0; q < 3; q++) vb_old[q * n + c] = vb[q];
	}
	const int dims[3] = {g->nx, g->ny, g->nz};
	const double h[3] = {g->dx, g->dy, g->dz};
	/* roundoff floor of the face divergence: 8 eps || sum_d 2|B_d| / h_d ||_2 */
	double fl = 0;
	for (long c = 0; c < n; c++) {
		const double t = 2 * fabs(s[5 * n + c]) / g->dx + 2 * fabs(s[6 * n + c]) / g->dy
			+ 2 * fabs(s[7 * n + c]) / g->dz;
		fl += t * t;
	}
	const double atol = 8 * DBL_EPSILON * sqrt(fl);
	const int e = or_poisson_periodic(dims, h, rhs, phi, tol, atol, max_cycles, rel_residual, cycles);
	/* B_face -= grad phi on faces; dimensions of length 1 have no gradient */
	for (long k = 0; k < D.nz; k++)
	for (long j = 0; j < D.ny; j++)
	for (long i = 0; i < D.nx; i++) {
		const long c = cid(&D, i, j, k);
		if (D.nx > 1) s[5 * n + c] -= (phi[c] - phi[cid(&D, i - 1, j, k)]) / g->dx;
		if (D.ny > 1) s[6 * n + c] -= (phi[c] - phi[cid(&D, i, j - 1, k)]) / g->dy;
		if (D.nz > 1) s[7 * n + c] -= (phi[c] - phi[cid(&D, i, j, k - 1)]) / g->dz;
	}
	/* constant-thermal-pressure energy refit for the new cell field */
	for (long k = 0; k < D.nz; k++)
	for (long j = 0; j < D.ny; j++)
	for (long i = 0; i < D.nx; i++) {
		const long c = cid(&D, i, j, k);
		double vb_new[3];
		vol_b(&D, s, i, j, k, vb_new);
		const double vo[3] = {vb_old[c], vb_old[n + c], vb_old[2 * n + c]};
		const double m[3] = {s[n + c], s[2 * n + c], s[3 * n + c]};
		refit_energy(s[c], m, &s[4 * n + c], vo, vb_new, g->gamma, g->mu0);
	}
	free(rhs);
	free(phi);
	free(vb_old);
	return e;
}

/* ====================================================================== */
/* Solar-wind box: non-periodic uniform grid with a one-cell boundary     */
/* layer, inflow wall, outflow copies, planet cells and a dipole field    */
/* ====================================================================== */

/*
State of a box cell, 14 doubles (component-major, 14 * n), the reference's
per-cell record for this path (source/mhd/variables.hpp:212-234):
    0 rho, 1-3 m, 4 E, 5-7 Vol_B, 8-13 Face_B(-x, +x, -y, +y, -z, +z)
Each cell keeps its own copy of each of its six faces, as the reference
does, so the averaging of sync_magnetic_field is restated copy by copy.
*/
#define BQ_RHO 0
#define BQ_M 1
#define BQ_E 4
#define BQ_VB 5
#define BQ_F 8

static inline long bid(const or_box* b, long i, long j, long k)
{
	return i + b->nx * (j + (long)b->ny * k);
}

static inline int in_box(const or_box* b, long i, long j, long k)
{
	return i >= 0 && j >= 0 && k >= 0 && i < b->nx && j < b->ny && k < b->nz;
}

/* cell min / max / centre coordinates of a level-0 cell */
static inline double cmin(const or_box* b, int d, long idx) { return b->start[d] + (double)idx * b->len[d]; }
static inline double cmax(const or_box* b, int d, long idx) { return b->start[d] + (double)(idx + 1) * b->len[d]; }
static inline double ccen(const or_box* b, int d, long idx) { return b->start[d] + ((double)idx + 0.5) * b->len[d]; }

/* at_inner_boundary, source/grid/solar_wind_box.hpp:66-117: (inside, outside) */
static void at_inner_boundary(const or_box* b, long i, long j, long k, int* inside, int* outside)
{
	const double sx = cmin(b, 0, i), sy = cmin(b, 1, j), sz = cmin(b, 2, k);
	const double ex = cmax(b, 0, i), ey = cmax(b, 1, j), ez = cmax(b, 2, k);
	const double r[8] = {
		sqrt(sx * sx + sy * sy + sz * sz), sqrt(sx * sx + sy * sy + ez * ez),
		sqrt(sx * sx + ey * ey + sz * sz), sqrt(sx * sx + ey * ey + ez * ez),
		sqrt(ex * ex + sy * sy + sz * sz), sqrt(ex * ex + sy * sy + ez * ez),
		sqrt(ex * ex + ey * ey + sz * sz), sqrt(ex * ex + ey * ey + ez * ez)};
	const double R = b->inner_radius;
	int in = 0, out = 0;
	for (int q = 0; q < 8; q++) {
		if (r[q] <= R) in = 1;
		if (r[q] > R) out = 1;
	}
	if (sx <= 0 && ex >= 0 && sy <= 0 && ey >= 0 && sz <= 0 && ez >= 0) in = 1;
	*inside = in;
	*outside = out;
}

int or_box_classify(const or_box* b, signed char* ctype, unsigned char* rule)
{
	const long n = (long)b->nx * b->ny * b->nz;
	const int sw_dim = abs(b->sw_dir) - 1;
	const long dims[3] = {b->nx, b->ny, b->nz};
	if (b->sw_dir == 0 || sw_dim > 2) return OR_ERR_SOLVER;
	for (long c = 0; c < n; c++) {
		ctype[c] = 1;
		rule[c] = OR_RULE_NONE;
	}
	/* prepare_grid, source/grid/solar_wind_box.hpp:262-367: inflow wall,
	   then face / edge / vertex cells of the outflow walls, each copying
	   from its interior neighbour (offset opposite to its wall sides) */
	for (long k = 0; k < b->nz; k++)
	for (long j = 0; j < b->ny; j++)
	for (long i = 0; i < b->nx; i++) {
		const long idx[3] = {i, j, k};
		const long c = bid(b, i, j, k);
		int side[3];
		for (int d = 0; d < 3; d++) {
			side[d] = idx[d] == 0 ? -1 : (idx[d] == dims[d] - 1 ? +1 : 0);
			if (dims[d] == 1) side[d] = 0;
		}
		if (side[0] || side[1] || side[2]) {
			ctype[c] = 0;
			rule[c] = (unsigned char)(1 + (1 - side[0]) + 3 * (1 - side[1]) + 9 * (1 - side[2]));
		}
		const int sws = b->sw_dir < 0 ? 0 : (int)(dims[sw_dim] - 1);
		if (idx[sw_dim] == sws) {
			ctype[c] = 0;
			rule[c] = OR_RULE_SW;
		}
	}
	/* planet cells, :369-406 */
	for (long k = 0; k < b->nz; k++)
	for (long j = 0; j < b->ny; j++)
	for (long i = 0; i < b->nx; i++) {
		int in, out;
		at_inner_boundary(b, i, j, k, &in, &out);
		if (in) {
			ctype[bid(b, i, j, k)] = 0;
			rule[bid(b, i, j, k)] = OR_RULE_PLANET;
		}
	}
	for (long k = 0; k < b->nz; k++)
	for (long j = 0; j < b->ny; j++)
	for (long i = 0; i < b->nx; i++) {
		const long c = bid(b, i, j, k);
		if (rule[c] != OR_RULE_PLANET) continue;
		int have_normal = 0;
		/* face and edge neighbours: offsets with one or two non-zero parts */
		for (int dz = -1; dz <= 1; dz++)
		for (int dy = -1; dy <= 1; dy++)
		for (int dx = -1; dx <= 1; dx++) {
			const int nz_ = (dx != 0) + (dy != 0) + (dz != 0);
			if (nz_ == 0 || nz_ == 3) continue;
			if (!in_box(b, i + dx, j + dy, k + dz)) continue;
			if (ctype[bid(b, i + dx, j + dy, k + dz)] == 1) have_normal = 1;
		}
		if (!have_normal) {
			ctype[c] = -1;
			rule[c] = OR_RULE_NONE;
		}
	}
	return OR_OK;
}

/* Background_Magnetic_Field::get_background_field for a constant field and
   point dipoles, source/background_magnetic_field.hpp:331-494 */
static void bg_field(const double pos[3], const double constant[3], double min_distance,
	int ndip, const double* moment, const double* dpos, double mu0, double out[3])
{
	double r3[3] = {constant[0], constant[1], constant[2]};
	for (int q = 0; q < ndip; q++) {
		const double* m = moment + 3 * q;
		const double r[3] = {pos[0] + (-dpos[3 * q]), pos[1] + (-dpos[3 * q + 1]), pos[2] + (-dpos[3 * q + 2])};
		const double r1 = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
		if (r1 < min_distance) {
			const double f = mu0 / (2 * M_PI * pow(min_distance, 3));
			for (int d = 0; d < 3; d++) r3[d] = r3[d] + m[d] * f;
		} else {
			const double inv = 1 / r1;
			const double u[3] = {r[0] * inv, r[1] * inv, r[2] * inv};
			const double dp = m[0] * u[0] + m[1] * u[1] + m[2] * u[2];
			const double f = mu0 / (4 * M_PI) / (r1 * r1 * r1);
			for (int d = 0; d < 3; d++) {
				const double proj = dp * u[d];
				r3[d] = r3[d] + (3 * proj + (-m[d])) * f;
			}
		}
	}
	out[0] = r3[0];
	out[1] = r3[1];
	out[2] = r3[2];
}

/* Bg_B(cell)(+d) at the centre of each cell's positive faces,
   source/mhd/solar_wind_box.hpp:120-171 */
void or_box_background(const or_box* b, const double constant[3], double min_distance,
	int ndip, const double* moment, const double* dpos, double* bg)
{
	const long n = (long)b->nx * b->ny * b->nz;
	for (long k = 0; k < b->nz; k++)
	for (long j = 0; j < b->ny; j++)
	for (long i = 0; i < b->nx; i++) {
		const long c = bid(b, i, j, k);
		const long idx[3] = {i, j, k};
		for (int d = 0; d < 3; d++) {
			double p[3];
			for (int e = 0; e < 3; e++) p[e] = ccen(b, e, idx[e]);
			p[d] = cmax(b, d, idx[d]);
			double v[3];
			bg_field(p, constant, min_distance, ndip, moment, dpos, b->mu0, v);
			for (int e = 0; e < 3; e++) bg[(3 * d + e) * n + c] = v[e];
		}
	}
}

static int box_pressure(const or_box* b, const double* s, long n, long c, double* p)
{
	const double U[8] = {s[c], s[n + c], s[2 * n + c], s[3 * n + c], s[4 * n + c],
		s[5 * n + c], s[6 * n + c], s[7 * n + c]};
	return or_get_pressure(U, b->gamma, b->mu0, p);
}

/* update_vol_B with constant thermal pressure for one cell,
   source/mhd/solve.hpp:1646-1672 */
static void box_refit(const or_box* b, double* s, long n, long c)
{
	if (s[c] <= 0) return;
	double p;
	box_pressure(b, s, n, c, &p);
	for (int d = 0; d < 3; d++) {
		s[(BQ_VB + d) * n + c] = 0.5 * (s[(BQ_F + 2 * d) * n + c] + s[(BQ_F + 2 * d + 1) * n + c]);
	}
	const double inv = 1 / s[c];
	const double v[3] = {s[n + c] * inv, s[2 * n + c] * inv, s[3 * n + c] * inv};
	const double B[3] = {s[5 * n + c], s[6 * n + c], s[7 * n + c]};
	s[4 * n + c] = or_total_energy(s[c], v, p, B, b->gamma, b->mu0);
}

/* sync_magnetic_field, source/common_functions.hpp:184-269: the two copies
   of each face between cells of type >= 0 are averaged, visited first from
   the cell with the smaller index, then from the other */
void or_box_sync(const or_box* b, double* s)
{
	const long n = (long)b->nx * b->ny * b->nz;
	for (long k = 0; k < b->nz; k++)
	for (long j = 0; j < b->ny; j++)
	for (long i = 0; i < b->nx; i++) {
		const long c = bid(b, i, j, k);
		if (b->ctype[c] < 0) continue;
		for (int d = 0; d < 3; d++) {
			const long ii = i + (d == 0), jj = j + (d == 1), kk = k + (d == 2);
			if (!in_box(b, ii, jj, kk)) continue;
			const long nb = bid(b, ii, jj, kk);
			if (b->ctype[nb] < 0) continue;
			double* a = &s[(BQ_F + 2 * d + 1) * n + c];
			double* o = &s[(BQ_F + 2 * d) * n + nb];
			const double diff = *a - *o;
			*a -= diff / 2;
			*o += diff / 2;
			const double diff2 = *o - *a;
			*o -= diff2 / 2;
			*a += diff2 / 2;
		}
	}
}

/* apply_boundaries_sw_box + apply_solar_wind_boundaries,
   source/mhd/solar_wind_box.hpp:219-577 (one solar-wind state) */
int or_box_boundaries(const or_box* b, double* s)
{
	const long n = (long)b->nx * b->ny * b->nz;
	const int sw_dim = abs(b->sw_dir) - 1;
	/* outflow copies (face, edge and vertex cells): they read only normal
	   cells, so their order does not matter */
	for (long k = 0; k < b->nz; k++)
	for (long j = 0; j < b->ny; j++)
	for (long i = 0; i < b->nx; i++) {
		const long c = bid(b, i, j, k);
		const int r = b->rule[c];
		if (r < 1 || r > 27) continue;
		const int o[3] = {(r - 1) % 3 - 1, ((r - 1) / 3) % 3 - 1, (r - 1) / 9 - 1};
		const long src = bid(b, i + o[0], j + o[1], k + o[2]);
		if (b->ctype[src] != 1) return OR_ERR_SOLVER;
		double p;
		const int e = box_pressure(b, s, n, src, &p);
		if (e) return e;
		s[c] = s[src];
		for (int q = 1; q <= 3; q++) s[q * n + c] = s[q * n + src];
		for (int d = 0; d < 3; d++) {
			const double lo = s[(BQ_F + 2 * d) * n + src], hi = s[(BQ_F + 2 * d + 1) * n + src];
			/* mirrored across the walls the cell touches, copied along the others */
			s[(BQ_F + 2 * d) * n + c] = o[d] ? hi : lo;
			s[(BQ_F + 2 * d + 1) * n + c] = o[d] ? lo : hi;
		}
		for (int d = 0; d < 3; d++) {
			s[(BQ_VB + d) * n + c] = 0.5 * (s[(BQ_F + 2 * d) * n + c] + s[(BQ_F + 2 * d + 1) * n + c]);
		}
		const double inv = 1 / s[c];
		const double v[3] = {s[n + c] * inv, s[2 * n + c] * inv, s[3 * n + c] * inv};
		const double B[3] = {s[5 * n + c], s[6 * n + c], s[7 * n + c]};
		s[4 * n + c] = or_total_energy(s[c], v, p, B, b->gamma, b->mu0);
	}
	/* planet cells, :520-541; the energy is computed with the momentum in
	   place of the velocity, as the reference does */
	for (long c = 0; c < n; c++) {
		if (b->rule[c] != OR_RULE_PLANET) continue;
		const double mass = b->inner_n * b->mp;
		for (int q = 0; q < 6; q++) s[(BQ_F + q) * n + c] = 0;
		for (int d = 0; d < 3; d++) s[(BQ_VB + d) * n + c] = 0;
		s[c] = mass;
		for (int d = 0; d < 3; d++) s[(1 + d) * n + c] = mass * b->inner_v[d];
		const double m[3] = {s[n + c], s[2 * n + c], s[3 * n + c]};
		const double B[3] = {0, 0, 0};
		s[4 * n + c] = or_total_energy(mass, m, b->inner_p, B, b->gamma, b->mu0);
	}
	/* inflow cells and the face each shares with its interior neighbour,
	   :219-299 */
	const double mass = b->sw_n * b->mp;
	for (long k = 0; k < b->nz; k++)
	for (long j = 0; j < b->ny; j++)
	for (long i = 0; i < b->nx; i++) {
		const long c = bid(b, i, j, k);
		if (b->rule[c] != OR_RULE_SW) continue;
		s[c] = mass;
		for (int d = 0; d < 3; d++) s[(1 + d) * n + c] = mass * b->sw_v[d];
		for (int d = 0; d < 3; d++) {
			s[(BQ_VB + d) * n + c] = b->sw_B[d];
			s[(BQ_F + 2 * d) * n + c] = b->sw_B[d];
			s[(BQ_F + 2 * d + 1) * n + c] = b->sw_B[d];
		}
		s[4 * n + c] = or_total_energy(mass, b->sw_v, b->sw_p, b->sw_B, b->gamma, b->mu0);

		const int sg = b->sw_dir > 0 ? 1 : -1;
		const long ni = i - sg * (sw_dim == 0), nj = j - sg * (sw_dim == 1), nk = k - sg * (sw_dim == 2);
		if (!in_box(b, ni, nj, nk)) continue;
		const long nb = bid(b, ni, nj, nk);
		const double inv = 1 / s[nb];
		const double v[3] = {s[n + nb] * inv, s[2 * n + nb] * inv, s[3 * n + nb] * inv};
		double p;
		const int e = box_pressure(b, s, n, nb, &p);
		if (e) return e;
		const int toward = BQ_F + 2 * sw_dim + (sg > 0), away = BQ_F + 2 * sw_dim + (sg < 0);
		s[toward * n + nb] = b->sw_B[sw_dim];
		s[(BQ_VB + sw_dim) * n + nb] = s[toward * n + nb] / 2 + s[away * n + nb] / 2;
		const double B[3] = {s[5 * n + nb], s[6 * n + nb], s[7 * n + nb]};
		s[4 * n + nb] = or_total_energy(s[nb], v, p, B, b->gamma, b->mu0);
	}
	return OR_OK;
}

int or_box_init(const or_box* b, double* s)
{
	const long n = (long)b->nx * b->ny * b->nz;
	/* initialize_plasma, source/mhd/solar_wind_box.hpp:62-205 */
	const double end_x = b->start[0] + (double)b->nx * b->len[0];
	const double mass = b->sw_n * b->mp;
	for (long k = 0; k < b->nz; k++)
	for (long j = 0; j < b->ny; j++)
	for (long i = 0; i < b->nx; i++) {
		const long c = bid(b, i, j, k);
		const double ratio = ccen(b, 0, i) / end_x;
		const double vf = (0.0 < ratio) ? ratio : 0.0;
		const double v[3] = {vf * b->sw_v[0], vf * b->sw_v[1], vf * b->sw_v[2]};
		s[c] = mass;
		for (int d = 0; d < 3; d++) s[(1 + d) * n + c] = mass * v[d];
		for (int d = 0; d < 3; d++) {
			s[(BQ_VB + d) * n + c] = b->sw_B[d];
			s[(BQ_F + 2 * d) * n + c] = b->sw_B[d];
			s[(BQ_F + 2 * d + 1) * n + c] = b->sw_B[d];
		}
		s[4 * n + c] = or_total_energy(mass, v, b->sw_p, b->sw_B, b->gamma, b->mu0);
	}
	/* update_vol_B(constant pressure) over cells of type >= 0,
	   tests/mhd/solar_wind_box.cpp:441-446 */
	for (long c = 0; c < n; c++) if (b->ctype[c] >= 0) box_refit(b, s, n, c);
	const int e = or_box_boundaries(b, s);
	if (e) return e;
	or_box_sync(b, s);
	return OR_OK;
}

int or_box_step(const or_box* b, int solver, double dt, double* s, double smax_dim[3],
	long zero_dim[3], long* bad_face)
{
	const long n = (long)b->nx * b->ny * b->nz;
	const double len[3] = {b->len[0], b->len[1], b->len[2]};
	double* P = (double*)calloc(24 * n, sizeof(double));
	if (!P) return OR_ERR_SOLVER;
	int err = OR_OK;
	long bad = -1;
	double sm[3] = {0, 0, 0};
	long zc[3] = {0, 0, 0};
	/* flux pass over faces whose both cells have type >= 0
	   (source/mhd/solve.hpp:262-393); the face is owned by its positive-side
	   cell c, between L = c - e_d and R = c */
	or_fc = 0;
	for (int d = 0; d < 3 && !err; d++) {
		for (long k = 0; k < b->nz; k++)
		for (long j = 0; j < b->ny; j++)
		for (long i = 0; i < b->nx; i++) {
			const long li = i - (d == 0), lj = j - (d == 1), lk = k - (d == 2);
			if (!in_box(b, li, lj, lk)) continue;
			const long c = bid(b, i, j, k), l = bid(b, li, lj, lk);
			if (b->ctype[c] < 0 || b->ctype[l] < 0) continue;
			or_fc++;
			double L[8], R[8], bgr[3] = {0, 0, 0}, F[8], sp = 0;
			const long cc[2] = {l, c};
			double* U[2] = {L, R};
			for (int q = 0; q < 2; q++) {
				const double m[3] = {s[n + cc[q]], s[2 * n + cc[q]], s[3 * n + cc[q]]};
				const double B[3] = {s[5 * n + cc[q]], s[6 * n + cc[q]], s[7 * n + cc[q]]};
				U[q][RHO] = s[cc[q]];
				U[q][NRJ] = s[4 * n + cc[q]];
				rot_to_face(d, m, U[q] + M0);
				rot_to_face(d, B, U[q] + B0);
			}
			if (b->bg) {
				const double bb[3] = {b->bg[(3 * d) * n + l], b->bg[(3 * d + 1) * n + l], b->bg[(3 * d + 2) * n + l]};
				rot_to_face(d, bb, bgr);
			}
			const int e = or_flux(solver, L, R, bgr, b->gamma, b->mu0, F, &sp);
			if (e) {
				if (!err || d * n + c < bad) {
					err = e;
					bad = d * n + c;
				}
				continue;
			}
			double Fg[8];
			Fg[RHO] = F[RHO];
			Fg[NRJ] = F[NRJ];
			rot_to_global(d, F + M0, Fg + M0);
			rot_to_global(d, F + B0, Fg + B0);
			for (int q = 0; q < 8; q++) P[(d * 8 + q) * n + c] = dt * Fg[q];
			if (sp <= 0) zc[d]++;
			if (sp > sm[d]) sm[d] = sp;
		}
	}
	if (bad_face) *bad_face = err ? bad : -1;
	if (smax_dim) for (int d = 0; d < 3; d++) smax_dim[d] = sm[d];
	if (zero_dim) for (int d = 0; d < 3; d++) zero_dim[d] = zc[d];
	if (err) {
		free(P);
		return err;
	}
	const double dx = len[0], dy = len[1], dz = len[2];
	/* flux of a face that was not computed (outside the box or next to a
	   cell of type -1) is zero */
#define PB(d, q, i, j, k) (in_box(b, (i), (j), (k)) ? P[((d) * 8 + (q)) * n + bid(b, (i), (j), (k))] : 0.0)
#define EZX(i, j, k) (+(PB(0, 6, i, j, k) * dz / 4))
#define EYX(i, j, k) (-(PB(0, 7, i, j, k) * dy / 4))
#define EXY(i, j, k) (+(PB(1, 7, i, j, k) * dx / 4))
#define EZY(i, j, k) (-(PB(1, 5, i, j, k) * dz / 4))
#define EXZ(i, j, k) (-(PB(2, 6, i, j, k) * dx / 4))
#define EYZ(i, j, k) (+(PB(2, 5, i, j, k) * dy / 4))
	double* out = (double*)malloc(sizeof(double) * 14 * n);
	if (!out) {
		free(P);
		return OR_ERR_SOLVER;
	}
	memcpy(out, s, sizeof(double) * 14 * n);
	for (long k = 0; k < b->nz; k++)
	for (long j = 0; j < b->ny; j++)
	for (long i = 0; i < b->nx; i++) {
		const long c = bid(b, i, j, k);
		if (b->ctype[c] <= 0) continue;
		/* Face_dB of the cell's six face copies (the 12 edge terms of each
		   face, in the order of the periodic step), update_mhd_state's
		   face change, :1589-1593 */
		double dB[3][2];
		for (int side = 0; side < 2; side++) {
			const long I = i + side, J = j + side, K = k + side;
			double r = 0;
			r += EZX(I, j + 1, k); r -= EZX(I, j - 1, k); r -= EYX(I, j, k + 1); r += EYX(I, j, k - 1);
			r += EZY(I, j + 1, k); r += EZY(I - 1, j + 1, k); r -= EZY(I, j, k); r -= EZY(I - 1, j, k);
			r -= EYZ(I, j, k + 1); r -= EYZ(I - 1, j, k + 1); r += EYZ(I, j, k); r += EYZ(I - 1, j, k);
			dB[0][side] = r;
			r = 0;
			r += EXY(i, J, k + 1); r -= EXY(i, J, k - 1); r -= EZY(i + 1, J, k); r += EZY(i - 1, J, k);
			r += EXZ(i, J, k + 1); r += EXZ(i, J - 1, k + 1); r -= EXZ(i, J, k); r -= EXZ(i, J - 1, k);
			r -= EZX(i + 1, J, k); r += EZX(i, J, k); r -= EZX(i + 1, J - 1, k); r += EZX(i, J - 1, k);
			dB[1][side] = r;
			r = 0;
			r += EYZ(i + 1, j, K); r -= EYZ(i - 1, j, K); r -= EXZ(i, j + 1, K); r += EXZ(i, j - 1, K);
			r += EYX(i + 1, j, K); r += EYX(i + 1, j, K - 1); r -= EYX(i, j, K); r -= EYX(i, j, K - 1);
			r -= EXY(i, j + 1, K); r -= EXY(i, j + 1, K - 1); r += EXY(i, j, K); r += EXY(i, j, K - 1);
			dB[2][side] = r;
		}
		const double area[3] = {dy * dz, dx * dz, dx * dy};
		for (int d = 0; d < 3; d++)
		for (int side = 0; side < 2; side++) {
			out[(BQ_F + 2 * d + side) * n + c] = s[(BQ_F + 2 * d + side) * n + c] + dB[d][side] / area[d];
		}
		/* conserved update, :1551-1587 */
		double lo[3][8], hi[3][8];
		for (int q = 0; q < 8; q++) {
			lo[0][q] = 0.0 + PB(0, q, i, j, k);
			hi[0][q] = 0.0 + PB(0, q, i + 1, j, k);
			lo[1][q] = 0.0 + PB(1, q, i, j, k);
			hi[1][q] = 0.0 + PB(1, q, i, j + 1, k);
			lo[2][q] = 0.0 + PB(2, q, i, j, k);
			hi[2][q] = 0.0 + PB(2, q, i, j, k + 1);
		}
		double rho = s[c];
		double m[3] = {s[n + c], s[2 * n + c], s[3 * n + c]};
		double E = s[4 * n + c];
		double vb[3] = {s[5 * n + c], s[6 * n + c], s[7 * n + c]};
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
		out[c] = rho;
		for (int q = 0; q < 3; q++) out[(1 + q) * n + c] = m[q];
		out[4 * n + c] = E;
		for (int q = 0; q < 3; q++) out[(BQ_VB + q) * n + c] = vb[q];
	}
#undef PB
#undef EZX
#undef EYX
#undef EXY
#undef EZY
#undef EXZ
#undef EYZ
	free(P);
	/* update_vol_B(substep), sync_magnetic_field, update_vol_B(0):
	   source/mhd/solve.hpp:1904-1916 */
	for (long c = 0; c < n; c++) if (b->ctype[c] >= 0) box_refit(b, out, n, c);
	or_box_sync(b, out);
	for (long c = 0; c < n; c++) if (b->ctype[c] >= 0) box_refit(b, out, n, c);
	memcpy(s, out, sizeof(double) * 14 * n);
	free(out);
	return OR_OK;
}

int or_box_run(const or_box* b, int solver, double cfl, double t_start, double t_end,
	long max_steps, double* s, long* steps_taken, double* t_final)
{
	double sm[3], dt;
	long zc[3], bad;
	/* final init with a zero time step, tests/mhd/solar_wind_box.cpp:464-473 */
	int e = or_box_step(b, solver, 0.0, s, sm, zc, &bad);
	double t = t_start;
	long steps = 0;
	while (!e && t < t_end && steps < max_steps) {
		/* lagged CFL step, as or_cfl_dt */
		double cdt = DBL_MAX;
		for (int d = 0; d < 3; d++) {
			if (zc[d] > 0) {
				cdt = 0;
				break;
			}
			cdt = smin(cdt, cfl * b->len[d] / sm[d]);
		}
		const double until_end = t_end - t;
		dt = (until_end <= cdt) ? until_end : cdt;
		e = or_box_step(b, solver, dt, s, sm, zc, &bad);
		if (e) break;
		t += dt;
		steps++;
		e = or_box_boundaries(b, s);
		if (e) break;
		or_box_sync(b, s);
	}
	if (steps_taken) *steps_taken = steps;
	if (t_final) *t_final = t;
	return e;
}

/* ======================================================================
   Uniform grids with value and copy boundaries (declarations and the
   layout in pamhd_oracle.h)
   ====================================================================== */

typedef struct { long id; int fn; int en[3]; } bnb_t;

#define BS(q, c) S[(long)(q) * n + (c)]
#define BF(f, c) S[(long)(8 + (f)) * n + (c)]
#define FI(f, c) finfo[(long)(f) * n + (c)]

static inline long bgrid_n(const or_bgrid* g) { return (long)g->nx * g->ny * g->nz; }

/* face index (-x, +x, -y, +y, -z, +z) of a direction -3..-1, +1..+3 */
static inline int fidx(int dir) { return 2 * ((dir < 0 ? -dir : dir) - 1) + (dir > 0); }

/* neighbours of cell (i, j, k): offsets -1..1 with z slowest and x fastest,
   periodic dimensions wrapped, closed ones cut; fn = +-(d + 1) for a face
   neighbour, en = (edge dimension, offset sign in the lower other
   dimension, in the higher one) for an edge neighbour (en[0] = -1 else) */
static int bnb_list(const or_bgrid* g, long i, long j, long k, bnb_t* out)
{
	const long N[3] = {g->nx, g->ny, g->nz}, idx[3] = {i, j, k};
	int m = 0;
	for (int oz = -1; oz <= 1; oz++)
	for (int oy = -1; oy <= 1; oy++)
	for (int ox = -1; ox <= 1; ox++) {
		const int o[3] = {ox, oy, oz};
		const int nzc = (ox != 0) + (oy != 0) + (oz != 0);
		if (nzc == 0) continue;
		long p[3];
		int ok = 1;
		for (int d = 0; d < 3 && ok; d++) {
			long v = idx[d] + o[d];
			if (v < 0 || v >= N[d]) {
				if (!g->periodic[d]) ok = 0;
				else v = (v + N[d]) % N[d];
			}
			p[d] = v;
		}
		if (!ok) continue;
		bnb_t* b = out + m++;
		b->id = p[0] + N[0] * (p[1] + N[1] * p[2]);
		b->fn = 0;
		b->en[0] = -1;
		b->en[1] = b->en[2] = 0;
		if (nzc == 1) {
			for (int d = 0; d < 3; d++) if (o[d]) b->fn = o[d] * (d + 1);
		} else if (nzc == 2) {
			const int e = !ox ? 0 : (!oy ? 1 : 2);
			const int d1 = e == 0 ? 1 : 0, d2 = e == 2 ? 1 : 2;
			b->en[0] = e;
			b->en[1] = o[d1];
			b->en[2] = o[d2];
		}
	}
	return m;
}

static inline void bcell_ijk(const or_bgrid* g, long c, long* i, long* j, long* k)
{
	*i = c % g->nx;
	*j = (c / g->nx) % g->ny;
	*k = c / ((long)g->nx * g->ny);
}

int or_bnd_classify(const or_bgrid* g, const unsigned char* in_value,
	const unsigned char* in_copy, signed char* vtype, signed char* ctype, signed char* finfo)
{
	const long n = bgrid_n(g);
	bnb_t nb[26];
	long i, j, k;
	for (int v = 0; v < 4; v++) {
		signed char* t = vtype + (long)v * n;
		const unsigned char* iv = in_value + (long)v * n;
		const unsigned char* ic = in_copy + (long)v * n;
		for (long c = 0; c < n; c++) t[c] = iv[c] ? 3 : (ic[c] ? 4 : 1);
		/* boundary cells without a normal neighbour (the cell itself aside)
		   are not solved */
		for (long c = 0; c < n; c++) {
			if (t[c] == 1) continue;
			bcell_ijk(g, c, &i, &j, &k);
			const int m = bnb_list(g, i, j, k, nb);
			int have = 0;
			for (int q = 0; q < m && !have; q++) {
				if (nb[q].id == c) continue;
				if (t[nb[q].id] == 1) have = 1;
			}
			if (!have) t[c] = 2;
		}
		/* copy cells need a normal neighbour to copy from */
		for (long c = 0; c < n; c++) {
			if (t[c] != 4) continue;
			bcell_ijk(g, c, &i, &j, &k);
			const int m = bnb_list(g, i, j, k, nb);
			int sources = 0;
			for (int q = 0; q < m; q++) if (t[nb[q].id] == 1) sources++;
			if (!sources) t[c] = 2;
		}
	}
	/* set_solver_info */
	for (long c = 0; c < n; c++) {
		ctype[c] = 1;
		for (int v = 0; v < 4; v++) {
			const int t = vtype[(long)v * n + c];
			if (t == 3 || t == 4) ctype[c] = 0;
		}
		const int ds = vtype[c] == 2;
		for (int v = 1; v < 4; v++) {
			if ((vtype[(long)v * n + c] == 2) != ds) return OR_ERR_SOLVER;
		}
		if (ds) ctype[c] = -1;
	}
	/* classify_faces: faces sharing an edge with a normal cell are normal */
	for (long c = 0; c < n; c++) {
		for (int f = 0; f < 6; f++) FI(f, c) = -99;
		if (ctype[c] == 1) {
			for (int f = 0; f < 6; f++) FI(f, c) = 1;
			continue;
		}
		bcell_ijk(g, c, &i, &j, &k);
		const int m = bnb_list(g, i, j, k, nb);
		for (int q = 0; q < m; q++) {
			if (ctype[nb[q].id] < 1) continue;
			const int fn = nb[q].fn;
			if (fn != 0) {
				for (int dir = -3; dir <= 3; dir++) {
					if (dir == 0 || dir == -fn) continue;
					FI(fidx(dir), c) = 1;
				}
			}
			if (nb[q].en[0] >= 0) {
				const int e = nb[q].en[0];
				const int dir1 = 1 + (e == 0 ? 1 : 0), dir2 = 1 + (e == 2 ? 1 : 2);
				FI(fidx(nb[q].en[1] < 0 ? -dir1 : +dir1), c) = 1;
				FI(fidx(nb[q].en[2] < 0 ? -dir2 : +dir2), c) = 1;
			}
		}
	}
	/* the other faces of a cell with a normal face are boundary faces */
	for (long c = 0; c < n; c++) {
		int normal = 0;
		for (int f = 0; f < 6; f++) if (FI(f, c) == 1) normal = 1;
		if (!normal) continue;
		for (int f = 0; f < 6; f++) if (FI(f, c) < 0) FI(f, c) = 0;
	}
	/* faces next to normal faces of neighbours are boundary faces */
	for (long c = 0; c < n; c++) {
		bcell_ijk(g, c, &i, &j, &k);
		const int m = bnb_list(g, i, j, k, nb);
		for (int q = 0; q < m; q++) {
			const long b = nb[q].id;
			const int fn = nb[q].fn;
			if (fn != 0) {
				for (int dir = -3; dir <= 3; dir++) {
					if (dir == 0) continue;
					if (FI(fidx(dir), b) < 1) continue;
					if (dir == fn && FI(fidx(dir), c) < 0) {
						FI(fidx(dir), c) = 0;
						continue;
					}
					if (dir == -fn && FI(fidx(fn), c) < 1) return OR_ERR_SOLVER;
					if (FI(fidx(dir), b) == 1 && FI(fidx(dir), c) < 0) FI(fidx(dir), c) = 0;
				}
			}
			if (nb[q].en[0] >= 0) {
				const int e = nb[q].en[0];
				const int dir1 = 1 + (e == 0 ? 1 : 0), dir2 = 1 + (e == 2 ? 1 : 2);
				const int s1 = nb[q].en[1] < 0 ? -1 : 1, s2 = nb[q].en[2] < 0 ? -1 : 1;
				if (FI(fidx(-s1 * dir1), b) == 1 || FI(fidx(-s2 * dir2), b) == 1) {
					if (FI(fidx(s1 * dir1), c) < 0) FI(fidx(s1 * dir1), c) = 0;
					if (FI(fidx(s2 * dir2), c) < 0) FI(fidx(s2 * dir2), c) = 0;
				}
			}
		}
	}
	for (long c = 0; c < n; c++)
		for (int f = 0; f < 6; f++) if (FI(f, c) < -1) FI(f, c) = -1;
	return OR_OK;
}

static inline void bvel(const double* S, long n, long c, double v[3])
{
	/* get_velocity, source/mhd/common.hpp:121-130 */
	if (BS(0, c) > 0) {
		const double inv = 1 / BS(0, c);
		for (int q = 0; q < 3; q++) v[q] = BS(1 + q, c) * inv;
	} else {
		v[0] = v[1] = v[2] = 0;
	}
}

static inline int bpressure(const or_bgrid* g, const double* S, long n, long c, double* p)
{
	const double U[8] = {BS(0, c), BS(1, c), BS(2, c), BS(3, c), BS(4, c), BS(5, c), BS(6, c), BS(7, c)};
	return or_get_pressure(U, g->gamma, g->mu0, p);
}

/* update_vol_B with constant thermal pressure for one cell (cells with
   rho <= 0 are skipped) */
static void bnd_refit(const or_bgrid* g, double* S, long n, long c)
{
	if (BS(0, c) <= 0) return;
	double p, v[3], B[3];
	bpressure(g, S, n, c, &p);
	for (int d = 0; d < 3; d++) {
		BS(5 + d, c) = 0.5 * (BF(2 * d, c) + BF(2 * d + 1, c));
		B[d] = BS(5 + d, c);
	}
	const double inv = 1 / BS(0, c);
	for (int q = 0; q < 3; q++) v[q] = BS(1 + q, c) * inv;
	BS(4, c) = or_total_energy(BS(0, c), v, p, B, g->gamma, g->mu0);
}

int or_bnd_apply(const or_bgrid* g, const unsigned char* val_mask, const double* n_val,
	const double* v_val, const double* p_val, const double* b_val, double* S)
{
	const long n = bgrid_n(g);
	const signed char* ctype = g->ctype;
	const signed char* finfo = g->finfo;
	bnb_t nb[26];
	long i, j, k;
	/* number density */
	for (long c = 0; c < n; c++) if (val_mask[c]) BS(0, c) = g->proton_mass * n_val[c];
	for (long c = 0; c < n; c++) {
		if (g->vtype[c] != 4) continue;
		bcell_ijk(g, c, &i, &j, &k);
		const int m = bnb_list(g, i, j, k, nb);
		double fs = 0, es = 0;
		size_t nf = 0, ne = 0;
		for (int q = 0; q < m; q++) {
			const long b = nb[q].id;
			if (ctype[b] < 1) continue;
			if (nb[q].fn != 0) { fs += BS(0, b); nf++; }
			if (nb[q].en[0] >= 0) { es += BS(0, b); ne++; }
		}
		if (nf > 0) BS(0, c) = fs / nf;
		else if (ne > 0) BS(0, c) = es / ne;
	}
	/* velocity */
	for (long c = 0; c < n; c++) {
		if (!val_mask[n + c]) continue;
		for (int q = 0; q < 3; q++) BS(1 + q, c) = v_val[(long)q * n + c] * BS(0, c);
	}
	for (long c = 0; c < n; c++) {
		if (g->vtype[n + c] != 4) continue;
		bcell_ijk(g, c, &i, &j, &k);
		const int m = bnb_list(g, i, j, k, nb);
		double fs[3] = {0, 0, 0}, es[3] = {0, 0, 0}, v[3];
		size_t nf = 0, ne = 0;
		for (int q = 0; q < m; q++) {
			const long b = nb[q].id;
			if (ctype[b] < 1) continue;
			if (nb[q].fn != 0) {
				bvel(S, n, b, v);
				for (int r = 0; r < 3; r++) fs[r] = fs[r] + v[r];
				nf++;
			}
			if (nb[q].en[0] >= 0) {
				bvel(S, n, b, v);
				for (int r = 0; r < 3; r++) es[r] = es[r] + v[r];
				ne++;
			}
		}
		/* mul(mul(Mas, sum), 1 / count) with an integer quotient: the
		   factor is 1 for one source and 0 for more (the reference's
		   integer division) */
		if (nf > 0) {
			const double f = (double)(1 / nf);
			for (int r = 0; r < 3; r++) BS(1 + r, c) = fs[r] * BS(0, c) * f;
		} else if (ne > 0) {
			const double f = (double)(1 / ne);
			for (int r = 0; r < 3; r++) BS(1 + r, c) = es[r] * BS(0, c) * f;
		}
	}
	/* pressure */
	for (long c = 0; c < n; c++) {
		if (!val_mask[2 * n + c]) continue;
		if (BS(0, c) <= 0) return OR_ERR_DENSITY;
		double v[3];
		bvel(S, n, c, v);
		const double B[3] = {BS(5, c), BS(6, c), BS(7, c)};
		BS(4, c) = or_total_energy(BS(0, c), v, p_val[c], B, g->gamma, g->mu0);
	}
	for (long c = 0; c < n; c++) {
		if (g->vtype[2 * n + c] != 4) continue;
		bcell_ijk(g, c, &i, &j, &k);
		const int m = bnb_list(g, i, j, k, nb);
		double fs = 0, es = 0, p;
		size_t nf = 0, ne = 0;
		for (int q = 0; q < m; q++) {
			const long b = nb[q].id;
			if (ctype[b] < 1) continue;
			if (nb[q].fn != 0) {
				if (bpressure(g, S, n, b, &p)) return OR_ERR_DENSITY;
				fs += p;
				nf++;
			}
			if (nb[q].en[0] >= 0) {
				if (bpressure(g, S, n, b, &p)) return OR_ERR_DENSITY;
				es += p;
				ne++;
			}
		}
		if (nf == 0 && ne == 0) continue;
		if (BS(0, c) <= 0) return OR_ERR_DENSITY;
		double v[3];
		bvel(S, n, c, v);
		const double B[3] = {BS(5, c), BS(6, c), BS(7, c)};
		BS(4, c) = or_total_energy(BS(0, c), v, nf > 0 ? fs / nf : es / ne, B, g->gamma, g->mu0);
	}
	/* face field: value boundaries on the boundary faces */
	for (long c = 0; c < n; c++) {
		if (!val_mask[3 * n + c]) continue;
		for (int f = 0; f < 6; f++) if (FI(f, c) == 0) BF(f, c) = b_val[(long)f * n + c];
	}
	/* face field: copy boundaries (uniform grids: neighbours of equal size) */
	for (long c = 0; c < n; c++) {
		if (g->vtype[3 * n + c] != 4) continue;
		int cnt[6] = {0, 0, 0, 0, 0, 0};
		for (int d = 0; d < 3; d++) {
			for (int s = 0; s < 2; s++) if (FI(2 * d + s, c) < 1) BF(2 * d + s, c) = 0;
			const int lo = FI(2 * d, c), hi = FI(2 * d + 1, c);
			if ((lo == 0 && hi == 0) || (lo == -1 && hi == -1)) return OR_ERR_SOLVER;
			if (lo == 0 && hi == 1) { BF(2 * d, c) += BF(2 * d + 1, c); cnt[2 * d]++; }
			if (lo == 1 && hi == 0) { BF(2 * d + 1, c) += BF(2 * d, c); cnt[2 * d + 1]++; }
		}
		bcell_ijk(g, c, &i, &j, &k);
		const int m = bnb_list(g, i, j, k, nb);
		for (int q = 0; q < m; q++) {
			const long b = nb[q].id;
			const int fn = nb[q].fn;
			if (fn == 0 && nb[q].en[0] < 0) continue;
			if (fn != 0) {
				for (int dir = -3; dir <= 3; dir++) {
					if (dir == 0) continue;
					const int f = fidx(dir);
					if (FI(f, c) == 0 && FI(f, b) == 1) { BF(f, c) += BF(f, b); cnt[f]++; }
				}
			}
			if (nb[q].en[0] >= 0) {
				const int e = nb[q].en[0];
				const int d1 = e == 0 ? 1 : 0, d2 = e == 2 ? 1 : 2;
				const int dirs[2] = {nb[q].en[1] * (d1 + 1), nb[q].en[2] * (d2 + 1)};
				for (int r = 0; r < 2; r++) {
					const int f = fidx(dirs[r]), fo = fidx(-dirs[r]);
					if (FI(f, c) == 0 && FI(fo, b) == 1) { BF(f, c) += BF(fo, b); cnt[f]++; }
				}
			}
		}
		for (int f = 0; f < 6; f++) {
			if (FI(f, c) != 0) continue;
			if (cnt[f] == 0) return OR_ERR_SOLVER;
			BF(f, c) /= cnt[f];
		}
	}
	/* update_vol_B(0) at constant pressure */
	for (long c = 0; c < n; c++) if (ctype[c] >= 0) bnd_refit(g, S, n, c);
	return OR_OK;
}

void or_bnd_sync(const or_bgrid* g, double* S)
{
	const long n = bgrid_n(g);
	const signed char* ctype = g->ctype;
	const long N[3] = {g->nx, g->ny, g->nz};
	for (int d = 0; d < 3; d++)
	for (long c = 0; c < n; c++) {
		long idx[3];
		bcell_ijk(g, c, idx, idx + 1, idx + 2);
		if (ctype[c] < 0) continue;
		long up = idx[d] + 1;
		if (up >= N[d]) {
			if (!g->periodic[d]) continue;
			up = 0;
		}
		idx[d] = up;
		const long b = idx[0] + N[0] * (idx[1] + N[1] * idx[2]);
		if (ctype[b] < 0) continue;
		double* pc = &BF(2 * d + 1, c);   /* c's + copy */
		double* pb = &BF(2 * d, b);       /* b's - copy */
		if (c <= b) {
			/* the visit from c comes first (a one-cell dimension: the
			   cell's -d visit, copies -d then +d) */
			if (c == b) {
				double diff = *pb - *pc; *pb -= diff / 2; *pc += diff / 2;
				diff = *pc - *pb; *pc -= diff / 2; *pb += diff / 2;
			} else {
				double diff = *pc - *pb; *pc -= diff / 2; *pb += diff / 2;
				diff = *pb - *pc; *pb -= diff / 2; *pc += diff / 2;
			}
		} else {
			double diff = *pb - *pc; *pb -= diff / 2; *pc += diff / 2;
			diff = *pc - *pb; *pc -= diff / 2; *pb += diff / 2;
		}
	}
}

/* the cell across the face on c's side dir (-3..+3), or -1 */
static long bnd_face_nb(const or_bgrid* g, long c, int dir)
{
	long idx[3];
	bcell_ijk(g, c, idx, idx + 1, idx + 2);
	const long N[3] = {g->nx, g->ny, g->nz};
	const int d = (dir < 0 ? -dir : dir) - 1;
	long v = idx[d] + (dir < 0 ? -1 : 1);
	if (v < 0 || v >= N[d]) {
		if (!g->periodic[d]) return -1;
		v = (v + N[d]) % N[d];
	}
	idx[d] = v;
	return idx[0] + N[0] * (idx[1] + N[1] * idx[2]);
}

/* each cell's period (2^level) and the smallest period around it, the
   substep length and the largest period: or_substep_periods over the cells
   of type >= 0 and the faces between them */
static int bnd_periods(const or_bgrid* g, double cfl, double max_dt, const int* substep_min,
	const int* substep_max, const double* maxv, int* period, int* minsub, double* sub_dt_out,
	int* max_period_out, long* bad_cell)
{
	const long n = bgrid_n(g);
	const signed char* ctype = g->ctype;
	const double len[3] = {g->dx, g->dy, g->dz};
	double* ts = (double*)malloc(sizeof(double) * n);
	int* lomin = (int*)malloc(sizeof(int) * n);
	int* smaxv = (int*)malloc(sizeof(int) * n);
	bnb_t nb[26];
	long i, j, k;
	int err = OR_OK;
	if (!ts || !lomin || !smaxv) { err = OR_ERR_SOLVER; goto done; }
	int lo = INT_MAX;
	for (long c = 0; c < n; c++) {
		lomin[c] = substep_min[c];
		smaxv[c] = substep_max[c];
		if (lomin[c] < lo) lo = lomin[c];
	}
	if (lo > 0) for (long c = 0; c < n; c++) lomin[c] -= lo;
	/* minimize_timestep over the faces between cells of type >= 0 */
	for (long c = 0; c < n; c++) {
		double t = DBL_MAX;
		if (ctype[c] >= 0) {
			for (int dir = -3; dir <= 3; dir++) {
				if (dir == 0) continue;
				const long b = bnd_face_nb(g, c, dir);
				if (b < 0 || ctype[b] < 0) continue;
				const int d = (dir < 0 ? -dir : dir) - 1;
				const double v = maxv[(long)d * n + (dir < 0 ? c : b)];
				if (v <= 0) t = 0;
				else t = smin(cfl * len[d] / v, t);
			}
		}
		ts[c] = t;
	}
	double smallest = DBL_MAX;
	for (long c = 0; c < n; c++) if (ctype[c] >= 0) smallest = smin(ts[c], smallest);
	double sub_dt;
	if (max_dt <= smallest) {
		for (long c = 0; c < n; c++) lomin[c] = smaxv[c] = 0;
		sub_dt = max_dt;
	} else {
		double ret = DBL_MAX;
		for (long c = 0; c < n; c++) {
			if (ctype[c] < 0) continue;
			ret = smin(ret, ts[c] / (double)(1 << lomin[c]));
		}
		for (long c = 0; c < n; c++) {
			if (ctype[c] < 0) continue;
			const double dt = ts[c];
			if (dt > 0) {
				const double x = smin(dt, max_dt) / ret;
				int step = (isfinite(x) && x > 0) ? ilogb(x) : 0;
				if (step < 0) step = 0;
				if (smaxv[c] > step) smaxv[c] = step;
				if (lomin[c] > smaxv[c] || smaxv[c] > 30) {
					if (bad_cell) *bad_cell = c;
					err = OR_ERR_SUBSTEP;
					goto done;
				}
			} else {
				lomin[c] = smaxv[c] = 0;
			}
		}
		sub_dt = ret;
	}
	/* restrict_substepping_period over face and edge neighbours of type >= 0 */
	for (long c = 0; c < n; c++) period[c] = ctype[c] >= 0 ? smaxv[c] : 0;
	for (;;) {
		long modified = 0;
		for (long c = 0; c < n; c++) {
			if (ctype[c] < 0) continue;
			bcell_ijk(g, c, &i, &j, &k);
			const int m = bnb_list(g, i, j, k, nb);
			for (int q = 0; q < m; q++) {
				const long b = nb[q].id;
				if (ctype[b] < 0) continue;
				if (nb[q].fn == 0 && nb[q].en[0] < 0) continue;
				if (period[c] > period[b] + 1) { period[c] = period[b] + 1; modified++; }
			}
		}
		if (!modified) break;
	}
	int maxp = -1;
	for (long c = 0; c < n; c++) {
		period[c] = 1 << period[c];
		if (ctype[c] >= 0 && period[c] > maxp) maxp = period[c];
	}
	for (long c = 0; c < n; c++) {
		int m_ = period[c];
		if (ctype[c] >= 0) {
			bcell_ijk(g, c, &i, &j, &k);
			const int m = bnb_list(g, i, j, k, nb);
			for (int q = 0; q < m; q++) {
				const long b = nb[q].id;
				if (ctype[b] < 0) continue;
				if (nb[q].fn == 0 && nb[q].en[0] < 0) continue;
				if (period[b] < m_) m_ = period[b];
			}
		}
		minsub[c] = m_;
	}
	*sub_dt_out = sub_dt;
	*max_period_out = maxp;
done:
	free(ts); free(lomin); free(smaxv);
	return err;
}

/* P of the d-face on the negative side of cell (i, j, k): periodic
   dimensions wrap, 0 beyond a closed end */
static inline double bnd_pf(const or_bgrid* g, const double* P, long n, int d, int q,
	long i, long j, long k)
{
	long idx[3] = {i, j, k};
	const long N[3] = {g->nx, g->ny, g->nz};
	for (int a = 0; a < 3; a++) {
		if (idx[a] < 0 || idx[a] >= N[a]) {
			if (!g->periodic[a]) return 0.0;
			idx[a] = (idx[a] + N[a]) % N[a];
		}
	}
	return P[((long)d * 8 + q) * n + idx[0] + N[0] * (idx[1] + N[1] * idx[2])];
}

int or_bnd_step(const or_bgrid* g, int solver, double time_step_factor, double max_dt,
	const int* substep_min, const int* substep_max, double* S, double* maxv,
	double* dt_out, int* max_period_out, long* bad_face)
{
	const long n = bgrid_n(g);
	const signed char* ctype = g->ctype;
	const double len[3] = {g->dx, g->dy, g->dz};
	const double area[3] = {g->dy * g->dz, g->dx * g->dz, g->dx * g->dy};
	const double dx = g->dx, dy = g->dy, dz = g->dz;
	int* period = (int*)malloc(sizeof(int) * n);
	int* minsub = (int*)malloc(sizeof(int) * n);
	double* slot = (double*)calloc((size_t)48 * n, sizeof(double));
	double* fdb = (double*)calloc((size_t)6 * n, sizeof(double));
	double* P = (double*)malloc(sizeof(double) * 24 * n);
	int err = OR_OK;
	if (!period || !minsub || !slot || !fdb || !P) { err = OR_ERR_SOLVER; goto done; }
	double sub_dt;
	int maxp;
	long bad_cell = -1;
	err = bnd_periods(g, time_step_factor, max_dt, substep_min, substep_max, maxv, period,
		minsub, &sub_dt, &maxp, &bad_cell);
	if (err) { if (bad_face) *bad_face = bad_cell; goto done; }

	double total_dt = 0;
	or_fc = 0;
	for (int sub = 1; sub <= maxp; sub++) {
		total_dt += sub_dt;
		/* fluxes of the faces due at this substep between cells of type
		   >= 0, owned by the cell on the negative side; P = 0 elsewhere */
		long bad = -1;
		for (int d = 0; d < 3 && !err; d++)
		for (long c = 0; c < n; c++) {
			for (int q = 0; q < 8; q++) P[(d * 8 + q) * n + c] = 0;
			const long o = bnd_face_nb(g, c, -(d + 1));
			if (o < 0 || ctype[o] < 0 || ctype[c] < 0) continue;
			const int m = minsub[o];
			if (sub % m != 0) continue;
			or_fc++;
			double L[8], R[8], bgr[3] = {0, 0, 0}, F[8], sp = 0;
			{
				const double ml[3] = {BS(1, o), BS(2, o), BS(3, o)}, Bl[3] = {BS(5, o), BS(6, o), BS(7, o)};
				const double mr[3] = {BS(1, c), BS(2, c), BS(3, c)}, Br[3] = {BS(5, c), BS(6, c), BS(7, c)};
				L[RHO] = BS(0, o); L[NRJ] = BS(4, o);
				R[RHO] = BS(0, c); R[NRJ] = BS(4, c);
				rot_to_face(d, ml, L + M0); rot_to_face(d, Bl, L + B0);
				rot_to_face(d, mr, R + M0); rot_to_face(d, Br, R + B0);
			}
			if (g->bg) {
				const double b[3] = {g->bg[(3 * d + 0) * n + o], g->bg[(3 * d + 1) * n + o],
					g->bg[(3 * d + 2) * n + o]};
				rot_to_face(d, b, bgr);
			}
			const int e = or_flux(solver, L, R, bgr, g->gamma, g->mu0, F, &sp);
			if (e) {
				if (!err || (long)d * n + c < bad) { err = e; bad = (long)d * n + c; }
				continue;
			}
			maxv[(long)d * n + c] = sp;
			double Fg[8];
			Fg[RHO] = F[RHO];
			Fg[NRJ] = F[NRJ];
			rot_to_global(d, F + M0, Fg + M0);
			rot_to_global(d, F + B0, Fg + B0);
			const double min_dt = sub_dt * m;
			for (int q = 0; q < 8; q++) P[(d * 8 + q) * n + c] = min_dt * Fg[q];
		}
		if (err) { if (bad_face) *bad_face = bad; goto done; }

		/* flux of the d-face on the negative side of the cell at offset
		   (a, b, e) from (i, j, k); 0 where that cell does not exist */
#define BPF(d, q, a, b, e) bnd_pf(g, P, n, (d), (q), i + (a), j + (b), k + (e))
#define EZX(a, b, e) (+(BPF(0, 6, a, b, e) * dz / 4))
#define EYX(a, b, e) (-(BPF(0, 7, a, b, e) * dy / 4))
#define EXY(a, b, e) (+(BPF(1, 7, a, b, e) * dx / 4))
#define EZY(a, b, e) (-(BPF(1, 5, a, b, e) * dz / 4))
#define EXZ(a, b, e) (-(BPF(2, 6, a, b, e) * dx / 4))
#define EYZ(a, b, e) (+(BPF(2, 5, a, b, e) * dy / 4))
		for (long c = 0; c < n; c++) {
			long i, j, k;
			bcell_ijk(g, c, &i, &j, &k);
			double db[6];
			for (int side = 0; side < 2; side++) {
				const long a = side;
				double r = 0;
				r += EZX(a, 1, 0); r -= EZX(a, -1, 0); r -= EYX(a, 0, 1); r += EYX(a, 0, -1);
				r += EZY(a, 1, 0); r += EZY(a - 1, 1, 0); r -= EZY(a, 0, 0); r -= EZY(a - 1, 0, 0);
				r -= EYZ(a, 0, 1); r -= EYZ(a - 1, 0, 1); r += EYZ(a, 0, 0); r += EYZ(a - 1, 0, 0);
				db[0 + side] = r;
				const long b = side;
				r = 0;
				r += EXY(0, b, 1); r -= EXY(0, b, -1); r -= EZY(1, b, 0); r += EZY(-1, b, 0);
				r += EXZ(0, b, 1); r += EXZ(0, b - 1, 1); r -= EXZ(0, b, 0); r -= EXZ(0, b - 1, 0);
				r -= EZX(1, b, 0); r += EZX(0, b, 0); r -= EZX(1, b - 1, 0); r += EZX(0, b - 1, 0);
				db[2 + side] = r;
				const long z = side;
				r = 0;
				r += EYZ(1, 0, z); r -= EYZ(-1, 0, z); r -= EXZ(0, 1, z); r += EXZ(0, -1, z);
				r += EYX(1, 0, z); r += EYX(1, 0, z - 1); r -= EYX(0, 0, z); r -= EYX(0, 0, z - 1);
				r -= EXY(0, 1, z); r -= EXY(0, 1, z - 1); r += EXY(0, 0, z); r += EXY(0, 0, z - 1);
				db[4 + side] = r;
			}
			for (int f = 0; f < 6; f++) fdb[(long)f * n + c] += db[f];
			for (int d = 0; d < 3; d++) {
				const long cp = bnd_face_nb(g, c, d + 1);
				for (int q = 0; q < 8; q++) {
					slot[((long)(2 * d) * 8 + q) * n + c] += P[(d * 8 + q) * n + c];
					if (cp >= 0) slot[((long)(2 * d + 1) * 8 + q) * n + c] += P[(d * 8 + q) * n + cp];
				}
			}
			const int due = sub % period[c] == 0;
			if (ctype[c] > 0 && !due) continue;   /* normal cells accumulate until due */
			if (ctype[c] > 0) {
				/* update_mhd_state */
				double rho = BS(0, c), m[3] = {BS(1, c), BS(2, c), BS(3, c)}, E = BS(4, c);
				double vb[3] = {BS(5, c), BS(6, c), BS(7, c)};
#define SL(dir, q) slot[((long)(dir) * 8 + (q)) * n + c]
				for (int d = 0; d < 3; d++) rho += (SL(2 * d, RHO) - SL(2 * d + 1, RHO)) / len[d];
				for (int d = 0; d < 3; d++) {
					const double il = 1 / len[d];
					for (int q = 0; q < 3; q++) m[q] = m[q] + (SL(2 * d, M0 + q) - SL(2 * d + 1, M0 + q)) * il;
				}
				for (int d = 0; d < 3; d++) E += (SL(2 * d, NRJ) - SL(2 * d + 1, NRJ)) / len[d];
				for (int d = 0; d < 3; d++) {
					const double il = 1 / len[d];
					for (int q = 0; q < 3; q++) vb[q] = vb[q] + (SL(2 * d, B0 + q) - SL(2 * d + 1, B0 + q)) * il;
				}
#undef SL
				for (int d = 0; d < 3; d++) {
					BF(2 * d, c) += fdb[(long)(2 * d) * n + c] / area[d];
					BF(2 * d + 1, c) += fdb[(long)(2 * d + 1) * n + c] / area[d];
				}
				BS(0, c) = rho;
				for (int q = 0; q < 3; q++) BS(1 + q, c) = m[q];
				BS(4, c) = E;
				for (int q = 0; q < 3; q++) BS(5 + q, c) = vb[q];
			}
			for (int dir = 0; dir < 6; dir++) {
				for (int q = 0; q < 8; q++) slot[((long)dir * 8 + q) * n + c] = 0;
				fdb[(long)dir * n + c] = 0;
			}
			/* update_vol_B(substep) of cells of type >= 0 due now */
			if (ctype[c] >= 0 && due) bnd_refit(g, S, n, c);
		}
#undef BPF
#undef EZX
#undef EYX
#undef EXY
#undef EZY
#undef EXZ
#undef EYZ
	}
	or_bnd_sync(g, S);
	for (long c = 0; c < n; c++) if (ctype[c] >= 0) bnd_refit(g, S, n, c);
	*dt_out = total_dt;
	*max_period_out = maxp;
done:
	free(period); free(minsub); free(slot); free(fdb); free(P);
	return err;
}

#undef BS
#undef BF
#undef FI
