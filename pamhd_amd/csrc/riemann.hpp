/*
Device Riemann solvers of the PAMHD ideal-MHD step for CDNA4 (gfx950).

Each solver takes the two face-normal-frame states of a face (negative side
L, positive side R) as FP64 conserved variables and returns the flux per unit
area and time plus the non-negative maximum signal speed, like the
reference's plug-in signature

    get_flux_X<Mas, Mom, Nrj, Mag>(MHD& L, MHD& R, const Vector& bg, gamma, mu0)
        -> std::tuple<MHD, double>           (source/mhd/rusanov.hpp:59-73 etc.)

Flux precision T is a template parameter:
  * T = double restates the reference's arithmetic in its operation order,
    built without FMA contraction, so results match the FP64 CPU path bit
    for bit;
  * T = float evaluates the wave speeds and fluxes in FP32, while the thermal
    pressure E - m^2/(2 rho) - B^2/(2 mu0) (cancellation-prone) and the
    positivity checks stay FP64.  Divisions and square roots use the
    hardware reciprocal / square-root instructions (about 1 ulp).

A domain error (the reference throws std::domain_error) is reported through
the returned status code instead (PAMHD_E_* of include/pamhd_mi355x.h).
*/
#pragma once

#include <hip/hip_runtime.h>
#include <cfloat>
#include <cmath>
#include <type_traits>

namespace pamhd_dev {

enum : int {
	ERR_OK = 0,
	ERR_DENSITY = 1,
	ERR_PRESSURE = 2,
	ERR_SPEED = 3,
	ERR_FLUX = 4,
	ERR_FIELD = 5,
	ERR_ROE_DENSITY = 6,
	ERR_ROE_PRESSURE = 7,
	ERR_SOLVER = 8
};

/* solver numbering of source/mhd/common.hpp:54-61 */
enum : int {
	RUSANOV = 0,
	RUSANOV_STAGGERED = 1,
	HLL_ATHENA = 2,
	HLLD_ATHENA = 3,
	ROE_ATHENA = 4,
	HYBRID = 5
};

enum { RHO = 0, M0 = 1, M1 = 2, M2 = 3, NRJ = 4, B0 = 5, B1 = 6, B2 = 7 };

/* std::max / std::min: the first argument wins ties */
template <class T> __device__ __forceinline__ T smax(T a, T b) { return (a < b) ? b : a; }
template <class T> __device__ __forceinline__ T smin(T a, T b) { return (b < a) ? b : a; }

/* precision-dependent primitives */
template <class T> struct Arith;

template <> struct Arith<double> {
	static __device__ __forceinline__ double div(double a, double b) { return a / b; }
	static __device__ __forceinline__ double rcp(double a) { return 1 / a; }
	static __device__ __forceinline__ double sqrt(double a) { return __builtin_sqrt(a); }
	static __device__ __forceinline__ bool isnormal(double a) { return __builtin_isnormal(a); }
	static __device__ __forceinline__ bool isfinite(double a) { return __builtin_isfinite(a); }
	static constexpr double eps = DBL_EPSILON;
};

template <> struct Arith<float> {
	static __device__ __forceinline__ float div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
	static __device__ __forceinline__ float rcp(float a) { return __builtin_amdgcn_rcpf(a); }
	static __device__ __forceinline__ float sqrt(float a) { return __builtin_amdgcn_sqrtf(a); }
	static __device__ __forceinline__ bool isnormal(float a) { return __builtin_isnormal(a); }
	static __device__ __forceinline__ bool isfinite(float a) { return __builtin_isfinite(a); }
	static constexpr float eps = FLT_EPSILON;
};

/*
Physical constants and the per-run values derived from them, computed once on
the host (IEEE division and square root round identically there):
0.5/mu0 and 1/mu0 of get_pressure / get_flux, sqrt(mu0) and 1/sqrt(mu0) of the
HLLD and Roe solvers.
*/
struct Phys {
	double gamma, mu0;
	double gm1;           /* gamma - 1 */
	double inv_mu0;       /* 1.0 / mu0 */
	double half_inv_mu0;  /* 0.5 / mu0 */
	double sqrt_mu0;      /* sqrt(mu0) */
	double inv_sqrt_mu0;  /* 1.0 / sqrt(mu0) */
	double two_mu0;       /* 2 * mu0 */
};

template <class T> __device__ __forceinline__ T dot3(const T* a, const T* b)
{
	return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

/* FP64 thermal pressure, source/mhd/common.hpp:76-115; half_inv_mu0 = 0.5/mu0 */
__host__ __device__ __forceinline__ double pressure64(const double U[8], double gm1, double half_inv_mu0)
{
	const double
		kinetic = 0.5 / U[RHO] * (U[M0] * U[M0] + U[M1] * U[M1] + U[M2] * U[M2]),
		magnetic = half_inv_mu0 * (U[B0] * U[B0] + U[B1] * U[B1] + U[B2] * U[B2]);
	return (U[NRJ] - kinetic - magnetic) * gm1;
}

/* FP64 total energy density from rho, v, p and B,
   source/mhd/common.hpp:254-311 (argument checks made by callers) */
__host__ __device__ __forceinline__ double total_energy64(double rho, const double v[3], double p,
	const double B[3], double gm1, double half_inv_mu0)
{
	const double
		kinetic = 0.5 * rho * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]),
		magnetic = half_inv_mu0 * (B[0] * B[0] + B[1] * B[1] + B[2] * B[2]);
	return p / gm1 + kinetic + magnetic;
}

/*
One side of a face in flux precision: conserved state, FP64-recovered
thermal pressure and velocity m * (1 / rho) (get_velocity,
source/mhd/common.hpp:118-130).
*/
template <class T> struct Side {
	T u[8];
	T p;
	T v[3];
	T ir;        /* 1 / rho (0 when rho <= 0) */
	double p64;
	double pt64; /* FP64 total pressure p + B^2 / 2 mu0 (FP32 sides only) */
	bool rho_pos;
};

template <class T> __device__ __forceinline__ void make_side(const double U[8], const Phys& ph,
	Side<T>& s)
{
	for (int q = 0; q < 8; q++) s.u[q] = (T)U[q];
	s.rho_pos = U[RHO] > 0;
	s.p64 = s.rho_pos ? pressure64(U, ph.gm1, ph.half_inv_mu0) : 0.0;
	s.p = (T)s.p64;
	if constexpr (std::is_same_v<T, float>) {
		s.pt64 = s.p64 + ph.half_inv_mu0 * (U[B0] * U[B0] + U[B1] * U[B1] + U[B2] * U[B2]);
	} else {
		s.pt64 = 0;
	}
	if (s.rho_pos) {
		const T inv = Arith<T>::rcp(s.u[RHO]);
		s.ir = inv;
		s.v[0] = s.u[M0] * inv;
		s.v[1] = s.u[M1] * inv;
		s.v[2] = s.u[M2] * inv;
	} else {
		s.ir = 0;
		s.v[0] = s.v[1] = s.v[2] = 0;
	}
}

/*
FP32 fast magnetosonic speed for the mixed path, from the squared speeds
with the cell's 1/rho: c_s^2 = gamma p / rho, v_A^2 = |B|^2 / (mu0 rho),
v_An^2 = B_n^2 / (mu0 rho), c_f^2 = (c_s^2 + v_A^2 + sqrt((c_s^2 + v_A^2)^2
- 4 c_s^2 v_An^2)) / 2: the same quantity as get_fast_magnetosonic_speed
(source/mhd/common.hpp:415-491; its to_sqrt is this discriminant written
with ratio^2 = v_An^2 / v_A^2) with two square roots and no division.
*/
__device__ __forceinline__ int fast_speed32(const Side<float>& s, const double U[8], const float bg[3],
	float gamma, float inv_mu0, float& cf)
{
	using A = Arith<float>;
	const float t0 = s.u[B0] + bg[0], t1 = s.u[B1] + bg[1], t2 = s.u[B2] + bg[2];
	if (!A::isfinite(t0)) return ERR_FIELD;
	if (!__builtin_isnormal(U[RHO]) || U[RHO] < 0) return ERR_DENSITY;
	if (!__builtin_isnormal(s.p64) || s.p64 < 0) return ERR_PRESSURE;
	const float im = s.ir * inv_mu0;
	const float cs2 = gamma * s.p * s.ir;
	const float va2 = (t0 * t0 + t1 * t1 + t2 * t2) * im;
	if (!(va2 > 0)) {
		cf = A::sqrt(cs2);
		return ERR_OK;
	}
	const float sum = cs2 + va2;
	if (!A::isnormal(sum) || sum < 0) return ERR_SPEED;
	const float disc = sum * sum - 4.0f * cs2 * (t0 * t0 * im);
	if (!A::isnormal(disc) || disc < 0) return ERR_SPEED;
	cf = A::sqrt(0.5f * (sum + A::sqrt(disc)));
	return ERR_OK;
}

/* fast magnetosonic speed with get_sound_speed / get_alfven_speed,
   source/mhd/common.hpp:324-491 */
template <class T> __device__ __forceinline__ int fast_speed(const Side<T>& s,
	const double U[8], const T bg[3], T gamma, T mu0, T& cf)
{
	using A = Arith<T>;
	if constexpr (std::is_same_v<T, float>) return fast_speed32(s, U, bg, gamma, A::rcp(mu0), cf);
	const T t[3] = {s.u[B0] + bg[0], s.u[B1] + bg[1], s.u[B2] + bg[2]};
	if (!A::isfinite(t[0])) return ERR_FIELD;
	const T mag = A::sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
	/* get_sound_speed: density and (FP64) pressure checks */
	if (!__builtin_isnormal(U[RHO]) || U[RHO] < 0) return ERR_DENSITY;
	if (!__builtin_isnormal(s.p64) || s.p64 < 0) return ERR_PRESSURE;
	const T sound = A::sqrt(A::div(gamma * s.p, s.u[RHO]));
	if (mag <= 0) {
		cf = sound;
		return ERR_OK;
	}
	const T alfven = A::div(mag, A::sqrt(mu0 * s.u[RHO]));
	const T
		sound2 = sound * sound,
		alfven2 = alfven * alfven,
		speeds2 = sound2 + alfven2;
	if (!A::isnormal(speeds2) || speeds2 < 0) return ERR_SPEED;
	const T ratio = A::div(t[0], mag);
	const T to_sqrt
		= sound2 * sound2 + alfven2 * alfven2
		+ T(2) * sound2 * alfven2 * (T(1) - T(2) * (ratio * ratio));
	if (!A::isnormal(to_sqrt) || to_sqrt < 0) return ERR_SPEED;
	cf = A::sqrt(T(0.5) * (speeds2 + A::sqrt(to_sqrt)));
	return ERR_OK;
}

/* ideal-MHD x-flux with background-field splitting, source/mhd/common.hpp:138-244 */
template <class T> __device__ __forceinline__ int physical_flux(const Side<T>& s,
	const double U[8], const T bg[3], T inv_mu0, T F[8])
{
	if (!__builtin_isnormal(U[RHO]) || U[RHO] <= 0) return ERR_DENSITY;
	if (!__builtin_isfinite(U[M0])) return ERR_FLUX;
	const T* B = s.u + B0;
	const T t[3] = {B[0] + bg[0], B[1] + bg[1], B[2] + bg[2]};
	const T
		inv = inv_mu0,
		pB0 = T(0.5) * inv * (bg[0] * bg[0] + bg[1] * bg[1] + bg[2] * bg[2]),
		pB1 = T(0.5) * inv * (B[0] * B[0] + B[1] * B[1] + B[2] * B[2]),
		pBt = T(0.5) * inv * (t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
	const T* v = s.v;
	F[RHO] = s.u[M0];
	for (int i = 0; i < 3; i++) {
		F[M0 + i] = s.u[M0 + i] * v[0] + (t[i] * t[0] + bg[i] * (-bg[0])) * (-inv);
	}
	F[M0] += s.p + pBt - pB0;
	F[NRJ]
		= v[0] * (s.u[NRJ] + s.p + pB1)
		- B[0] * dot3(v, B) * inv
		+ inv * (
			B[1] * (v[0] * bg[1] - v[1] * bg[0])
			+ B[2] * (v[0] * bg[2] - v[2] * bg[0]));
	for (int i = 0; i < 3; i++) {
		F[B0 + i] = t[i] * v[0] + v[i] * (-t[0]);
	}
	F[B0] = 0;
	return ERR_OK;
}

/* source/mhd/rusanov.hpp:59-214 */
template <class T> __device__ __forceinline__ int flux_rusanov(const double Ld[8],
	const double Rd[8], const T bg[3], const Phys& ph, T F[8], T& s_out)
{
	using A = Arith<T>;
	const T gamma = (T)ph.gamma, mu0 = (T)ph.mu0, inv_mu0 = (T)ph.inv_mu0;
	if (!__builtin_isnormal(Ld[RHO]) || Ld[RHO] < 0) return ERR_DENSITY;
	if (!__builtin_isnormal(Rd[RHO]) || Rd[RHO] < 0) return ERR_DENSITY;
	Side<T> L, R;
	make_side<T>(Ld, ph, L);
	make_side<T>(Rd, ph, R);
	T cfL, cfR;
	int e = fast_speed<T>(L, Ld, bg, gamma, mu0, cfL);
	if (e) return e;
	e = fast_speed<T>(R, Rd, bg, gamma, mu0, cfR);
	if (e) return e;
	const T
		max_fast = smax(cfL, cfR),
		s = smax(fabs(A::div(L.u[M0], L.u[RHO])) + max_fast,
			fabs(A::div(R.u[M0], R.u[RHO])) + max_fast);
	if (!A::isfinite(max_fast) || !A::isfinite(s)) return ERR_SPEED;
	T FL[8], FR[8];
	e = physical_flux<T>(L, Ld, bg, inv_mu0, FL);
	if (e) return e;
	e = physical_flux<T>(R, Rd, bg, inv_mu0, FR);
	if (e) return e;
	if (!A::isfinite(FL[RHO]) || !A::isfinite(FR[RHO])) return ERR_FLUX;

	F[RHO] = (FL[RHO] + FR[RHO]) / T(2) - (R.u[RHO] - L.u[RHO]) * (s / T(2));
	F[NRJ] = (FL[NRJ] + FR[NRJ]) / T(2) - (R.u[NRJ] - L.u[NRJ]) * (s / T(2));
	const T hs = -s / T(2);
	for (int i = 0; i < 3; i++) {
		F[M0 + i] = (FL[M0 + i] + FR[M0 + i]) * T(0.5) + (R.u[M0 + i] - L.u[M0 + i]) * hs;
		F[B0 + i] = (FL[B0 + i] + FR[B0 + i]) * T(0.5) + (R.u[B0 + i] - L.u[B0 + i]) * hs;
	}
	if (!A::isfinite(F[RHO])) return ERR_FLUX;
	s_out = s;
	return ERR_OK;
}

/* source/mhd/hll_athena.hpp:57-275 */
template <class T> __device__ __forceinline__ int flux_hll(const double Ld[8],
	const double Rd[8], const T bg[3], const Phys& ph, T F[8], T& s_out)
{
	using A = Arith<T>;
	const T gamma = (T)ph.gamma, mu0 = (T)ph.mu0, inv_mu0 = (T)ph.inv_mu0;
	if (Ld[RHO] <= 0 || Rd[RHO] <= 0) return ERR_DENSITY; /* get_pressure throws */
	Side<T> L, R;
	make_side<T>(Ld, ph, L);
	make_side<T>(Rd, ph, R);
	const T
		pmL = A::div(dot3(L.u + B0, L.u + B0), (T)ph.two_mu0),
		pmR = A::div(dot3(R.u + B0, R.u + B0), (T)ph.two_mu0);
	T cfL, cfR;
	int e = fast_speed<T>(L, Ld, bg, gamma, mu0, cfL);
	if (e) return e;
	e = fast_speed<T>(R, Rd, bg, gamma, mu0, cfR);
	if (e) return e;
	const T
		ms = smax(cfL, cfR),
		msn = (L.v[0] <= R.v[0]) ? L.v[0] - ms : R.v[0] - ms,
		msp = (L.v[0] <= R.v[0]) ? R.v[0] + ms : L.v[0] + ms,
		bm = smin(msn, T(0)),
		bp = smax(msp, T(0));
	if (!__builtin_isnormal(L.p64) || L.p64 < 0) return ERR_PRESSURE;
	if (!A::isfinite(pmL) || pmL < 0) return ERR_PRESSURE;
	if (!A::isfinite(msn)) return ERR_SPEED;
	if (!__builtin_isnormal(R.p64) || R.p64 < 0) return ERR_PRESSURE;
	if (!A::isfinite(pmR) || pmR < 0) return ERR_PRESSURE;
	if (!A::isfinite(msp)) return ERR_SPEED;

	if (!A::isnormal(bp - bm) || bp - bm < 0) {
		for (int q = 0; q < 8; q++) F[q] = 0;
		s_out = 0;
		return ERR_OK;
	}
	T FL[8], FR[8];
	e = physical_flux<T>(L, Ld, bg, inv_mu0, FL);
	if (e) return e;
	e = physical_flux<T>(R, Rd, bg, inv_mu0, FR);
	if (e) return e;
	FL[RHO] -= L.u[RHO] * bm;
	FL[NRJ] -= L.u[NRJ] * bm;
	FR[RHO] -= R.u[RHO] * bp;
	FR[NRJ] -= R.u[NRJ] * bp;
	for (int i = 0; i < 3; i++) {
		FL[M0 + i] = FL[M0 + i] + L.u[M0 + i] * (-bm);
		FL[B0 + i] = FL[B0 + i] + L.u[B0 + i] * (-bm);
		FR[M0 + i] = FR[M0 + i] + R.u[M0 + i] * (-bp);
		FR[B0 + i] = FR[B0 + i] + R.u[B0 + i] * (-bp);
	}
	FR[B0] = FL[B0] = 0;
	const T den = bp - bm, sum = bp + bm;
	F[RHO] = (FL[RHO] + FR[RHO]) / T(2) + A::div((FL[RHO] - FR[RHO]) * sum, den) / T(2.0);
	F[NRJ] = (FL[NRJ] + FR[NRJ]) / T(2) + A::div((FL[NRJ] - FR[NRJ]) * sum, den) / T(2.0);
	const T w = A::div(sum, den) / T(2.0);
	for (int i = 0; i < 3; i++) {
		F[M0 + i] = (FL[M0 + i] + FR[M0 + i]) * T(0.5) + (FL[M0 + i] - FR[M0 + i]) * w;
		F[B0 + i] = (FL[B0 + i] + FR[B0 + i]) * T(0.5) + (FL[B0 + i] - FR[B0 + i]) * w;
	}
	s_out = smax(fabs(bp), fabs(bm));
	return ERR_OK;
}

/*
HLL and Rusanov split into a per-cell part and a per-face part, for faces
without a background field: fast speed, physical flux and the side checks
depend on one cell only, so a kernel may compute them once per cell and
pass them to both of the cell's faces along a direction.  The per-face
parts return exactly what flux_hll / flux_rusanov return for the same two
states (the same calls on the same inputs, the checks in the same order).
*/
template <class T> struct FluxCell {
	Side<T> s;
	T cf;        /* fast speed */
	T aux;       /* HLL: B^2 / 2 mu0; Rusanov: |m_x / rho| */
	T F[8];      /* physical flux */
	int e_rho;   /* the solver's density check */
	int e_fs;    /* fast_speed's error */
	int e_pf;    /* physical_flux's error */
};

template <int SOLVER, class T> __device__ __forceinline__ void flux_cell(const double U[8],
	const Phys& ph, FluxCell<T>& c)
{
	using A = Arith<T>;
	const T zero[3] = {0, 0, 0};
	if constexpr (SOLVER == RUSANOV) {
		c.e_rho = (!__builtin_isnormal(U[RHO]) || U[RHO] < 0) ? ERR_DENSITY : ERR_OK;
	} else {
		c.e_rho = U[RHO] <= 0 ? ERR_DENSITY : ERR_OK;
	}
	make_side<T>(U, ph, c.s);
	c.cf = 0;
	c.e_fs = fast_speed<T>(c.s, U, zero, (T)ph.gamma, (T)ph.mu0, c.cf);
	if constexpr (SOLVER == RUSANOV) {
		c.aux = fabs(A::div(c.s.u[M0], c.s.u[RHO]));
	} else {
		c.aux = A::div(dot3(c.s.u + B0, c.s.u + B0), (T)ph.two_mu0);
	}
	for (int q = 0; q < 8; q++) c.F[q] = 0;
	c.e_pf = physical_flux<T>(c.s, U, zero, (T)ph.inv_mu0, c.F);
}

/* flux_rusanov's per-face part */
template <class T> __device__ __forceinline__ int flux_rusanov_face(const FluxCell<T>& L,
	const FluxCell<T>& R, T F[8], T& s_out)
{
	using A = Arith<T>;
	if (L.e_rho) return L.e_rho;
	if (R.e_rho) return R.e_rho;
	if (L.e_fs) return L.e_fs;
	if (R.e_fs) return R.e_fs;
	const T
		max_fast = smax(L.cf, R.cf),
		s = smax(L.aux + max_fast, R.aux + max_fast);
	if (!A::isfinite(max_fast) || !A::isfinite(s)) return ERR_SPEED;
	if (L.e_pf) return L.e_pf;
	if (R.e_pf) return R.e_pf;
	const T* FL = L.F;
	const T* FR = R.F;
	if (!A::isfinite(FL[RHO]) || !A::isfinite(FR[RHO])) return ERR_FLUX;
	F[RHO] = (FL[RHO] + FR[RHO]) / T(2) - (R.s.u[RHO] - L.s.u[RHO]) * (s / T(2));
	F[NRJ] = (FL[NRJ] + FR[NRJ]) / T(2) - (R.s.u[NRJ] - L.s.u[NRJ]) * (s / T(2));
	const T hs = -s / T(2);
	for (int i = 0; i < 3; i++) {
		F[M0 + i] = (FL[M0 + i] + FR[M0 + i]) * T(0.5) + (R.s.u[M0 + i] - L.s.u[M0 + i]) * hs;
		F[B0 + i] = (FL[B0 + i] + FR[B0 + i]) * T(0.5) + (R.s.u[B0 + i] - L.s.u[B0 + i]) * hs;
	}
	if (!A::isfinite(F[RHO])) return ERR_FLUX;
	s_out = s;
	return ERR_OK;
}

/* flux_hll's per-face part */
template <class T> __device__ __forceinline__ int flux_hll_face(const FluxCell<T>& Lc,
	const FluxCell<T>& Rc, T F[8], T& s_out)
{
	using A = Arith<T>;
	const Side<T>& L = Lc.s;
	const Side<T>& R = Rc.s;
	if (Lc.e_rho || Rc.e_rho) return ERR_DENSITY;
	if (Lc.e_fs) return Lc.e_fs;
	if (Rc.e_fs) return Rc.e_fs;
	const T
		pmL = Lc.aux,
		pmR = Rc.aux,
		ms = smax(Lc.cf, Rc.cf),
		msn = (L.v[0] <= R.v[0]) ? L.v[0] - ms : R.v[0] - ms,
		msp = (L.v[0] <= R.v[0]) ? R.v[0] + ms : L.v[0] + ms,
		bm = smin(msn, T(0)),
		bp = smax(msp, T(0));
	if (!__builtin_isnormal(L.p64) || L.p64 < 0) return ERR_PRESSURE;
	if (!A::isfinite(pmL) || pmL < 0) return ERR_PRESSURE;
	if (!A::isfinite(msn)) return ERR_SPEED;
	if (!__builtin_isnormal(R.p64) || R.p64 < 0) return ERR_PRESSURE;
	if (!A::isfinite(pmR) || pmR < 0) return ERR_PRESSURE;
	if (!A::isfinite(msp)) return ERR_SPEED;

	if (!A::isnormal(bp - bm) || bp - bm < 0) {
		for (int q = 0; q < 8; q++) F[q] = 0;
		s_out = 0;
		return ERR_OK;
	}
	if (Lc.e_pf) return Lc.e_pf;
	if (Rc.e_pf) return Rc.e_pf;
	T FL[8], FR[8];
	for (int q = 0; q < 8; q++) {
		FL[q] = Lc.F[q];
		FR[q] = Rc.F[q];
	}
	FL[RHO] -= L.u[RHO] * bm;
	FL[NRJ] -= L.u[NRJ] * bm;
	FR[RHO] -= R.u[RHO] * bp;
	FR[NRJ] -= R.u[NRJ] * bp;
	for (int i = 0; i < 3; i++) {
		FL[M0 + i] = FL[M0 + i] + L.u[M0 + i] * (-bm);
		FL[B0 + i] = FL[B0 + i] + L.u[B0 + i] * (-bm);
		FR[M0 + i] = FR[M0 + i] + R.u[M0 + i] * (-bp);
		FR[B0 + i] = FR[B0 + i] + R.u[B0 + i] * (-bp);
	}
	FR[B0] = FL[B0] = 0;
	const T den = bp - bm, sum = bp + bm;
	F[RHO] = (FL[RHO] + FR[RHO]) / T(2) + A::div((FL[RHO] - FR[RHO]) * sum, den) / T(2.0);
	F[NRJ] = (FL[NRJ] + FR[NRJ]) / T(2) + A::div((FL[NRJ] - FR[NRJ]) * sum, den) / T(2.0);
	const T w = A::div(sum, den) / T(2.0);
	for (int i = 0; i < 3; i++) {
		F[M0 + i] = (FL[M0 + i] + FR[M0 + i]) * T(0.5) + (FL[M0 + i] - FR[M0 + i]) * w;
		F[B0 + i] = (FL[B0 + i] + FR[B0 + i]) * T(0.5) + (FL[B0 + i] - FR[B0 + i]) * w;
	}
	s_out = smax(fabs(bp), fabs(bm));
	return ERR_OK;
}

template <int SOLVER, class T> __device__ __forceinline__ int flux_face(const FluxCell<T>& L,
	const FluxCell<T>& R, T F[8], T& s_out)
{
	if constexpr (SOLVER == RUSANOV) return flux_rusanov_face<T>(L, R, F, s_out);
	else return flux_hll_face<T>(L, R, F, s_out);
}

/* one side's star state, source/mhd/hlld_athena.hpp:316-401 (eqns 43-48) */
template <class T> __device__ __forceinline__ void hlld_star(const Side<T>& s, T nm, T sfd,
	T smd, T sm, T ds, T Bx, T Bx2, T inv, T ptst, T S[8], T& vdbs)
{
	using A = Arith<T>;
	const T tmp = s.u[RHO] * sfd * smd - Bx2 * inv;
	if (fabs(tmp) < A::eps * ptst) {
		for (int i = 0; i < 3; i++) {
			S[M0 + i] = s.v[i] * ds;
			S[B0 + i] = s.u[B0 + i];
		}
	} else {
		const T tmp1 = A::div(Bx * (sm - s.v[0]) * inv, tmp);
		for (int i = 0; i < 3; i++) S[M0 + i] = (s.v[i] + s.u[B0 + i] * (-tmp1)) * ds;
		const T tmp2 = A::div(s.u[RHO] * (sfd * sfd) - Bx2 * inv, tmp);
		for (int i = 0; i < 3; i++) S[B0 + i] = s.u[B0 + i] * tmp2;
	}
	S[RHO] = ds;
	S[M0] = ds * sm;
	S[B0] = Bx;
	vdbs = A::div(dot3(S + M0, S + B0), ds);
	S[NRJ] = A::div(
		sfd * s.u[NRJ]
		- (s.p + nm) * s.v[0]
		+ ptst * sm
		+ Bx * inv * (dot3(s.v, s.u + B0) - vdbs), smd);
}

/*
HLLD's per-cell part: the state checks, thermal pressure, 1 / rho and fast
speed of one side in the face frame.  HLLD ignores the background field, so
the part is the same for both faces of a cell along one direction and a
kernel may compute it once per cell.  Returns the stage of the first failed
check (0 none, 1 density, 2 pressure, 3 fast speed) with its error code.
flux_hlld checks the density of both sides, then the pressures, then the
fast speeds, so the error of a face is that of the side whose check fails
first, the left side on a tie (first_side_error).
*/
template <class T> __device__ __forceinline__ int hlld_side(const double U[8], const Phys& ph,
	Side<T>& s, T& cf, int& code)
{
	code = ERR_OK;
	if (U[RHO] <= 0) {
		code = ERR_DENSITY;
		return 1;
	}
	make_side<T>(U, ph, s);
	if (s.p64 <= 0) {
		code = ERR_PRESSURE;
		return 2;
	}
	const T zero[3] = {0, 0, 0};
	int e;
	if constexpr (std::is_same_v<T, float>) {
		e = fast_speed32(s, U, zero, (float)ph.gamma, (float)ph.inv_mu0, cf);
	} else {
		e = fast_speed<T>(s, U, zero, (T)ph.gamma, (T)ph.mu0, cf);
	}
	if (e) {
		code = e;
		return 3;
	}
	return 0;
}

__device__ __forceinline__ int first_side_error(int stage_l, int code_l, int stage_r, int code_r)
{
	return (stage_l && (!stage_r || stage_l <= stage_r)) ? code_l : code_r;
}

/*
Per-cell part of HLLD in one face direction, as a record passed between the
threads of a kernel: the FP64 state in the face frame, the FP64 thermal
pressure, the fast speed and hlld_side's stage and error code; for the
mixed path also 1 / rho in FP64.  unpack_side rebuilds the Side exactly as
make_side built it.
*/
template <class T> struct XPart {
	double u[8];
	double p64;
	double ir64;   /* 1 / rho in FP64 (mixed path) */
	double cf;     /* fast speed (the mixed path's to ~1e-14 of the FP64 one) */
	int sc;        /* stage * 16 + code */
};

/* sqrt(x) within ~1e-14 relative: FP32 reciprocal square root refined by
   one Newton step in FP64 (the error squares: (1e-7)^2), for x in FP32's
   normal range; IEEE FP64 square root outside it */
__device__ __forceinline__ double sqrt_refined(double x)
{
	if (!(x > 1e-30 && x < 1e30)) return __builtin_sqrt(x);
	const double y0 = (double)__builtin_amdgcn_rsqf((float)x);
	const double y1 = y0 * (1.5 - (0.5 * x) * (y0 * y0));
	return x * y1;
}

/*
Fast magnetosonic speed of the mixed path in FP64 without an FP64 division
or square root sequence: c_s^2 = gamma p / rho, v_A^2 = |B|^2 / (mu0 rho),
v_An^2 = B_n^2 / (mu0 rho) with the cell's FP64 1 / rho, and c_f^2 =
(c_s^2 + v_A^2 + sqrt((c_s^2 + v_A^2)^2 - 4 c_s^2 v_An^2)) / 2, the
quantity of get_fast_magnetosonic_speed (source/mhd/common.hpp:415-491)
with its checks; square roots by sqrt_refined.  Accurate wave speeds keep
the mixed path's CFL time step equal to the FP64 one to ~1e-14 and its
flux error proportional to the jumps.
*/
__device__ __forceinline__ int fast_speed_mixed(const double U[8], double p64, double ir64,
	const Phys& ph, double& cf)
{
	if (!__builtin_isfinite(U[B0])) return ERR_FIELD;
	if (!__builtin_isnormal(U[RHO]) || U[RHO] < 0) return ERR_DENSITY;
	if (!__builtin_isnormal(p64) || p64 < 0) return ERR_PRESSURE;
	const double im = ir64 * ph.inv_mu0;
	const double cs2 = ph.gamma * p64 * ir64;
	const double van2 = U[B0] * U[B0] * im;
	const double va2 = (U[B0] * U[B0] + U[B1] * U[B1] + U[B2] * U[B2]) * im;
	if (!(va2 > 0)) {
		cf = sqrt_refined(cs2);
		return ERR_OK;
	}
	const double sum = cs2 + va2;
	if (!__builtin_isnormal(sum) || sum < 0) return ERR_SPEED;
	const double disc = sum * sum - 4.0 * cs2 * van2;
	if (!__builtin_isnormal(disc) || disc < 0) return ERR_SPEED;
	cf = sqrt_refined(0.5 * (sum + sqrt_refined(disc)));
	return ERR_OK;
}

template <class T> __device__ __forceinline__ void hlld_part(const double U[8], const Phys& ph,
	XPart<T>& x)
{
#pragma unroll
	for (int q = 0; q < 8; q++) x.u[q] = U[q];
	int stage, code = ERR_OK;
	if constexpr (std::is_same_v<T, float>) {
		/* hlld_side's checks in its order: density, pressure, fast speed */
		x.p64 = 0.0;
		x.ir64 = 0.0;
		x.cf = 0.0;
		if (U[RHO] <= 0) {
			stage = 1;
			code = ERR_DENSITY;
		} else {
			const double p = pressure64(U, ph.gm1, ph.half_inv_mu0);
			if (p <= 0) {
				stage = 2;
				code = ERR_PRESSURE;
			} else {
				const double ir = 1 / U[RHO];
				double cf = 0;
				code = fast_speed_mixed(U, p, ir, ph, cf);
				stage = code ? 3 : 0;
				if (!stage) {
					x.p64 = p;
					x.ir64 = ir;
					x.cf = cf;
				}
			}
		}
	} else {
		Side<T> s;
		T cf = 0;
		stage = hlld_side<T>(U, ph, s, cf, code);
		x.p64 = stage == 0 ? s.p64 : 0.0;
		x.ir64 = 0.0;
		x.cf = cf;
	}
	x.sc = stage * 16 + code;
}

template <class T> __device__ __forceinline__ void unpack_side(const XPart<T>& x, const Phys& ph,
	Side<T>& s)
{
#pragma unroll
	for (int q = 0; q < 8; q++) s.u[q] = (T)x.u[q];
	s.rho_pos = x.u[RHO] > 0;
	s.p64 = x.p64;
	s.p = (T)x.p64;
	if constexpr (std::is_same_v<T, float>) {
		s.pt64 = s.p64 + ph.half_inv_mu0 * (x.u[B0] * x.u[B0] + x.u[B1] * x.u[B1] + x.u[B2] * x.u[B2]);
	} else {
		s.pt64 = 0;
	}
	const T inv = Arith<T>::rcp(s.u[RHO]);
	s.ir = inv;
	s.v[0] = s.u[M0] * inv;
	s.v[1] = s.u[M1] * inv;
	s.v[2] = s.u[M2] * inv;
}

/* FP64 ideal-MHD x-flux without background field (get_flux,
   source/mhd/common.hpp:138-244 with bg = 0) from the state, 1 / rho and
   the total pressure */
__device__ __forceinline__ void physical_flux64(const double U[8], const double v[3], double pt,
	double inv_mu0, double F[8])
{
	const double* B = U + B0;
	F[RHO] = U[M0];
#pragma unroll
	for (int i = 0; i < 3; i++) F[M0 + i] = U[M0 + i] * v[0] - B[i] * B[0] * inv_mu0;
	F[M0] += pt;
	F[NRJ] = v[0] * (U[NRJ] + pt) - B[0] * (v[0] * B[0] + v[1] * B[1] + v[2] * B[2]) * inv_mu0;
	F[B0] = 0;
#pragma unroll
	for (int i = 1; i < 3; i++) F[B0 + i] = B[i] * v[0] - v[i] * B[0];
}

/* one side's star state of the mixed HLLD as differences from the side's
   own state (eqns 43-48 of the HLLD method restated): with d = S - u,
   e = S - S_M and delta = S_M - u,
     rho* - rho = rho delta / e,  m_n* - m_n = (rho* - rho) S_M + rho delta,
     v_t* - v_t = -B_t tmp1,  B_t* - B_t = B_t rho d delta / tmp,
     E* - E = (delta (E + p_T + rho d S_M) + B_n (v.B - v*.B*) / mu0) / e,
   tmp = rho d e - B_n^2 / mu0, tmp1 = B_n delta / (mu0 tmp); every term is
   a product with delta or with a difference of the two sides, so FP32
   rounding errors scale with the jump, not with the state */
struct MixedStar {
	float dU[8];     /* U* - U (dU[B0] unused) */
	float vs[3];     /* v* - v of the transverse components (vs[0] unused) */
	float Bs[3];     /* B* of the transverse components */
	float dB[3];     /* B* - B of the transverse components */
	float rho_s;     /* rho* */
};

__device__ __forceinline__ void mixed_star(const XPart<float>& P, const double v64[3], double pt64,
	float d, float e, float delta, float sm, float Bx, float dbx, float inv, float ptst, MixedStar& S)
{
	using A = Arith<float>;
	const float rho = (float)P.u[RHO];
	const float ie = A::rcp(e);
	const float dr = rho * delta * ie;
	S.rho_s = rho + dr;
	S.dU[RHO] = dr;
	S.dU[M0] = dr * sm + rho * delta;
	const float v[3] = {(float)v64[0], (float)v64[1], (float)v64[2]};
	const float B[3] = {(float)P.u[B0], (float)P.u[B1], (float)P.u[B2]};
	const float tmp = rho * d * e - Bx * Bx * inv;
	float tmp1 = 0;
	if (fabsf(tmp) < A::eps * ptst) {
		for (int i = 1; i < 3; i++) {
			S.dU[M0 + i] = v[i] * dr;
			S.dB[i] = 0;
			S.vs[i] = 0;
		}
	} else {
		const float it = A::rcp(tmp);
		tmp1 = Bx * delta * inv * it;
		const float g = rho * d * delta * it;
		for (int i = 1; i < 3; i++) {
			S.vs[i] = -B[i] * tmp1;
			S.dU[M0 + i] = v[i] * dr + S.vs[i] * S.rho_s;
			S.dB[i] = B[i] * g;
		}
	}
	/* v.B - v*.B*, with dbx = B_n(face) - B_n(cell) */
	float vb = -delta * Bx - v[0] * dbx;
	for (int i = 1; i < 3; i++) {
		S.Bs[i] = B[i] + S.dB[i];
		S.dU[B0 + i] = S.dB[i];
		vb += -v[i] * S.dB[i] + tmp1 * B[i] * S.Bs[i];
	}
	S.dU[B0] = 0;
	S.dU[NRJ] = (delta * ((float)P.u[NRJ] + (float)pt64 + rho * d * sm) + Bx * inv * vb) * ie;
}

/*
HLLD of source/mhd/hlld_athena.hpp:54-559 for the mixed path, written as
F = F_L + S_L (U*_L - U_L) [+ S*_L (U**_L - U*_L)] (and the mirror forms
from F_R) with the physical fluxes F_L, F_R, the velocities, the jumps of
the normal velocity, total pressure and transverse field and velocity in
FP64, and the wave speeds and the star and double-star differences in
FP32 in the delta form of mixed_star: the FP32 rounding then scales with
the jumps across the face, not with |F|.  The same branches and degenerate
cases as the FP64 solver; returns FP64 fluxes.  Both parts must have
passed hlld_side's checks.
*/
__device__ __forceinline__ int flux_hlld_mixed_face(const XPart<float>& L, const XPart<float>& R,
	const Phys& ph, double F[8], double& s_out)
{
	using A = Arith<float>;
	if (!__builtin_isnormal(L.u[RHO]) || L.u[RHO] <= 0) return ERR_DENSITY;
	if (!__builtin_isfinite(L.u[M0])) return ERR_FLUX;
	if (!__builtin_isnormal(R.u[RHO]) || R.u[RHO] <= 0) return ERR_DENSITY;
	if (!__builtin_isfinite(R.u[M0])) return ERR_FLUX;
	const double inv64 = ph.inv_mu0;
	const double vL64[3] = {L.u[M0] * L.ir64, L.u[M1] * L.ir64, L.u[M2] * L.ir64};
	const double vR64[3] = {R.u[M0] * R.ir64, R.u[M1] * R.ir64, R.u[M2] * R.ir64};
	const double ptL64 = L.p64 + ph.half_inv_mu0 * (L.u[B0] * L.u[B0] + L.u[B1] * L.u[B1] + L.u[B2] * L.u[B2]);
	const double ptR64 = R.p64 + ph.half_inv_mu0 * (R.u[B0] * R.u[B0] + R.u[B1] * R.u[B1] + R.u[B2] * R.u[B2]);
	double FL[8], FR[8];
	physical_flux64(L.u, vL64, ptL64, inv64, FL);
	physical_flux64(R.u, vR64, ptR64, inv64, FR);
	if (!__builtin_isfinite(FL[RHO]) || !__builtin_isfinite(FR[RHO])) return ERR_FLUX;

	/* outer wave speeds in FP64 (accurate fast speeds): the CFL speed and
	   the supersonic tests of the FP64 solver */
	const double ms = L.cf > R.cf ? L.cf : R.cf;
	const double dv64 = vR64[0] - vL64[0];
	const bool lo = vL64[0] <= vR64[0];
	const double
		msn = lo ? vL64[0] - ms : vR64[0] - ms,
		msp = lo ? vR64[0] + ms : vL64[0] + ms;
	if (msn >= 0.0) {
		for (int q = 0; q < 8; q++) F[q] = FL[q];
		s_out = fabs(msp);
		return ERR_OK;
	} else if (msp <= 0.0) {
		for (int q = 0; q < 8; q++) F[q] = FR[q];
		s_out = fabs(msn);
		return ERR_OK;
	}
	const float inv = (float)ph.inv_mu0, inv_sqrt = (float)ph.inv_sqrt_mu0;
	const float vL = (float)vL64[0];
	const float dv = (float)dv64;
	/* S_L - u_L and S_R - u_R */
	const float sfdn = (float)(lo ? -ms : dv64 - ms), sfdp = (float)(lo ? ms : ms - dv64);
	const float rhoL = (float)L.u[RHO], rhoR = (float)R.u[RHO];
	const float dpt = (float)(ptR64 - ptL64);
	/* S_M - u_L and S_M - u_R (eqn 38 of the method, differenced) */
	const float iden = A::rcp(sfdp * rhoR - sfdn * rhoL);
	const float delL = (sfdp * rhoR * dv - dpt) * iden, delR = (sfdn * rhoL * dv - dpt) * iden;
	const float sm = vL + delL;
	const double sm64 = vL64[0] + (double)delL;
	const float smdn = sfdn - delL, smdp = sfdp - delR;
	const double Bx64 = (L.u[B0] + R.u[B0]) * 0.5;
	const float Bx = (float)Bx64, Bx2 = Bx * Bx;
	const float ptst = (float)ptL64 + rhoL * sfdn * delL;
	MixedStar Sn, Sp;
	mixed_star(L, vL64, ptL64, sfdn, smdn, delL, sm, Bx, (float)(Bx64 - L.u[B0]), inv, ptst, Sn);
	mixed_star(R, vR64, ptR64, sfdp, smdp, delR, sm, Bx, (float)(Bx64 - R.u[B0]), inv, ptst, Sp);
	const float dssn = A::sqrt(Sn.rho_s), dssp = A::sqrt(Sp.rho_s);
	const double
		ssn = sm64 - (double)(fabsf(Bx) * inv_sqrt * A::rcp(dssn)),
		ssp = sm64 + (double)(fabsf(Bx) * inv_sqrt * A::rcp(dssp));

	/* double-star states minus single-star states, eqns 59-63; they are
	   used where the FP64 solver uses them (its threshold: with a weak
	   normal field the region between S*_L and S*_R is thin but its states
	   differ from the single-star ones by the jumps) */
	float D2n[8] = {0, 0, 0, 0, 0, 0, 0, 0}, D2p[8] = {0, 0, 0, 0, 0, 0, 0, 0};
	if (!(Bx2 * 0.5f < (float)DBL_EPSILON * ptst)) {
		const float isds = A::rcp(dssn + dssp);
		const float sgn = (Bx > 0) ? 1.0f : -1.0f;
		const float c2 = sgn * dssn * dssp * (float)ph.sqrt_mu0;
		float vdn = 0, vdp = 0;   /* v*.B* - v**.B** on each side */
		for (int i = 1; i < 3; i++) {
			/* v*_R - v*_L and B*_R - B*_L from the FP64 jumps */
			const float jv = (float)(vR64[i] - vL64[i]) + (Sp.vs[i] - Sn.vs[i]);
			const float jb = (float)(R.u[B0 + i] - L.u[B0 + i]) + (Sp.dB[i] - Sn.dB[i]);
			const float dvn = (dssp * jv + jb * sgn * inv_sqrt) * isds;
			const float dvp = (-dssn * jv + jb * sgn * inv_sqrt) * isds;
			const float dbn = (dssn * jb + jv * c2) * isds;
			const float dbp = (-dssp * jb + jv * c2) * isds;
			D2n[M0 + i] = Sn.rho_s * dvn;
			D2p[M0 + i] = Sp.rho_s * dvp;
			D2n[B0 + i] = dbn;
			D2p[B0 + i] = dbp;
			const float vsn = (float)vL64[i] + Sn.vs[i], vsp = (float)vR64[i] + Sp.vs[i];
			vdn -= vsn * dbn + dvn * Sn.Bs[i] + dvn * dbn;
			vdp -= vsp * dbp + dvp * Sp.Bs[i] + dvp * dbp;
		}
		D2n[NRJ] = -dssn * sgn * vdn * inv_sqrt;
		D2p[NRJ] = dssp * sgn * vdp * inv_sqrt;
	}
	/* F = F_L + S_L (U*_L - U_L) [+ S*_L (U**_L - U*_L)] or the mirror form,
	   the FP32 differences scaled by FP64 wave speeds */
	if (ssn >= 0.0) {
		for (int q = 0; q < 8; q++) F[q] = FL[q] + (double)Sn.dU[q] * msn;
	} else if (sm64 >= 0.0) {
		for (int q = 0; q < 8; q++) F[q] = FL[q] + ((double)D2n[q] * ssn + (double)Sn.dU[q] * msn);
	} else if (ssp > 0.0) {
		for (int q = 0; q < 8; q++) F[q] = FR[q] + ((double)D2p[q] * ssp + (double)Sp.dU[q] * msp);
	} else {
		for (int q = 0; q < 8; q++) F[q] = FR[q] + (double)Sp.dU[q] * msp;
	}
	F[B0] = 0;
	s_out = fabs(msn) > fabs(msp) ? fabs(msn) : fabs(msp);
	return ERR_OK;
}

/* the mixed HLLD of one face from the two states */
__device__ __forceinline__ int flux_hlld_mixed(const double Ld[8], const double Rd[8], const Phys& ph,
	double F[8], double& s_out)
{
	XPart<float> L, R;
	hlld_part<float>(Ld, ph, L);
	hlld_part<float>(Rd, ph, R);
	const int sl = L.sc >> 4, sr = R.sc >> 4;
	if (sl | sr) return first_side_error(sl, L.sc & 15, sr, R.sc & 15);
	return flux_hlld_mixed_face(L, R, ph, F, s_out);
}

/* HLLD's per-face part in FP64, after both sides passed hlld_side */
template <class T> __device__ __forceinline__ int flux_hlld_face(const Side<T>& L, const Side<T>& R,
	T cfL, T cfR, const double Ld[8], const double Rd[8], const Phys& ph, T F[8], T& s_out)
{
	using A = Arith<T>;
	const T mu0 = (T)ph.mu0, inv_mu0 = (T)ph.inv_mu0;
	const T zero[3] = {0, 0, 0};
	const T
		inv = inv_mu0,
		inv_sqrt = (T)ph.inv_sqrt_mu0,
		Bx = (L.u[B0] + R.u[B0]) / T(2),
		Bx2 = Bx * Bx;
	int e;
	const T ms = smax(cfL, cfR);
	const T* vL = L.v;
	const T* vR = R.v;
	const T
		msn = (vL[0] <= vR[0]) ? vL[0] - ms : vR[0] - ms,
		msp = (vL[0] <= vR[0]) ? vR[0] + ms : vL[0] + ms;

	T FL[8], FR[8];
	e = physical_flux<T>(L, Ld, zero, inv_mu0, FL);
	if (e) return e;
	e = physical_flux<T>(R, Rd, zero, inv_mu0, FR);
	if (e) return e;

	if (msn >= T(0)) {
		for (int q = 0; q < 8; q++) F[q] = FL[q];
		s_out = fabs(msp);
		return ERR_OK;
	} else if (msp <= T(0)) {
		for (int q = 0; q < 8; q++) F[q] = FR[q];
		s_out = fabs(msn);
		return ERR_OK;
	}

	const T
		sfdn = msn - vL[0],
		sfdp = msp - vR[0],
		nmn = dot3(L.u + B0, L.u + B0) * (inv / T(2)),
		nmp = dot3(R.u + B0, R.u + B0) * (inv / T(2)),
		sm = A::div(
			sfdp * R.u[M0]
			- sfdn * L.u[M0]
			- (R.p + nmp)
			+ (L.p + nmn),
			sfdp * R.u[RHO]
			- sfdn * L.u[RHO]),
		smdn = msn - sm,
		smdp = msp - sm,
		dsn = A::div(L.u[RHO] * sfdn, smdn),
		dssn = A::sqrt(dsn),
		dsp = A::div(R.u[RHO] * sfdp, smdp),
		dssp = A::sqrt(dsp),
		/* sqrt(rho* mu0): with mu0 = 1 the product is exact, so it is the
		   square root just taken (a uniform branch) */
		sqmn = ph.mu0 == 1.0 ? dssn : A::sqrt(dsn * mu0),
		sqmp = ph.mu0 == 1.0 ? dssp : A::sqrt(dsp * mu0),
		ssn = sm - A::div(fabs(Bx), sqmn),
		ssp = sm + A::div(fabs(Bx), sqmp),
		ptst = L.p + nmn + L.u[RHO] * sfdn * (sm - vL[0]);

	T Sn[8], Sp[8], vdbsn, vdbsp;
	hlld_star<T>(L, nmn, sfdn, smdn, sm, dsn, Bx, Bx2, inv, ptst, Sn, vdbsn);
	hlld_star<T>(R, nmp, sfdp, smdp, sm, dsp, Bx, Bx2, inv, ptst, Sp, vdbsp);

	T vsn[3], vsp[3];
	{
		const T in = A::rcp(dsn), ip = A::rcp(dsp);
		for (int i = 0; i < 3; i++) {
			vsn[i] = Sn[M0 + i] * in;
			vsp[i] = Sp[M0 + i] * ip;
		}
	}

	T S2n[8], S2p[8];
	if (Bx2 / T(2) < A::eps * ptst) {
		for (int q = 0; q < 8; q++) {
			S2n[q] = Sn[q];
			S2p[q] = Sp[q];
		}
	} else {
		const T isds = A::rcp(dssn + dssp);
		const T sgn = (Bx > 0) ? T(1) : T(-1);
		S2n[RHO] = Sn[RHO];
		S2p[RHO] = Sp[RHO];
		T v2[3];
		const T c1 = sgn * inv_sqrt;
		for (int i = 0; i < 3; i++) {
			v2[i] = ((vsn[i] * dssn + vsp[i] * dssp) + (Sp[B0 + i] - Sn[B0 + i]) * c1) * isds;
		}
		v2[0] = sm;
		for (int i = 0; i < 3; i++) {
			S2n[M0 + i] = v2[i] * Sn[RHO];
			S2p[M0 + i] = v2[i] * Sp[RHO];
		}
		S2n[M0] = Sn[M0];
		S2p[M0] = Sp[M0];
		const T c2 = sgn * dssn * dssp * (T)ph.sqrt_mu0;
		for (int i = 0; i < 3; i++) {
			const T b = ((Sp[B0 + i] * dssn + Sn[B0 + i] * dssp) + (vsp[i] - vsn[i]) * c2) * isds;
			S2n[B0 + i] = b;
			S2p[B0 + i] = b;
		}
		S2n[B0] = S2p[B0] = Bx;
		const T vdb2 = dot3(v2, S2n + B0);
		S2n[NRJ] = Sn[NRJ] - dssn * sgn * (vdbsn - vdb2) * inv_sqrt;
		S2p[NRJ] = Sp[NRJ] + dssp * sgn * (vdbsp - vdb2) * inv_sqrt;
	}

	if (ssn >= T(0)) {
		for (int q = 0; q < 8; q++) F[q] = FL[q] + (Sn[q] - L.u[q]) * msn;
	} else if (sm >= T(0)) {
		const T d = ssn - msn;
		for (int q = 0; q < 8; q++) F[q] = FL[q] + S2n[q] * ssn - Sn[q] * d - L.u[q] * msn;
	} else if (ssp > T(0)) {
		const T d = ssp - msp;
		for (int q = 0; q < 8; q++) F[q] = FR[q] + S2p[q] * ssp - Sp[q] * d - R.u[q] * msp;
	} else {
		for (int q = 0; q < 8; q++) F[q] = FR[q] + (Sp[q] - R.u[q]) * msp;
	}
	F[B0] = 0;
	s_out = smax(fabs(msn), fabs(msp));
	return ERR_OK;
}

/* source/mhd/hlld_athena.hpp:54-559 (ignores the background field, :77) in
   FP64; the mixed path's HLLD is flux_hlld_mixed */
template <class T> __device__ __forceinline__ int flux_hlld(const double Ld[8],
	const double Rd[8], const T* /*bg*/, const Phys& ph, T F[8], T& s_out)
{
	static_assert(std::is_same_v<T, double>, "FP64 HLLD; the mixed form is flux_hlld_mixed");
	Side<T> L, R;
	T cfL = 0, cfR = 0;
	int cL, cR;
	const int sL = hlld_side<T>(Ld, ph, L, cfL, cL);
	const int sR = hlld_side<T>(Rd, ph, R, cfR, cR);
	if (sL | sR) return first_side_error(sL, cL, sR, cR);
	return flux_hlld_face<T>(L, R, cfL, cfR, Ld, Rd, ph, F, s_out);
}

/*
Roe flux in scaled units (B / sqrt(mu0)), source/mhd/roe_athena.hpp:341-559,
with the eigensystem of :56-296 evaluated on the fly: the 7x7 left and right
eigenvector matrices are never materialised; the projection a = L dU and the
dissipation sum over the right eigenvectors read their entries from the
closed forms (eqs. B17-B29 of the Athena method paper).
*/
template <class T> struct RoeEsys {
	T v1, v2, v3, b2, b3, vsq, hp, g1, g2, x, d;
	T cf, cs, vax, bet2, bet3, bet2_star, bet3_star, bet_starsq, vbet;
	T alpha_f, alpha_s, sqrtd, isqrtd, s, qf, qs, af_prime, as_prime, afpbb, aspbb;
	T bt_star;
	/* left-eigenvector scalars */
	T nalpha_f, nalpha_s, cff, css, nqf, nqs, af, as, afpb, aspb, q2_star, q3_star, vqstr, norm2;
	T ev[7];

	__device__ __forceinline__ void init(T d_, T v1_, T v2_, T v3_, T h, T b1, T b2_, T b3_,
		T x_, T y, T gamma)
	{
		using A = Arith<T>;
		const T tiny = A::eps;
		d = d_; v1 = v1_; v2 = v2_; v3 = v3_; b2 = b2_; b3 = b3_; x = x_;
		g1 = gamma - T(1);
		g2 = gamma - T(2);
		const T di = A::rcp(d);
		vsq = v1 * v1 + v2 * v2 + v3 * v3;
		const T btsq = b2 * b2 + b3 * b3;
		const T bt_starsq = (g1 - g2 * y) * btsq;
		const T vaxsq = b1 * b1 * di;
		hp = h - (vaxsq + btsq * di);
		const T twid_asq = smax((g1 * (hp - T(0.5) * vsq) - g2 * x), tiny);
		const T ct2 = bt_starsq * di;
		const T tsum = vaxsq + ct2 + twid_asq;
		const T tdif = vaxsq + ct2 - twid_asq;
		const T cf2_cs2 = A::sqrt(tdif * tdif + T(4.0) * twid_asq * ct2);
		const T cfsq = T(0.5) * (tsum + cf2_cs2);
		cf = A::sqrt(cfsq);
		const T cssq = A::div(twid_asq * vaxsq, cfsq);
		cs = A::sqrt(cssq);
		const T bt = A::sqrt(btsq);
		bt_star = A::sqrt(bt_starsq);
		if (bt == T(0)) {
			bet2 = T(1);
			bet3 = T(0);
		} else {
			bet2 = A::div(b2, bt);
			bet3 = A::div(b3, bt);
		}
		bet2_star = A::div(bet2, A::sqrt(g1 - g2 * y));
		bet3_star = A::div(bet3, A::sqrt(g1 - g2 * y));
		bet_starsq = bet2_star * bet2_star + bet3_star * bet3_star;
		vbet = v2 * bet2_star + v3 * bet3_star;
		if ((cfsq - cssq) == T(0)) {
			alpha_f = T(1);
			alpha_s = T(0);
		} else if ((twid_asq - cssq) <= T(0)) {
			alpha_f = T(0);
			alpha_s = T(1);
		} else if ((cfsq - twid_asq) <= T(0)) {
			alpha_f = T(1);
			alpha_s = T(0);
		} else {
			alpha_f = A::sqrt(A::div(twid_asq - cssq, cfsq - cssq));
			alpha_s = A::sqrt(A::div(cfsq - twid_asq, cfsq - cssq));
		}
		sqrtd = A::sqrt(d);
		isqrtd = A::rcp(sqrtd);
		s = (b1 > 0) ? T(1) : T(-1);
		const T twid_a = A::sqrt(twid_asq);
		qf = cf * alpha_f * s;
		qs = cs * alpha_s * s;
		af_prime = twid_a * alpha_f * isqrtd;
		as_prime = twid_a * alpha_s * isqrtd;
		afpbb = af_prime * bt_star * bet_starsq;
		aspbb = as_prime * bt_star * bet_starsq;
		vax = A::sqrt(vaxsq);
		ev[0] = v1 - cf;
		ev[1] = v1 - vax;
		ev[2] = v1 - cs;
		ev[3] = v1;
		ev[4] = v1 + cs;
		ev[5] = v1 + vax;
		ev[6] = v1 + cf;

		T norm = A::div(T(0.5), twid_asq);
		cff = norm * alpha_f * cf;
		css = norm * alpha_s * cs;
		nqf = qf * norm;
		nqs = qs * norm;
		af = norm * af_prime * d;
		as = norm * as_prime * d;
		afpb = norm * af_prime * bt_star;
		aspb = norm * as_prime * bt_star;
		norm *= g1;
		nalpha_f = alpha_f * norm;
		nalpha_s = alpha_s * norm;
		q2_star = A::div(bet2_star, bet_starsq);
		q3_star = A::div(bet3_star, bet_starsq);
		vqstr = (v2 * q2_star + v3 * q3_star);
		norm2 = norm * T(2.0);
	}

	/* row n of the left eigenvector matrix (eq. B29) */
	__device__ __forceinline__ void left_row(int n, T r[7]) const
	{
		const T vh = vsq - hp;
		const T l10 = T(0.5) * (v2 * bet3 - v3 * bet2);
		const T l12 = T(-0.5) * bet3, l13 = T(0.5) * bet2;
		const T l15 = T(-0.5) * sqrtd * bet3 * s, l16 = T(0.5) * sqrtd * bet2 * s;
		const T l05 = as * q2_star - nalpha_f * b2, l06 = as * q3_star - nalpha_f * b3;
		const T l25 = -af * q2_star - nalpha_s * b2, l26 = -af * q3_star - nalpha_s * b3;
		switch (n) {
		case 0:
			r[0] = nalpha_f * vh + cff * (cf + v1) - nqs * vqstr - aspb;
			r[1] = -nalpha_f * v1 - cff;
			r[2] = -nalpha_f * v2 + nqs * q2_star;
			r[3] = -nalpha_f * v3 + nqs * q3_star;
			r[4] = nalpha_f;
			r[5] = l05;
			r[6] = l06;
			break;
		case 1:
			r[0] = l10; r[1] = T(0); r[2] = l12; r[3] = l13; r[4] = T(0); r[5] = l15; r[6] = l16;
			break;
		case 2:
			r[0] = nalpha_s * vh + css * (cs + v1) + nqf * vqstr + afpb;
			r[1] = -nalpha_s * v1 - css;
			r[2] = -nalpha_s * v2 - nqf * q2_star;
			r[3] = -nalpha_s * v3 - nqf * q3_star;
			r[4] = nalpha_s;
			r[5] = l25;
			r[6] = l26;
			break;
		case 3:
			r[0] = T(1.0) - norm2 * (T(0.5) * vsq - Arith<T>::div(g2 * x, g1));
			r[1] = norm2 * v1;
			r[2] = norm2 * v2;
			r[3] = norm2 * v3;
			r[4] = -norm2;
			r[5] = norm2 * b2;
			r[6] = norm2 * b3;
			break;
		case 4:
			r[0] = nalpha_s * vh + css * (cs - v1) - nqf * vqstr + afpb;
			r[1] = -nalpha_s * v1 + css;
			r[2] = -nalpha_s * v2 + nqf * q2_star;
			r[3] = -nalpha_s * v3 + nqf * q3_star;
			r[4] = nalpha_s;
			r[5] = l25;
			r[6] = l26;
			break;
		case 5:
			r[0] = -l10; r[1] = T(0); r[2] = -l12; r[3] = -l13; r[4] = T(0); r[5] = l15; r[6] = l16;
			break;
		default:
			r[0] = nalpha_f * vh + cff * (cf - v1) + nqs * vqstr - aspb;
			r[1] = -nalpha_f * v1 + cff;
			r[2] = -nalpha_f * v2 - nqs * q2_star;
			r[3] = -nalpha_f * v3 - nqs * q3_star;
			r[4] = nalpha_f;
			r[5] = l05;
			r[6] = l06;
			break;
		}
	}

	/* column n of the right eigenvector matrix (eq. B21) */
	__device__ __forceinline__ void right_col(int n, T c[7]) const
	{
		const T r41 = -(v2 * bet3 - v3 * bet2);
		switch (n) {
		case 0:
			c[0] = alpha_f;
			c[1] = alpha_f * ev[0];
			c[2] = alpha_f * v2 + qs * bet2_star;
			c[3] = alpha_f * v3 + qs * bet3_star;
			c[4] = alpha_f * (hp - v1 * cf) + qs * vbet + aspbb;
			c[5] = as_prime * bet2_star;
			c[6] = as_prime * bet3_star;
			break;
		case 1:
			c[0] = T(0);
			c[1] = T(0);
			c[2] = -bet3;
			c[3] = bet2;
			c[4] = r41;
			c[5] = -bet3 * s * isqrtd;
			c[6] = bet2 * s * isqrtd;
			break;
		case 2:
			c[0] = alpha_s;
			c[1] = alpha_s * ev[2];
			c[2] = alpha_s * v2 - qf * bet2_star;
			c[3] = alpha_s * v3 - qf * bet3_star;
			c[4] = alpha_s * (hp - v1 * cs) - qf * vbet - afpbb;
			c[5] = -af_prime * bet2_star;
			c[6] = -af_prime * bet3_star;
			break;
		case 3:
			c[0] = T(1.0);
			c[1] = v1;
			c[2] = v2;
			c[3] = v3;
			c[4] = T(0.5) * vsq + Arith<T>::div(g2 * x, g1);
			c[5] = T(0);
			c[6] = T(0);
			break;
		case 4:
			c[0] = alpha_s;
			c[1] = alpha_s * ev[4];
			c[2] = alpha_s * v2 + qf * bet2_star;
			c[3] = alpha_s * v3 + qf * bet3_star;
			c[4] = alpha_s * (hp + v1 * cs) + qf * vbet - afpbb;
			c[5] = -af_prime * bet2_star;
			c[6] = -af_prime * bet3_star;
			break;
		case 5:
			c[0] = T(0);
			c[1] = T(0);
			c[2] = bet3;
			c[3] = -bet2;
			c[4] = -r41;
			c[5] = -bet3 * s * isqrtd;
			c[6] = bet2 * s * isqrtd;
			break;
		default:
			c[0] = alpha_f;
			c[1] = alpha_f * ev[6];
			c[2] = alpha_f * v2 - qs * bet2_star;
			c[3] = alpha_f * v3 - qs * bet3_star;
			c[4] = alpha_f * (hp + v1 * cf) - qs * vbet + aspbb;
			c[5] = as_prime * bet2_star;
			c[6] = as_prime * bet3_star;
			break;
		}
	}
};

template <class T> __device__ __forceinline__ int roe_fluxes(const T Ul[7], const T Ur[7],
	const T Wl[7], const T Wr[7], T Bxi, T gamma, T flux[7])
{
	using A = Arith<T>;
	enum { D = 0, X = 1, Y = 2, Z = 3, E = 4, BY = 5, BZ = 6 };
	const T sqrtdl = A::sqrt(Wl[D]);
	const T sqrtdr = A::sqrt(Wr[D]);
	const T isdlpdr = A::rcp(sqrtdl + sqrtdr);
	const T droe = sqrtdl * sqrtdr;
	const T v1roe = (sqrtdl * Wl[X] + sqrtdr * Wr[X]) * isdlpdr;
	const T v2roe = (sqrtdl * Wl[Y] + sqrtdr * Wr[Y]) * isdlpdr;
	const T v3roe = (sqrtdl * Wl[Z] + sqrtdr * Wr[Z]) * isdlpdr;
	const T b2roe = (sqrtdr * Wl[BY] + sqrtdl * Wr[BY]) * isdlpdr;
	const T b3roe = (sqrtdr * Wl[BZ] + sqrtdl * Wr[BZ]) * isdlpdr;
	const T dby = Wl[BY] - Wr[BY], dbz = Wl[BZ] - Wr[BZ], sd = sqrtdl + sqrtdr;
	const T x = A::div(T(0.5) * (dby * dby + dbz * dbz), sd * sd);
	const T y = A::div(T(0.5) * (Wl[D] + Wr[D]), droe);
	const T pbl = T(0.5) * (Bxi * Bxi + Wl[BY] * Wl[BY] + Wl[BZ] * Wl[BZ]);
	const T pbr = T(0.5) * (Bxi * Bxi + Wr[BY] * Wr[BY] + Wr[BZ] * Wr[BZ]);
	const T hroe = (A::div(Ul[E] + Wl[E] + pbl, sqrtdl) + A::div(Ur[E] + Wr[E] + pbr, sqrtdr)) * isdlpdr;

	RoeEsys<T> es;
	es.init(droe, v1roe, v2roe, v3roe, hroe, Bxi, b2roe, b3roe, x, y, gamma);

	T Fl[7], Fr[7];
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
	Fl[X] -= T(0.5) * (Bxi * Bxi - Wl[BY] * Wl[BY] - Wl[BZ] * Wl[BZ]);
	Fr[X] -= T(0.5) * (Bxi * Bxi - Wr[BY] * Wr[BY] - Wr[BZ] * Wr[BZ]);
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

	if (es.ev[0] >= T(0)) {
		for (int q = 0; q < 7; q++) flux[q] = Fl[q];
		return ERR_OK;
	}
	if (es.ev[6] <= T(0)) {
		for (int q = 0; q < 7; q++) flux[q] = Fr[q];
		return ERR_OK;
	}

	T dU[7], a[7];
	for (int m = 0; m < 7; m++) dU[m] = Ur[m] - Ul[m];
	for (int n = 0; n < 7; n++) {
		T row[7];
		es.left_row(n, row);
		T acc = T(0);
		for (int m = 0; m < 7; m++) acc += row[m] * dU[m];
		a[n] = acc;
	}

	T u[7];
	for (int m = 0; m < 7; m++) u[m] = Ul[m];
	for (int n = 0; n < 6; n++) {
		T col[7];
		es.right_col(n, col);
		for (int m = 0; m < 7; m++) u[m] += a[n] * col[m];
		if (es.ev[n + 1] > es.ev[n]) {
			if (u[0] <= T(0)) return ERR_ROE_DENSITY;
			T p = u[4] - A::div(T(0.5) * (u[1] * u[1] + u[2] * u[2] + u[3] * u[3]), u[0]);
			p -= T(0.5) * (u[5] * u[5] + u[6] * u[6] + Bxi * Bxi);
			if (p < T(0)) return ERR_ROE_PRESSURE;
		}
	}

	for (int q = 0; q < 7; q++) flux[q] = T(0.5) * (Fl[q] + Fr[q]);
	for (int m = 0; m < 7; m++) {
		const T coeff = T(0.5) * a[m] * smax(fabs(es.ev[m]), T(0));
		T col[7];
		es.right_col(m, col);
		for (int q = 0; q < 7; q++) flux[q] -= coeff * col[q];
	}
	return ERR_OK;
}

/* source/mhd/roe_athena.hpp:567-818 (ignores the background field, :584) */
template <class T> __device__ __forceinline__ int flux_roe(const double Ld[8],
	const double Rd[8], const T* /*bg*/, const Phys& ph, T F[8], T& s_out)
{
	using A = Arith<T>;
	const T gamma = (T)ph.gamma, mu0 = (T)ph.mu0;
	const T zero[3] = {0, 0, 0};
	if (Ld[RHO] <= 0) return ERR_DENSITY;
	if (Rd[RHO] <= 0) return ERR_DENSITY;
	Side<T> L, R;
	make_side<T>(Ld, ph, L);
	make_side<T>(Rd, ph, R);
	if (L.p64 <= 0) return ERR_PRESSURE;
	if (R.p64 <= 0) return ERR_PRESSURE;

	const T sq = (T)ph.sqrt_mu0;
	const T Bxi = A::div(L.u[B0] + R.u[B0], T(2.0) * sq);
	const T Ul[7] = {L.u[RHO], L.u[M0], L.u[M1], L.u[M2], L.u[NRJ],
		A::div(L.u[B1], sq), A::div(L.u[B2], sq)};
	const T Ur[7] = {R.u[RHO], R.u[M0], R.u[M1], R.u[M2], R.u[NRJ],
		A::div(R.u[B1], sq), A::div(R.u[B2], sq)};
	/* pressure with vacuum permeability 1 on the unscaled field
	   (roe_athena.hpp:705-721), recovered in FP64 */
	const T pl1 = (T)pressure64(Ld, ph.gm1, 0.5);
	const T pr1 = (T)pressure64(Rd, ph.gm1, 0.5);
	const T Wl[7] = {L.u[RHO], L.v[0], L.v[1], L.v[2], pl1, Ul[5], Ul[6]};
	const T Wr[7] = {R.u[RHO], R.v[0], R.v[1], R.v[2], pr1, Ur[5], Ur[6]};

	T f7[7];
	int e = roe_fluxes<T>(Ul, Ur, Wl, Wr, Bxi, gamma, f7);
	if (e) return e;
	F[RHO] = f7[0];
	F[M0] = f7[1];
	F[M1] = f7[2];
	F[M2] = f7[3];
	F[NRJ] = f7[4];
	F[B0] = 0;
	F[B1] = f7[5] * sq;
	F[B2] = f7[6] * sq;

	T cfL, cfR;
	e = fast_speed<T>(L, Ld, zero, gamma, mu0, cfL);
	if (e) return e;
	e = fast_speed<T>(R, Rd, zero, gamma, mu0, cfR);
	if (e) return e;
	const T ms = smax(cfL, cfR);
	const T
		msn = (L.v[0] <= R.v[0]) ? L.v[0] - ms : R.v[0] - ms,
		msp = (L.v[0] <= R.v[0]) ? R.v[0] + ms : L.v[0] + ms;
	if (msn >= T(0)) {
		s_out = fabs(msp);
	} else if (msp <= T(0)) {
		s_out = fabs(msn);
	} else {
		s_out = smax(fabs(msn), fabs(msp));
	}
	return ERR_OK;
}

template <int SOLVER, class T> __device__ __forceinline__ int riemann(const double L[8],
	const double R[8], const T bg[3], const Phys& ph, T F[8], T& s)
{
	if constexpr (SOLVER == RUSANOV) return flux_rusanov<T>(L, R, bg, ph, F, s);
	else if constexpr (SOLVER == HLL_ATHENA) return flux_hll<T>(L, R, bg, ph, F, s);
	else if constexpr (SOLVER == HLLD_ATHENA) return flux_hlld<T>(L, R, bg, ph, F, s);
	else if constexpr (SOLVER == ROE_ATHENA) return flux_roe<T>(L, R, bg, ph, F, s);
	else return ERR_SOLVER;
}

/* a solver in flux precision T with FP64 results: T = float gives the FP32
   forms, except HLLD, whose mixed form (flux_hlld_mixed) returns FP64
   physical fluxes plus FP32 fan corrections */
template <int SOLVER, class T> __device__ __forceinline__ int riemann_d(const double L[8],
	const double R[8], const T bg[3], const Phys& ph, double F[8], double& s)
{
	if constexpr (SOLVER == HLLD_ATHENA && std::is_same_v<T, float>) {
		return flux_hlld_mixed(L, R, ph, F, s);
	} else {
		T Ft[8], st = 0;
		const int e = riemann<SOLVER, T>(L, R, bg, ph, Ft, st);
		for (int q = 0; q < 8; q++) F[q] = (double)Ft[q];
		s = (double)st;
		return e;
	}
}

inline Phys make_phys(double gamma, double mu0)
{
	Phys p;
	p.gamma = gamma;
	p.mu0 = mu0;
	p.gm1 = gamma - 1;
	p.inv_mu0 = 1.0 / mu0;
	p.half_inv_mu0 = 0.5 / mu0;
	p.sqrt_mu0 = std::sqrt(mu0);
	p.inv_sqrt_mu0 = 1.0 / std::sqrt(mu0);
	p.two_mu0 = 2 * mu0;
	return p;
}

} // namespace pamhd_dev
