/*
Harness that compiles the reference's own Rusanov and HLL flux headers
(source/mhd/rusanov.hpp, source/mhd/hll_athena.hpp and what they include:
source/mhd/common.hpp, source/common_functions.hpp,
source/common_variables.hpp) into a shared library, so the oracle's
restatement can be compared with the reference's compiled arithmetic.

TEST INFRASTRUCTURE ONLY.  The reference sources are read in place from
the reference checkout at build time and never copied; the library is
written to oracle/_ref/.  HLLD and Roe are not built: their headers include
gensimcell.hpp and (Roe) mhd/variables.hpp -> grid/amr.hpp -> muparserx and
rapidjson, none of which exist here.

The caller side (state container and variable tags) is this file's own:
the solvers are templates over any container with operator[](Tag)
returning Tag::data_type& (source/mhd/variables.hpp:240-246).
*/
#include <array>
#include <functional>
#include <iostream>
#include <stdexcept>
#include <tuple>

#include "mhd/rusanov.hpp"
#include "mhd/hll_athena.hpp"

namespace {

struct Mas_T { using data_type = double; };
struct Mom_T { using data_type = std::array<double, 3>; };
struct Nrj_T { using data_type = double; };
struct Mag_T { using data_type = std::array<double, 3>; };

struct State {
	double rho = 0;
	std::array<double, 3> mom{0, 0, 0};
	double nrj = 0;
	std::array<double, 3> mag{0, 0, 0};
	double& operator[](const Mas_T&) { return rho; }
	std::array<double, 3>& operator[](const Mom_T&) { return mom; }
	double& operator[](const Nrj_T&) { return nrj; }
	std::array<double, 3>& operator[](const Mag_T&) { return mag; }
	const double& operator[](const Mas_T&) const { return rho; }
	const std::array<double, 3>& operator[](const Mom_T&) const { return mom; }
	const double& operator[](const Nrj_T&) const { return nrj; }
	const std::array<double, 3>& operator[](const Mag_T&) const { return mag; }
};

State load(const double* u)
{
	State s;
	s.rho = u[0];
	s.mom = {u[1], u[2], u[3]};
	s.nrj = u[4];
	s.mag = {u[5], u[6], u[7]};
	return s;
}

void store(const State& s, double* f)
{
	f[0] = s.rho;
	f[1] = s.mom[0];
	f[2] = s.mom[1];
	f[3] = s.mom[2];
	f[4] = s.nrj;
	f[5] = s.mag[0];
	f[6] = s.mag[1];
	f[7] = s.mag[2];
}

template <class Fn> int call(Fn fn, const double* L, const double* R, const double* bg,
	double* F, double* smax)
{
	State l = load(L), r = load(R);
	const std::array<double, 3> b{bg[0], bg[1], bg[2]};
	try {
		State flux;
		double s = 0;
		std::tie(flux, s) = fn(l, r, b);
		store(flux, F);
		*smax = s;
		return 0;
	} catch (const std::domain_error&) {
		return 1;
	}
}

} // namespace

extern "C" int ref_flux_rusanov(const double* L, const double* R, const double* bg,
	double gamma, double mu0, double* F, double* smax)
{
	return call([&](State& l, State& r, const std::array<double, 3>& b) {
		return pamhd::mhd::get_flux_rusanov<Mas_T, Mom_T, Nrj_T, Mag_T>(l, r, b, gamma, mu0);
	}, L, R, bg, F, smax);
}

extern "C" int ref_flux_hll(const double* L, const double* R, const double* bg,
	double gamma, double mu0, double* F, double* smax)
{
	return call([&](State& l, State& r, const std::array<double, 3>& b) {
		return pamhd::mhd::athena::get_flux_hll<Mas_T, Mom_T, Nrj_T, Mag_T>(l, r, b, gamma, mu0);
	}, L, R, bg, F, smax);
}

/* A batch of faces in the layout of or_face_flux_batch (component-major
   L, R (8 * n), bg (3 * n) or NULL, F (8 * n), smax (n)), for timing the
   reference's compiled flux against the restatement's on the same faces:
   which = 0 Rusanov, 2 HLL (the solver numbering of
   source/mhd/common.hpp:54-61).  Returns the number of faces whose solver
   threw (their flux and speed are zero). */
extern "C" long long ref_flux_batch(int which, size_t n, const double* L, const double* R,
	const double* bg, double gamma, double mu0, double* F, double* smax)
{
	long long bad = 0;
	for (size_t f = 0; f < n; f++) {
		double l[8], r[8], b[3] = {0, 0, 0}, out[8] = {0, 0, 0, 0, 0, 0, 0, 0}, s = 0;
		for (int q = 0; q < 8; q++) {
			l[q] = L[q * n + f];
			r[q] = R[q * n + f];
		}
		if (bg) for (int q = 0; q < 3; q++) b[q] = bg[q * n + f];
		const int e = which == 0 ? ref_flux_rusanov(l, r, b, gamma, mu0, out, &s)
			: ref_flux_hll(l, r, b, gamma, mu0, out, &s);
		if (e) {
			bad++;
			for (int q = 0; q < 8; q++) out[q] = 0;
			s = 0;
		}
		for (int q = 0; q < 8; q++) F[q * n + f] = out[q];
		smax[f] = s;
	}
	return bad;
}
