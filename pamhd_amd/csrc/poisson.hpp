/*
Geometric multigrid solver of the periodic 7-point Poisson problem
L phi = f on the device, used by divergence cleaning (the counterpart of
the Poisson solve inside pamhd::divergence::remove,
source/divergence/remove.hpp:609-769).

Levels halve every even dimension >= 4 whose spacing is within 1.5x of the
level's finest active spacing (semi-coarsening), until nothing can be halved
or a level has <= 64 cells.  Smoothing is red-black Gauss-Seidel when every
active dimension of the level is even, damped Jacobi (omega 0.8) otherwise;
restriction averages the children, prolongation is trilinear in the
coarsened dimensions; V(2,2) cycles with 60 sweeps on the coarsest level.
Every kernel computes each value in a fixed operation order, so results do
not depend on the launch shape; the residual norms of the stopping test are
deterministic one-pass reductions (block moments merged pairwise).
*/
#ifndef PAMHD_POISSON_HPP
#define PAMHD_POISSON_HPP

#include <hip/hip_runtime.h>

#include <vector>

namespace pamhd_dev {

/* index of cell (i, j, k) (in range) of a level: x fastest, or, on a level
   smoothed red-black with an even x size (packed), the two colours
   ((i + j + k) & 1) as separate halves, each x fastest over the colour's
   cells, so that a colour pass reads and writes whole cache lines (the 7
   neighbours of a cell are of the other colour) */
__host__ __device__ inline int mg_index(int packed, int n0, int n1, int half_n, int i, int j, int k)
{
	if (!packed) return i + n0 * (j + n1 * k);
	return ((i + j + k) & 1) * half_n + (i >> 1) + (n0 >> 1) * (j + n1 * k);
}

struct MgLevel {
	int n[3];
	int packed = 0;  /* colour-split layout (mg_index) */
	double h[3];
	double hh[3];    /* h * h */
	double diag;     /* sum over active dimensions of 2 / (h * h) */
	int ncell;
	double* phi = nullptr;
	double* f = nullptr;
	double* r = nullptr;
};

struct PoissonMG {
	std::vector<MgLevel> lv;
	double* partial = nullptr;      /* per-block partial sums */
	double* norm_part = nullptr;    /* per-block residual moments of the stopping test */
	size_t norm_part_len = 0;       /* doubles */
	size_t partial_len = 0;
	double* result = nullptr;       /* device: 2 doubles */
	double* host_result = nullptr;  /* pinned: 2 doubles */
	int fused_smoother = 0;         /* 1: LDS-tiled fused red-black sweeps where possible */
	int tail = -1;                  /* first of the bottom levels cycled in one block, or -1 */
};

/* allocates the level hierarchy for a dims[0] x dims[1] x dims[2] grid */
hipError_t mg_create(PoissonMG& mg, const int dims[3], const double h[3]);
void mg_destroy(PoissonMG& mg);

/*
Solves L phi = f with f already in lv[0].f (phi starts from 0) until
||residual - mean||_2 <= max(tol ||f - mean||_2, atol) or max_cycles
V-cycles.  *converged tells which; *rel gets ||r - mean|| / ||f - mean||.
*/
hipError_t mg_solve(PoissonMG& mg, hipStream_t st, double tol, double atol, int max_cycles,
	double* rel, int* cycles, bool* converged);

/* deterministic sum over n values: mode 0 sums a, mode 1 sums (a - shift)^2,
   mode 2 sums |a|;
   synchronises the stream and returns the value in *out */
hipError_t mg_sum(PoissonMG& mg, hipStream_t st, const double* a, int n, int mode,
	double shift, double* out);

/* out[0], out[1] = the sums of the even and odd entries of the nb pairs
   (device memory, e.g. per-block partial sums in mg.norm_part); synchronises
   the stream */
hipError_t mg_sum_pairs(PoissonMG& mg, hipStream_t st, const double* pairs, int nb, double out[2]);
/* launches one V-cycle on the hierarchy (exposed for timing) */
void mg_vcycle(PoissonMG& mg, hipStream_t st, int level);

/* one smoothing sweep on a level (exposed for timing and tests) */
void mg_smooth(PoissonMG& mg, hipStream_t st, int level, int sweeps);

} // namespace pamhd_dev

#endif
