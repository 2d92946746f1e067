// spec_sum.hip -- left-to-right FP64 sums, bitwise, computed in parallel.
//
// The reference adds its norm and finalisation sums one value after another
// (Norm, src/hash_align.c:153-154 and 158-159; src/EM.c:362-363).  Every
// partial sum is rounded, so only that exact sequence of additions gives the
// same bits, and one device lane adding 200,000 dependent values takes
// milliseconds.  These kernels produce the same bits in tens of
// microseconds.
//
// Split a chain of n values into chunks of L.  For chunk j the map from
// its start value to its end value,
//     F_j(s) = fl(...fl(fl(s + y_a) + y_a+1)... + y_b),
// is monotone in s and cheap to evaluate, but the true start s_j is only
// known once chunks 0..j-1 are done.  So:
//   1. chunk sums c_j (any order; they only steer guesses);
//   2. g0_j = sum_{i<j} c_i; the drift d_j = F_j(g0_j) - g0_j is run once
//      per chunk, all chunks at once.  It captures the chunk's rounding
//      (small terms absorbed by a large running sum) at the right
//      magnitude, which plain chunk sums miss by up to ~10^4 ulps;
//   3. g1_j = sum_{i<j} d_i; every chunk runs F_j from the W consecutive
//      doubles centred on g1_j (one candidate per lane), all at once;
//   4. one workgroup per chain walks the chunks in order from s_0 = +0:
//      the exact s_j is looked up among chunk j's candidates by its
//      ordered bit pattern, which yields F_j(s_j) bitwise; a start outside
//      the window (a miss) reruns the chunk from s_j, so the result never
//      depends on how good the guess was.
// Steps 2-4 are pure FP64 additions in the reference's order on values
// already rounded as the reference rounds them; step 1's and the guesses'
// rounding only decides hits and misses.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hem_dev.h"

#define SBLOCK 256
#define WAVE 64
#define SPEC_W HEMK_SPEC_W
#define SPEC_PG 32                          // chunks per LDS page
#define SPEC_PRE (SPEC_PG * SPEC_W / SBLOCK)  // page values per thread

static_assert(SPEC_W == WAVE, "one candidate per lane of one wave");
static_assert(SPEC_PG <= SBLOCK, "page bases are loaded one per thread");

namespace {

// Ordered integer view of a double: consecutive keys are consecutive
// doubles (-0.0 and +0.0 are neighbours), and the map is a bijection on
// all 2^64 bit patterns, so unkey(key(x)) has x's bits.
__device__ __forceinline__ long long dkey(double x)
{
	const long long b = __double_as_longlong(x);
	return b >= 0 ? b : -(b & 0x7FFFFFFFFFFFFFFFLL) - 1;
}

__device__ __forceinline__ double dunkey(long long k)
{
	const long long b =
	    k >= 0 ? k : ((-(k + 1)) | (long long)0x8000000000000000ULL);
	return __longlong_as_double(b);
}

// Chain slot q of a launch -> the row it sums (rows of x are ld apart).
// Pairs: slots 2i, 2i+1 are rows i and R + i of replicate list entry i.
__device__ __forceinline__ bool chain_row(const hemk_spec &sp, uint32_t q,
                                          uint32_t &row)
{
	uint32_t rep = sp.pairs ? q >> 1 : q;
	if (sp.alist) {
		if (rep >= sp.alist[0])
			return false;
		rep = sp.alist[1 + rep];
	} else if (q >= sp.nchain) {
		return false;
	}
	row = sp.pairs ? (q & 1) * sp.R + rep : rep;
	return true;
}

// The value the chain adds at position i of its row: x, (x - m)^2 with
// m = mres[row] / n (the mean of src/hash_align.c:156), or x * x.
__device__ __forceinline__ double chain_val(int xf, double x, double m)
{
#pragma clang fp contract(off)
	if (xf == HEMK_SPEC_CENTRED)
		return (x - m) * (x - m);
	if (xf == HEMK_SPEC_SQUARE)
		return x * x;
	return x;
}

__device__ __forceinline__ double chain_mean(const hemk_spec &sp,
                                             uint32_t row)
{
	return sp.xf == HEMK_SPEC_CENTRED ? sp.mres[row] / (double)sp.n : 0.0;
}

// acc + x[0] + ... + x[cnt-1] left to right (per lane; x in LDS, read
// eight ahead of the additions).
__device__ __forceinline__ double run_chain(double acc, const double *x,
                                            uint32_t cnt)
{
#pragma clang fp contract(off)
	uint32_t j = 0;
	if (cnt >= 8) {
		double a[8];
#pragma unroll
		for (int q = 0; q < 8; ++q)
			a[q] = x[q];
		for (j = 8; j + 8 <= cnt; j += 8) {
			double b[8];
#pragma unroll
			for (int q = 0; q < 8; ++q)
				b[q] = x[j + q];
#pragma unroll
			for (int q = 0; q < 8; ++q)
				acc += a[q];
#pragma unroll
			for (int q = 0; q < 8; ++q)
				a[q] = b[q];
		}
#pragma unroll
		for (int q = 0; q < 8; ++q)
			acc += a[q];
	}
	for (; j < cnt; ++j)
		acc += x[j];
	return acc;
}

// Sum of the first j values of v[] (guesses only), fixed order, the same
// bits in every lane of the wave.
__device__ __forceinline__ double wave_prefix(const double *v, uint32_t j)
{
	const int lane = threadIdx.x % WAVE;
	double g = 0.0;
	for (uint32_t i = lane; i < j; i += WAVE)
		g += v[i];
#pragma unroll
	for (int m = WAVE / 2; m >= 1; m >>= 1)
		g += __shfl_xor(g, m, WAVE);
	return g;
}

// Stage chunk j of a row (its chain values) into LDS.
__device__ __forceinline__ uint32_t stage_chunk(const hemk_spec &sp,
                                                uint32_t row, uint32_t j,
                                                double m, double *ys,
                                                int nthreads)
{
	const uint32_t beg = j * sp.L;
	const uint32_t cnt = min(sp.L, sp.n - beg);
	const double *x = sp.x + (size_t)row * sp.ld + beg;
	for (uint32_t i = threadIdx.x; i < cnt; i += nthreads)
		ys[i] = chain_val(sp.xf, x[i], m);
	return cnt;
}

__device__ __forceinline__ bool skip(const hemk_spec &sp,
                                     const hemk_ctl *ctl)
{
	return ctl && ctl->done;
}

// Step 1: chunk sums.  Grid (P, chains), SBLOCK threads.
__global__ __launch_bounds__(SBLOCK) void k_spec_csum(hemk_spec sp,
                                                      const hemk_ctl *ctl)
{
#pragma clang fp contract(off)
	__shared__ double red[SBLOCK / WAVE];
	uint32_t row;
	if (skip(sp, ctl) || !chain_row(sp, blockIdx.y, row))
		return;
	const uint32_t j = blockIdx.x;
	const uint32_t beg = j * sp.L;
	const uint32_t end = min(sp.n, beg + sp.L);
	const double m = chain_mean(sp, row);
	const double *x = sp.x + (size_t)row * sp.ld;
	double a = 0.0;
	for (uint32_t i = beg + threadIdx.x; i < end; i += SBLOCK)
		a += chain_val(sp.xf, x[i], m);
#pragma unroll
	for (int k = WAVE / 2; k >= 1; k >>= 1)
		a += __shfl_xor(a, k, WAVE);
	if (threadIdx.x % WAVE == 0)
		red[threadIdx.x / WAVE] = a;
	__syncthreads();
	if (threadIdx.x == 0) {
		double t = red[0];
#pragma unroll
		for (int w = 1; w < SBLOCK / WAVE; ++w)
			t += red[w];
		sp.csum[(size_t)blockIdx.y * sp.P + j] = t;
	}
}

// Plain solve, pass 1: r = u1 - u and v = (u2 - u1) - r (src/EM.c:333-336)
// into rows 0 and 1 of x, and the chunk sums of their chain values (IDENT
// or SQUARE).  Grid (P, 2).
__global__ __launch_bounds__(SBLOCK) void k_spec_rv(
    hemk_spec sp, const double *__restrict__ u, const double *__restrict__ u1,
    const double *__restrict__ u2, double *__restrict__ xw,
    const hemk_ctl *ctl)
{
#pragma clang fp contract(off)
	__shared__ double red[SBLOCK / WAVE];
	if (skip(sp, ctl))
		return;
	const uint32_t q = blockIdx.y;
	const uint32_t j = blockIdx.x;
	const uint32_t beg = j * sp.L;
	const uint32_t end = min(sp.n, beg + sp.L);
	double a = 0.0;
	for (uint32_t i = beg + threadIdx.x; i < end; i += SBLOCK) {
		const double r = u1[i] - u[i];
		const double val = q == 0 ? r : (u2[i] - u1[i]) - r;
		xw[(size_t)q * sp.ld + i] = val;
		a += chain_val(sp.xf, val, 0.0);
	}
#pragma unroll
	for (int k = WAVE / 2; k >= 1; k >>= 1)
		a += __shfl_xor(a, k, WAVE);
	if (threadIdx.x % WAVE == 0)
		red[threadIdx.x / WAVE] = a;
	__syncthreads();
	if (threadIdx.x == 0) {
		double t = red[0];
#pragma unroll
		for (int w = 1; w < SBLOCK / WAVE; ++w)
			t += red[w];
		sp.csum[(size_t)q * sp.P + j] = t;
	}
}

// Step 2: drift of each chunk from the chunk-sum guess.  Grid (P, chains),
// one wave.
__global__ __launch_bounds__(WAVE) void k_spec_drift(hemk_spec sp,
                                                     const hemk_ctl *ctl)
{
#pragma clang fp contract(off)
	__shared__ double ys[HEMK_SPEC_LMAX];
	uint32_t row;
	if (skip(sp, ctl) || !chain_row(sp, blockIdx.y, row))
		return;
	const uint32_t j = blockIdx.x;
	const size_t cb = (size_t)blockIdx.y * sp.P;
	const double g0 = wave_prefix(sp.csum + cb, j);
	const uint32_t cnt =
	    stage_chunk(sp, row, j, chain_mean(sp, row), ys, WAVE);
	__syncthreads();
	const double f = run_chain(g0, ys, cnt);
	if (threadIdx.x == 0)
		sp.drift[cb + j] = f - g0;
}

// Step 3: every chunk from SPEC_W consecutive starts around the drift
// guess.  Grid (P, chains), one wave; lane c starts at key(g1) - W/2 + c.
__global__ __launch_bounds__(WAVE) void k_spec_window(hemk_spec sp,
                                                      const hemk_ctl *ctl)
{
#pragma clang fp contract(off)
	__shared__ double ys[HEMK_SPEC_LMAX];
	uint32_t row;
	if (skip(sp, ctl) || !chain_row(sp, blockIdx.y, row))
		return;
	const uint32_t j = blockIdx.x;
	const size_t cb = (size_t)blockIdx.y * sp.P;
	const double g1 = wave_prefix(sp.drift + cb, j);
	// wrapping arithmetic: a window that straddles the ends of the key
	// range only produces misses (the lookup checks for overflow)
	const long long base =
	    (long long)((unsigned long long)dkey(g1) - SPEC_W / 2);
	const uint32_t cnt =
	    stage_chunk(sp, row, j, chain_mean(sp, row), ys, WAVE);
	__syncthreads();
	const double s = run_chain(
	    dunkey((long long)((unsigned long long)base + threadIdx.x)), ys,
	    cnt);
	sp.tab[(cb + j) * SPEC_W + threadIdx.x] = s;
	if (threadIdx.x == 0)
		sp.base[cb + j] = base;
}

// Step 4: walk the chunks in order.  One workgroup per chain; every thread
// holds the same running value, so every branch is uniform.  Candidate
// tables stream through two LDS pages (the next page is in registers
// while the current one is resolved).
__global__ __launch_bounds__(SBLOCK) void k_spec_resolve(hemk_spec sp,
                                                         hemk_ctl *ctl)
{
#pragma clang fp contract(off)
	__shared__ double tb[2][SPEC_PG * SPEC_W];
	__shared__ long long bs[2][SPEC_PG];
	__shared__ double ys[HEMK_SPEC_LMAX];
	uint32_t row;
	if (skip(sp, ctl) || !chain_row(sp, blockIdx.x, row))
		return;
	const size_t cb = (size_t)blockIdx.x * sp.P;
	const double *tab = sp.tab + cb * SPEC_W;
	const long long *bse = sp.base + cb;
	const double m = chain_mean(sp, row);
	const uint32_t npg = (sp.P + SPEC_PG - 1) / SPEC_PG;
	const int tid = threadIdx.x;
	double pre[SPEC_PRE];
	long long preb = 0;

	// page 0
	{
		const uint32_t lim = min(sp.P, (uint32_t)SPEC_PG) * SPEC_W;
#pragma unroll
		for (int i = 0; i < SPEC_PRE; ++i) {
			const uint32_t f = tid + SBLOCK * i;
			pre[i] = f < lim ? tab[f] : 0.0;
		}
		if (tid < SPEC_PG && (uint32_t)tid < sp.P)
			preb = bse[tid];
#pragma unroll
		for (int i = 0; i < SPEC_PRE; ++i)
			tb[0][tid + SBLOCK * i] = pre[i];
		if (tid < SPEC_PG)
			bs[0][tid] = preb;
		__syncthreads();
	}

	double s = 0.0;
	uint32_t misses = 0;
	for (uint32_t pg = 0; pg < npg; ++pg) {
		const int buf = pg & 1;
		const uint32_t j0 = pg * SPEC_PG;
		const uint32_t j1 = min(sp.P, j0 + SPEC_PG);
		const bool more = pg + 1 < npg;
		if (more) {
			const uint32_t n0 = j0 + SPEC_PG;
			const uint32_t lim =
			    (min(sp.P, n0 + SPEC_PG) - n0) * SPEC_W;
			const double *t = tab + (size_t)n0 * SPEC_W;
#pragma unroll
			for (int i = 0; i < SPEC_PRE; ++i) {
				const uint32_t f = tid + SBLOCK * i;
				pre[i] = f < lim ? t[f] : 0.0;
			}
			if (tid < SPEC_PG && n0 + tid < sp.P)
				preb = bse[n0 + tid];
		}
		for (uint32_t j = j0; j < j1; ++j) {
			long long idx;
			if (!__builtin_sub_overflow(dkey(s), bs[buf][j - j0],
			                            &idx) &&
			    idx >= 0 && idx < SPEC_W) {
				s = tb[buf][(j - j0) * SPEC_W + idx];
				continue;
			}
			// the exact start lies outside the window: add the
			// chunk from it
			const uint32_t cnt = stage_chunk(sp, row, j, m, ys, SBLOCK);
			__syncthreads();
			s = run_chain(s, ys, cnt);
			__syncthreads();
			++misses;
		}
		if (more) {
#pragma unroll
			for (int i = 0; i < SPEC_PRE; ++i)
				tb[buf ^ 1][tid + SBLOCK * i] = pre[i];
			if (tid < SPEC_PG)
				bs[buf ^ 1][tid] = preb;
		}
		__syncthreads();
	}
	if (tid == 0) {
		sp.res[row] = s;
		if (misses && sp.miss)
			atomicAdd(sp.miss, misses);
	}
}

} // namespace

extern "C" {

uint32_t hemk_spec_chunk(uint32_t n)
{
	// ~2 dependent chunk runs of L adds against ~n/L dependent lookups
	uint32_t L = 256;
	while (L < HEMK_SPEC_LMAX && (uint64_t)L * L < (uint64_t)n * 4)
		L *= 2;
	return L;
}

int hemk_spec_sum(const hemk_spec *sp, int produce, const double *u,
                  const double *u1, const double *u2, hemk_ctl *ctl,
                  void *s)
{
	hipStream_t st = (hipStream_t)s;
	if (sp->n == 0 || sp->nchain == 0 || sp->L == 0 ||
	    sp->L > HEMK_SPEC_LMAX || sp->L % 8 != 0 ||
	    sp->P != (sp->n + sp->L - 1) / sp->L ||
	    (produce && (sp->nchain != 2 || sp->pairs || sp->alist)))
		return (int)hipErrorInvalidValue;
	dim3 grid(sp->P, sp->nchain);
	if (produce)
		k_spec_rv<<<grid, SBLOCK, 0, st>>>(*sp, u, u1, u2,
		                                   (double *)sp->x, ctl);
	else
		k_spec_csum<<<grid, SBLOCK, 0, st>>>(*sp, ctl);
	k_spec_drift<<<grid, WAVE, 0, st>>>(*sp, ctl);
	k_spec_window<<<grid, WAVE, 0, st>>>(*sp, ctl);
	k_spec_resolve<<<sp->nchain, SBLOCK, 0, st>>>(*sp, ctl);
	return (int)hipGetLastError();
}

} // extern "C"
