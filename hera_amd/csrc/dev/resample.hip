// resample.hip -- bootstrap resampling of fragment counts (gfx950).
//
// A replicate redraws all N fragments with replacement: the count vector
// over the K categories (unique reads of each transcript, then each class)
// is Multinomial(N, w / N).  Drawn exactly, in two levels: N draws over
// weight-balanced chunks of categories (level 1), then, inside each chunk,
// its m_j draws over the chunk's categories (level 2; a chunk holding a
// single category takes all m_j without drawing).  Every draw is a
// Philox-4x32-10 word turned into an exact uniform integer by Lemire's
// multiply-and-reject method, and histograms are integer sums, so the counts
// depend only on (seed, weights) -- not on the launch shape or timing.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../host/philox.h"
#include "hem_dev.h"

#define RBLOCK 256

namespace {

// Uniform integer in [0, M) (M >= 1) for draw (i, j) of a replicate.
__device__ __forceinline__ uint32_t draw_bounded(uint64_t key, uint32_t i,
                                                 uint32_t j, uint32_t M)
{
	const uint32_t thresh = (0u - M) % M;  // 2^32 mod M
	for (uint32_t att = 0;; ++att) {
		const hem_u32x4 h = hem_philox(key, i, j, att, 0x62737472u);
		for (int w = 0; w < 4; ++w) {
			const uint64_t m = (uint64_t)h.v[w] * M;
			if ((uint32_t)m >= thresh)
				return (uint32_t)(m >> 32);
		}
	}
}

// Index of the last entry <= y in a non-decreasing array a[0..n].
__device__ __forceinline__ uint32_t upper_index(const uint32_t *a, uint32_t n,
                                                uint32_t y)
{
	uint32_t lo = 0, hi = n;  // a[lo] <= y < a[hi]
	while (hi - lo > 1) {
		const uint32_t mid = (lo + hi) >> 1;
		if (a[mid] <= y)
			lo = mid;
		else
			hi = mid;
	}
	return lo;
}

__global__ __launch_bounds__(RBLOCK) void k_fill_u32(uint32_t *p, size_t n,
                                                     uint32_t v)
{
	for (size_t i = (size_t)blockIdx.x * RBLOCK + threadIdx.x; i < n;
	     i += (size_t)gridDim.x * RBLOCK)
		p[i] = v;
}

// Level 1: block (segment s, replicate r) draws fragments
// [s*seg, min(N, (s+1)*seg)) over the B chunks; m[r*B + j] += counts.
__global__ __launch_bounds__(RBLOCK) void k_resample_l1(
    const uint32_t *__restrict__ chunk_cw, uint32_t B, uint32_t N,
    uint32_t seg, const uint64_t *__restrict__ seeds, uint32_t *__restrict__ m)
{
	extern __shared__ uint32_t sh[];
	uint32_t *cw = sh;          // B + 1
	uint32_t *hist = sh + B + 1; // B
	const uint32_t r = blockIdx.y;
	const uint64_t key = seeds[r];
	for (uint32_t i = threadIdx.x; i <= B; i += RBLOCK)
		cw[i] = chunk_cw[i];
	for (uint32_t i = threadIdx.x; i < B; i += RBLOCK)
		hist[i] = 0;
	__syncthreads();
	const uint64_t i0 = (uint64_t)blockIdx.x * seg;
	const uint64_t i1 = i0 + seg < N ? i0 + seg : N;
	for (uint64_t i = i0 + threadIdx.x; i < i1; i += RBLOCK) {
		const uint32_t y = draw_bounded(key, (uint32_t)i, 0xFFFFFFFFu, N);
		atomicAdd(&hist[upper_index(cw, B, y)], 1u);
	}
	__syncthreads();
	for (uint32_t i = threadIdx.x; i < B; i += RBLOCK)
		if (hist[i])
			atomicAdd(&m[(size_t)r * B + i], hist[i]);
}

// Level 2: block (chunk j, replicate r) distributes m_j over the chunk's
// categories and writes out[r*K + k] for its categories.
__global__ __launch_bounds__(RBLOCK) void k_resample_l2(
    const uint32_t *__restrict__ w, const uint32_t *__restrict__ chunk_start,
    uint32_t B, uint32_t K, const uint64_t *__restrict__ seeds,
    const uint32_t *__restrict__ m, uint32_t *__restrict__ out)
{
	__shared__ uint32_t cdf[HEMK_RS_CHUNK + 1];
	__shared__ uint32_t hist[HEMK_RS_CHUNK];
	const uint32_t j = blockIdx.x, r = blockIdx.y;
	const uint32_t k0 = chunk_start[j], k1 = chunk_start[j + 1];
	const uint32_t nk = k1 - k0;
	const uint32_t mj = m[(size_t)r * B + j];
	uint32_t *o = out + (size_t)r * K + k0;
	if (nk == 1) {
		if (threadIdx.x == 0)
			o[0] = mj;
		return;
	}
	for (uint32_t i = threadIdx.x; i < nk; i += RBLOCK)
		hist[i] = 0;
	if (mj == 0) {
		for (uint32_t i = threadIdx.x; i < nk; i += RBLOCK)
			o[i] = 0;
		return;
	}
	// exclusive prefix of the chunk's weights (serial: nk <= HEMK_RS_CHUNK)
	if (threadIdx.x == 0) {
		uint32_t acc = 0;
		for (uint32_t i = 0; i < nk; ++i) {
			cdf[i] = acc;
			acc += w[k0 + i];
		}
		cdf[nk] = acc;
	}
	__syncthreads();
	const uint32_t W = cdf[nk];
	const uint64_t key = seeds[r];
	for (uint32_t i = threadIdx.x; i < mj; i += RBLOCK) {
		const uint32_t y = draw_bounded(key, i, j, W);
		atomicAdd(&hist[upper_index(cdf, nk, y)], 1u);
	}
	__syncthreads();
	for (uint32_t i = threadIdx.x; i < nk; i += RBLOCK)
		o[i] = hist[i];
}

// Row gather: out[q*R + r] = in[idx[q]*R + r] (class rows into tile order).
__global__ __launch_bounds__(RBLOCK) void k_gather_rows_u32(
    const uint32_t *__restrict__ in, const uint32_t *__restrict__ idx,
    uint32_t rows, uint32_t R, uint32_t *__restrict__ out)
{
	const size_t n = (size_t)rows * R;
	for (size_t i = (size_t)blockIdx.x * RBLOCK + threadIdx.x; i < n;
	     i += (size_t)gridDim.x * RBLOCK) {
		const size_t q = i / R, r = i % R;
		out[i] = in[(size_t)idx[q] * R + r];
	}
}

inline uint32_t cdiv(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

// ---- hera's sampler (src/EM.c:98-150) ----------------------------------
// The reference lays the unique reads out as N slots (create_distribution,
// src/EM.c:98-113: slot k belongs to the transcript t with
// cum[t-1] <= k < cum[t], cum the running sum of the unique counts) and
// draws, for n = N-1 down to 1, slot rand_r() % n (src/EM.c:134-135):
// N - 1 draws from glibc's rand_r, which advances its 32-bit state by three
// steps of the LCG x -> 1103515245 x + 12345 (mod 2^32) per call and
// returns 11 + 10 + 10 bits taken from bits 16.. of the three states.  The
// LCG has full period, so the state before draw i is 3i steps past the
// replicate's first state: every draw is computed on its own.
#define HERA_A 1103515245u
#define HERA_C 12345u

// (A, C) with x_k = A x_0 + C after k steps.
__device__ __forceinline__ void lcg_jump(uint32_t k, uint32_t &A, uint32_t &C)
{
	uint32_t a = HERA_A, c = HERA_C;   // the 2^b-step map
	A = 1u;
	C = 0u;
	while (k) {
		if (k & 1u) {
			C = a * C + c;
			A = a * A;
		}
		c = a * c + c;
		a = a * a;
		k >>= 1;
	}
}

__device__ __forceinline__ uint32_t hera_rand(uint32_t &x)
{
	uint32_t r;
	x = HERA_A * x + HERA_C;
	r = (x >> 16) % 2048u;
	x = HERA_A * x + HERA_C;
	r = (r << 10) ^ ((x >> 16) % 1024u);
	x = HERA_A * x + HERA_C;
	r = (r << 10) ^ ((x >> 16) % 1024u);
	return r;
}

// bt[b] = the transcript of slot b << sh (first t with cum[t] > b << sh).
__global__ __launch_bounds__(RBLOCK) void k_hera_buckets(
    const uint32_t *__restrict__ cum, uint32_t T, uint32_t sh, uint32_t nb,
    uint32_t *__restrict__ bt)
{
	for (uint32_t b = blockIdx.x * RBLOCK + threadIdx.x; b < nb;
	     b += gridDim.x * RBLOCK) {
		const uint64_t k = (uint64_t)b << sh;
		uint32_t lo = 0, hi = T;
		while (lo < hi) {
			const uint32_t mid = lo + (hi - lo) / 2;
			if (cum[mid] > k)
				hi = mid;
			else
				lo = mid + 1;
		}
		bt[b] = lo;
	}
}

// Rows out[r][K]: zero unique counts, the table's class counts.
__global__ __launch_bounds__(RBLOCK) void k_hera_rows(
    const uint32_t *__restrict__ cls, uint32_t T, uint32_t C, uint32_t R,
    uint32_t *__restrict__ out)
{
	const uint32_t K = T + C;
	const uint64_t n = (uint64_t)R * K;
	for (uint64_t i = (uint64_t)blockIdx.x * RBLOCK + threadIdx.x; i < n;
	     i += (uint64_t)gridDim.x * RBLOCK) {
		const uint32_t k = (uint32_t)(i % K);
		out[i] = k < T ? 0u : cls[k - T];
	}
}

// Draws: thread handles HERA_DPT consecutive draws of replicate
// blockIdx.y; replicate r's first state is 3 (N-1) r steps past seed.
#define HERA_DPT 32
__global__ __launch_bounds__(RBLOCK) void k_hera_draw(
    const uint32_t *__restrict__ cum, uint32_t T, uint32_t N,
    const uint32_t *__restrict__ bt, uint32_t sh, uint32_t seed,
    uint32_t *__restrict__ out, uint32_t K)
{
	const uint32_t M = N - 1;   // draws per replicate (N >= 2)
	const uint32_t r = blockIdx.y;
	const uint32_t i0 =
	    (blockIdx.x * RBLOCK + threadIdx.x) * (uint32_t)HERA_DPT;
	if (i0 >= M)
		return;
	uint32_t A, C;
	// steps mod 2^32 (the period): 3 (M r + i0)
	lcg_jump((uint32_t)(3ull * ((uint64_t)M * r + i0)), A, C);
	uint32_t x = A * seed + C;
	uint32_t *row = out + (size_t)r * K;
	const uint32_t i1 = min(M, i0 + (uint32_t)HERA_DPT);
	for (uint32_t i = i0; i < i1; ++i) {
		const uint32_t slot = hera_rand(x) % (M - i);
		uint32_t t = bt[slot >> sh];
		while (cum[t] <= slot)
			++t;
		atomicAdd(row + t, 1u);
	}
}

} // namespace

extern "C" {

// out[r][off_u + k] = in[r][k] (k < Tq) and out[r][off_c + k] =
// in[r][Tq + k] (k < Cq): one sample's resampled rows into the handle's
// unique and class blocks (rows K words long).
__global__ __launch_bounds__(RBLOCK) void k_scatter_counts(
    const uint32_t *__restrict__ in, uint32_t R, uint32_t Tq, uint32_t Cq,
    uint32_t *__restrict__ out, uint32_t K, uint32_t off_u, uint32_t off_c)
{
	const uint32_t Kq = Tq + Cq;
	const uint64_t n = (uint64_t)R * Kq;
	for (uint64_t i = (uint64_t)blockIdx.x * RBLOCK + threadIdx.x; i < n;
	     i += (uint64_t)gridDim.x * RBLOCK) {
		const uint32_t r = (uint32_t)(i / Kq), k = (uint32_t)(i % Kq);
		const uint32_t o = k < Tq ? off_u + k : off_c + (k - Tq);
		out[(size_t)r * K + o] = in[i];
	}
}

int hemk_scatter_counts(const uint32_t *in, uint32_t R, uint32_t Tq,
                        uint32_t Cq, uint32_t *out, uint32_t K,
                        uint32_t off_u, uint32_t off_c, void *s)
{
	uint32_t grid = cdiv((uint64_t)R * (Tq + Cq), RBLOCK);
	if (!grid)
		return 0;
	if (grid > 16384)
		grid = 16384;
	k_scatter_counts<<<grid, RBLOCK, 0, (hipStream_t)s>>>(in, R, Tq, Cq, out,
	                                                      K, off_u, off_c);
	return (int)hipGetLastError();
}

int hemk_hera_resample(const uint32_t *cum, uint32_t T, uint32_t N,
                       uint32_t *bt, uint32_t sh, const uint32_t *cls,
                       uint32_t C, uint32_t seed, uint32_t R, uint32_t *out,
                       void *s)
{
	hipStream_t st = (hipStream_t)s;
	const uint32_t K = T + C;
	if (T == 0 || R == 0 || R > 65535 || sh > 31)
		return (int)hipErrorInvalidValue;
	uint32_t grid = cdiv((uint64_t)R * K, RBLOCK);
	k_hera_rows<<<grid > 16384 ? 16384 : grid, RBLOCK, 0, st>>>(cls, T, C,
	                                                            R, out);
	if (N >= 2) {
		const uint32_t nb = (uint32_t)(((uint64_t)(N - 1) >> sh) + 1);
		grid = cdiv(nb, RBLOCK);
		k_hera_buckets<<<grid > 8192 ? 8192 : grid, RBLOCK, 0, st>>>(
		    cum, T, sh, nb, bt);
		k_hera_draw<<<dim3(cdiv(cdiv(N - 1, HERA_DPT), RBLOCK), R),
		              RBLOCK, 0, st>>>(cum, T, N, bt, sh, seed, out, K);
	}
	return (int)hipGetLastError();
}

int hemk_fill_u32(uint32_t *p, size_t n, uint32_t v, void *s)
{
	if (!n)
		return 0;
	uint32_t grid = cdiv(n, RBLOCK);
	if (grid > 8192)
		grid = 8192;
	k_fill_u32<<<grid, RBLOCK, 0, (hipStream_t)s>>>(p, n, v);
	return (int)hipGetLastError();
}

int hemk_resample(const uint32_t *w, uint32_t K, const uint32_t *chunk_start,
                  const uint32_t *chunk_cw, uint32_t B, uint32_t N,
                  const uint64_t *seeds, uint32_t R, uint32_t *m,
                  uint32_t *out, void *s)
{
	hipStream_t st = (hipStream_t)s;
	int e = hemk_fill_u32(m, (size_t)R * B, 0, s);
	if (e)
		return e;
	if (N) {
		const uint32_t seg = 1u << 20;
		dim3 g1(cdiv(N, seg), R);
		const size_t lds = ((size_t)2 * B + 1) * 4;
		k_resample_l1<<<g1, RBLOCK, lds, st>>>(chunk_cw, B, N, seg, seeds, m);
		e = (int)hipGetLastError();
		if (e)
			return e;
	}
	dim3 g2(B, R);
	k_resample_l2<<<g2, RBLOCK, 0, st>>>(w, chunk_start, B, K, seeds, m, out);
	return (int)hipGetLastError();
}

int hemk_gather_rows_u32(const uint32_t *in, const uint32_t *idx,
                         uint32_t rows, uint32_t R, uint32_t *out, void *s)
{
	if (!rows || !R)
		return 0;
	uint32_t grid = cdiv((uint64_t)rows * R, RBLOCK);
	if (grid > 16384)
		grid = 16384;
	k_gather_rows_u32<<<grid, RBLOCK, 0, (hipStream_t)s>>>(in, idx, rows, R,
	                                                       out);
	return (int)hipGetLastError();
}

} // extern "C"
