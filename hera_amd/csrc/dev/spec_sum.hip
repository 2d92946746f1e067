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
//   2. each chunk's drift d_j, its end minus its start when run from the
//      guess g0_j = sum_{i<j} c_i: exactly F_j(g0_j) - g0_j for signed
//      rows (whose running sum can return close to zero, where a start a
//      few ulps off changes the rounding), and for sums of squares the
//      rounded increments fl(s + y) - s over a block scan from g0_j;
//      plain chunk sums miss the rounding by up to ~10^4 ulps;
//   3. g_j = sum_{i<j} d_i (a fixed expression, so every kernel forms the
//      same bits); every chunk runs F_j from the W consecutive doubles
//      centred on g_j (one candidate per lane; W = 64 or 128 for signed
//      rows, 16 for squares), all at once, and looks each result up in
//      the next chunk's window (a link, or none);
//   4. one workgroup per chain follows the links from s_0 = +0, skipping
//      ahead over links composed in LDS; where the exact start of a chunk
//      lies outside its window (a miss), that chunk is run again from it
//      and the next ones from windows shifted by the guess's error, so the
//      result never depends on how good the guesses were.
// Steps 3-4 are FP64 additions in the reference's order on values rounded
// as the reference rounds them; steps 1-2 only decide hits and misses.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>

#include "hem_dev.h"
#include "spec_common.h"

namespace {

// Ordered integer view of a double: consecutive keys are consecutive
// doubles (-0.0 and +0.0 are neighbours), and the map is a bijection on
// all 2^64 bit patterns, so dunkey(dkey(x)) has x's bits.
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

// The same key from the two 32-bit halves (scalar-unit friendly): negative
// patterns flip every bit but the sign.
__device__ __forceinline__ unsigned long long ukey(double x)
{
	const unsigned long long b = (unsigned long long)__double_as_longlong(x);
	const uint32_t hi = (uint32_t)(b >> 32), lo = (uint32_t)b;
	const uint32_t t = (uint32_t)((int32_t)hi >> 31);
	return ((unsigned long long)(hi ^ (t & 0x7FFFFFFFu)) << 32) | (lo ^ t);
}

__device__ __forceinline__ double readlane_d(double v, uint32_t l)
{
	const long long b = __double_as_longlong(v);
	const int lo = __builtin_amdgcn_readlane((int)b, (int)l);
	const int hi = __builtin_amdgcn_readlane((int)(b >> 32), (int)l);
	return __longlong_as_double(
	    (long long)(((unsigned long long)(unsigned)hi << 32) |
	                (unsigned)lo));
}

// Offset of key k in the window starting at key b, as a 64-bit wrapping
// difference.  A window is the W keys b, b+1, ... taken with the same
// wrapping addition, so offset < W identifies the candidate whose start has
// exactly the bits of the looked-up value, wherever the window lies.
__device__ __forceinline__ bool in_window(unsigned long long k, long long b,
                                          uint32_t w, uint32_t &idx)
{
	const unsigned long long d = k - (unsigned long long)b;
	idx = (uint32_t)d;
	return (uint32_t)(d >> 32) == 0 && (uint32_t)d < w;
}

__device__ __forceinline__ long long kadd(long long a, long long b)
{
	return (long long)((unsigned long long)a + (unsigned long long)b);
}

// Step 1: chunk sums.  Grid (P, chains), SBLOCK threads.
template <int XF>
__global__ __launch_bounds__(SBLOCK) void k_spec_csum(hemk_spec sp,
                                                      const hemk_ctl *ctl)
{
#pragma clang fp contract(off)
	uint32_t row;
	if (skip(ctl) || !chain_row(sp, blockIdx.y, row))
		return;
	const uint32_t j = blockIdx.x;
	const uint32_t beg = j * sp.L;
	const uint32_t end = min(sp.n, beg + sp.L);
	const double m = chain_mean<XF>(sp, row);
	const double *x = sp.x + (size_t)row * sp.ld;
	double a = 0.0, b = 0.0;
	for (uint32_t i = beg + threadIdx.x; i < end; i += SBLOCK) {
		a += xval<XF>(x[i], m);
		b += x[i] * x[i];
	}
	a = block_sum(a);
	if (XF == HEMK_SPEC_IDENT && sp.csq)
		b = block_sum(b);
	if (threadIdx.x == 0) {
		sp.csum[(size_t)blockIdx.y * sp.P + j] = a;
		if (XF == HEMK_SPEC_IDENT && sp.csq)
			sp.csq[(size_t)blockIdx.y * sp.P + j] = b;
	}
}

// Plain solve, pass 1: r = u1 - u and v = (u2 - u1) - r (src/EM.c:333-336)
// into rows 0 and 1 of x, and the chunk sums of their chain values (XF
// IDENT or SQUARE).  Grid (P, 2).
template <int XF>
__global__ __launch_bounds__(SBLOCK) void k_spec_rv(
    hemk_spec sp, const double *__restrict__ u, const double *__restrict__ u1,
    const double *__restrict__ u2, double *__restrict__ xw,
    const hemk_ctl *ctl)
{
#pragma clang fp contract(off)
	if (skip(ctl))
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
		a += xval<XF>(val, 0.0);
	}
	a = block_sum(a);
	if (threadIdx.x == 0)
		sp.csum[(size_t)q * sp.P + j] = a;
}

// Step 2: drift of each chunk.  Grid (P, chains), SBLOCK threads; thread t
// holds L/SBLOCK consecutive values.  The running sum before each value is
// g0_j plus a block scan of the chunk (approximate, like g0 itself); the
// drift adds up the rounded increments fl(s + y) - s.
template <int XF>
__global__ __launch_bounds__(SBLOCK) void k_spec_drift(hemk_spec sp,
                                                       const hemk_ctl *ctl)
{
#pragma clang fp contract(off)
	__shared__ double wtot[SBLOCK / WAVE];
	uint32_t row;
	if (skip(ctl) || !chain_row(sp, blockIdx.y, row))
		return;
	const uint32_t j = blockIdx.x;
	const size_t cb = (size_t)blockIdx.y * sp.P;
	const int tid = threadIdx.x, lane = tid % WAVE, w = tid / WAVE;
	const double m = chain_mean<XF>(sp, row);
	double g = 0.0;
	if (XF == HEMK_SPEC_CENTRED && sp.gsq) {
		// chunks before j are full: sum of (x - m)^2 over one is
		// sq - 2 m s + L m^2
		const double lm2 = (double)sp.L * m * m;
		for (uint32_t i = tid; i < j; i += SBLOCK)
			g += (sp.gsq[cb + i] - 2.0 * m * sp.gsum[cb + i]) + lm2;
	} else {
		for (uint32_t i = tid; i < j; i += SBLOCK)
			g += sp.csum[cb + i];
	}
	const double g0 = block_sum(g);
	const uint32_t E = sp.L / SBLOCK;   // 1..8
	const uint32_t beg = j * sp.L + tid * E;
	const double *x = sp.x + (size_t)row * sp.ld;
	double y[8], tot = 0.0;
#pragma unroll
	for (int e = 0; e < 8; ++e) {
		y[e] = (e < (int)E && beg + e < sp.n) ? xval<XF>(x[beg + e], m)
		                                       : 0.0;
		tot += y[e];
	}
	// exclusive scan of the thread totals
	double inc = tot;
#pragma unroll
	for (int k = 1; k < WAVE; k <<= 1) {
		const double o = __shfl_up(inc, k, WAVE);
		if (lane >= k)
			inc += o;
	}
	if (lane == WAVE - 1)
		wtot[w] = inc;
	__syncthreads();
	double off = inc - tot;
	for (int k = 0; k < w; ++k)
		off += wtot[k];
	double sh = g0 + off, q = 0.0;
#pragma unroll
	for (int e = 0; e < 8; ++e) {
		const double nx = sh + y[e];
		q += nx - sh;
		sh = nx;
	}
	q = block_sum(q);
	if (tid == 0)
		sp.drift[cb + j] = q;
}

// Step 2 for signed chains: each chunk's drift run exactly, d_j = F_j(g0_j)
// - g0_j, from the sum g0_j of the earlier chunk sums.  A chunk's rounding
// depends on where its start lies; a start within a few ulps of the true
// one reproduces most of it, so these drifts leave about 5x fewer misses
// than the block-scan drift on signed rows whose running sum returns close
// to zero.  Grid (ceil(P / 4), chains), SBLOCK threads, one wave per chunk
// (every lane runs the same chain; its values are read by broadcast).
template <int XF>
__global__ __launch_bounds__(SBLOCK) void k_spec_seqdrift(hemk_spec sp,
                                                          const hemk_ctl *ctl)
{
#pragma clang fp contract(off)
	constexpr int NW = SBLOCK / WAVE;
	__shared__ double ys[NW * HEMK_SPEC_LMAX];
	uint32_t row;
	if (skip(ctl) || !chain_row(sp, blockIdx.y, row))
		return;
	const int w = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
	const uint32_t j = blockIdx.x * NW + w;
	const size_t cb = (size_t)blockIdx.y * sp.P;
	const bool live = j < sp.P;
	const uint32_t cnt = live ? min(sp.L, sp.n - j * sp.L) : 0;
	double v[CH_REG];
	chunk_load(sp.x + (size_t)row * sp.ld + (size_t)j * sp.L, cnt, v);
	// g0 = the earlier chunk sums (any order: it only steers the guess),
	// eight loads per lane in flight with the chunk's
	double g = 0.0;
	for (uint32_t i = lane; live && i < j; i += 8 * WAVE) {
		double c[8];
#pragma unroll
		for (int q = 0; q < 8; ++q)
			c[q] = i + q * WAVE < j ? sp.csum[cb + i + q * WAVE] : 0.0;
#pragma unroll
		for (int q = 0; q < 8; ++q)
			g += c[q];
	}
#pragma unroll
	for (int k = WAVE / 2; k >= 1; k >>= 1)
		g += __shfl_xor(g, k, WAVE);
	double *yw = ys + w * HEMK_SPEC_LMAX;
	chunk_store<XF>(v, cnt, chain_mean<XF>(sp, row), yw);
	__syncthreads();
	if (!live)
		return;
	const double s = run_chain(g, yw, cnt);
	if (lane == 0)
		sp.drift[cb + j] = s - g;
}

// Guess of chunk j's start from the drifts d:
//   g(j) = ((B_k + d_64k) + d_64k+1) + ... + d_j-1,   k = j / 64,
//   B_k  = ((bs_0 + bs_1) + ...) + bs_k-1,  bs_i = xor-tree sum of block i.
// Whoever needs the guess of a chunk evaluates this same expression, so the
// window base that chunk j computes for chunk j + 1 is bitwise the one
// chunk j + 1 runs from.  The wave's chunks j0 .. j0 + CPV - 1 share block
// k = j0 / 64 (CPV divides 64); every lane gets g(j) and g(j + 1) for its j.
__device__ __forceinline__ double tree_sum(double v)
{
#pragma clang fp contract(off)
#pragma unroll
	for (int k = WAVE / 2; k >= 1; k >>= 1)
		v += __shfl_xor(v, k, WAVE);
	return v;
}

template <int CPV>
__device__ __forceinline__ void chunk_guess(const double *d, uint32_t P,
                                            uint32_t j0, uint32_t j,
                                            double &gj, double &gn)
{
#pragma clang fp contract(off)
	const int lane = threadIdx.x % WAVE;
	const uint32_t k = j0 / WAVE;
	double B = 0.0;
	uint32_t b = 0;
	// full blocks before k, eight at a time (loads first, then the trees,
	// then B in block order); block k's drifts load with the last group
	double dv = 0.0;
	for (;;) {
		double t[8];
#pragma unroll
		for (int q = 0; q < 8; ++q)
			if (b + q < k)
				t[q] = d[(b + q) * WAVE + lane];
		const bool lastg = b + 8 >= k;
		if (lastg && k * WAVE + lane < P)
			dv = d[k * WAVE + lane];
#pragma unroll
		for (int q = 0; q < 8; ++q)
			if (b + q < k)
				t[q] = tree_sum(t[q]);
#pragma unroll
		for (int q = 0; q < 8; ++q)
			if (b + q < k)
				B += t[q];
		if (lastg)
			break;
		b += 8;
	}
	const double bs = tree_sum(dv);
	double g = B;
	const uint32_t top = j0 + CPV - 1 - k * WAVE;   // largest j - 64k
	for (uint32_t t = 0; t < top; ++t) {
		const double y = readlane_d(dv, t);
		if (k * WAVE + t < j)
			g += y;
	}
	gj = g;
	const uint32_t lj = j - k * WAVE;
	const double dj = __shfl(dv, (int)lj, WAVE);
	gn = lj + 1 == WAVE ? B + bs : g + dj;
}

// Step 3: every chunk from W consecutive starts around its guess.  Grid
// (P, chains), one workgroup of max(W, 64) threads per chunk (its values
// staged by the first wave in one round of loads); thread c < W runs the
// chunk from key(g) - W/2 + c.  Besides the
// results (tab) and the window base, each candidate's result is looked up
// in the next chunk's window: nx holds its offset there, or
// HEMK_SPEC_NOLINK.
template <int W, int XF>
__global__ __launch_bounds__(W > WAVE ? W : WAVE) void k_spec_window(
    hemk_spec sp, const hemk_ctl *ctl)
{
#pragma clang fp contract(off)
	__shared__ double ys[HEMK_SPEC_LMAX];
	uint32_t row;
	if (skip(ctl) || !chain_row(sp, blockIdx.y, row))
		return;
	const int c = threadIdx.x;
	const uint32_t j = blockIdx.x;
	const size_t cb = (size_t)blockIdx.y * sp.P;
	const uint32_t cnt = min(sp.L, sp.n - j * sp.L);
	const bool stager = c < WAVE;
	double v[CH_REG];
	if (stager)
		chunk_load(sp.x + (size_t)row * sp.ld + (size_t)j * sp.L, cnt,
		           v);
	double g, gn;
	chunk_guess<1>(sp.drift + cb, sp.P, j, j, g, gn);
	if (stager)
		chunk_store<XF>(v, cnt, chain_mean<XF>(sp, row), ys);
	__syncthreads();
	if (c >= W)
		return;
	const long long base = kadd(dkey(g), -(W / 2));
	const double s = run_chain(dunkey(kadd(base, c)), ys, cnt);
	sp.tab[(cb + j) * W + c] = s;
	uint32_t link = HEMK_SPEC_NOLINK, idx;
	if (j + 1 < sp.P &&
	    in_window(ukey(s), kadd(dkey(gn), -(W / 2)), W, idx))
		link = idx;
	sp.nx[(cb + j) * W + c] = (uint8_t)link;
	if (c == 0)
		sp.base[cb + j] = base;
}

// Repairs after a miss.  A miss at chunk j means the guess g_j was off by
// delta = s_j - g_j, and the following guesses are usually off by about the
// same amount.  Waves 0 and 1 run chunks j .. j+RC-1 at once: group 0 from
// the exact s_j, group k from the RW doubles centred on g_{j+k} + delta;
// results go to rtl[k*RW + c], window bases to rbl[k].  Eight chunks per
// repair (rather than four) roughly halve the longest run of repairs in a
// row, which sets the resolve time when many rows resolve at once.
#define RW 16
#define RC 8
#define REP_LDS (RC * HEMK_SPEC_LMAX * 8)   // a repair's chunk values
static_assert(RC * RW <= SBLOCK, "repair lanes fit the workgroup");
// chunks whose links the resolve kernel holds in LDS at once: 64 KB of
// links for W = 64 and W = 128
#define SEGP(W) ((W) > 64 ? 512 : 1024)

template <int W, int XF>
__device__ __forceinline__ void repair(const hemk_spec &sp, uint32_t row,
                                       size_t cb, uint32_t j, double m,
                                       double s, double *ys, double *rtl,
                                       long long *rbl)
{
#pragma clang fp contract(off)
	const uint32_t nch = min((uint32_t)RC, sp.P - j);
	const uint32_t tot = min(nch * sp.L, sp.n - j * sp.L);
	stage<XF>(sp.x + (size_t)row * sp.ld + (size_t)j * sp.L, tot, m, ys,
	          SBLOCK);
	__syncthreads();
	if (threadIdx.x < RC * RW) {
		const int lane = threadIdx.x, k = lane / RW, c = lane % RW;
		const uint32_t jj = j + k;
		const double delta = s - dunkey(kadd(sp.base[cb + j], W / 2));
		long long b = 0;
		double st = s;
		if (k > 0 && jj < sp.P) {
			const double g = dunkey(kadd(sp.base[cb + jj], W / 2)) + delta;
			b = kadd(dkey(g), -(RW / 2));
			st = dunkey(kadd(b, c));
		}
		double r = 0.0;
		if (jj < sp.P)
			r = run_chain(st, ys + k * sp.L,
			              min(sp.L, sp.n - jj * sp.L));
		rtl[lane] = r;
		if (c == 0)
			rbl[k] = b;
	}
	__syncthreads();
}

__device__ __forceinline__ uint32_t ufirst(uint32_t v)
{
	return (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
}

// Step 4: follow the chain through the windows, one workgroup per chain.
// The walk holds either the exact start s of chunk j or the offset e of
// that start in chunk j's window.  With an offset, the next chunk's offset
// is nx[j][e]; composed links over 16 and 256 chunks (built in LDS per
// segment of SEGP(W) chunks) let it skip ahead, and the result of the last
// chunk is tab[P-1][e].  A missing link means chunk j+1's exact start
// tab[j][e] lies outside its window: the walk takes that value, and a start
// that is in neither its window nor the current repair's starts a repair.
// Every wave walks redundantly (all branches uniform), so all threads can
// stage a repair's values.
template <int W, int XF>
__global__ __launch_bounds__(SBLOCK) void k_spec_resolve(hemk_spec sp,
                                                         const hemk_ctl *ctl)
{
#pragma clang fp contract(off)
	constexpr uint32_t SEG = SEGP(W);
	__shared__ uint8_t nxl[SEG * W];
	__shared__ uint8_t c1[SEG / 16 * W];
	__shared__ uint8_t c2[SEG / 256 * W];
	__shared__ double ys[REP_LDS / 8];
	__shared__ double rtl[RC * RW];
	__shared__ long long rbl[RC];
	uint32_t row;
	if (skip(ctl) || !chain_row(sp, blockIdx.x, row))
		return;
	const int tid = threadIdx.x;
	const size_t cb = (size_t)blockIdx.x * sp.P;
	const double m = chain_mean<XF>(sp, row);
	const uint32_t P = sp.P;
	double s = 0.0;
	uint32_t j = 0, e = 0, rj = 0, misses = 0;
	bool have = false, rep = false;
	for (uint32_t jb = 0; jb < P; jb += SEG) {
		const uint32_t ns = min(SEG, P - jb);
		const uint32_t last = P - 1 - jb;
		__syncthreads();
		{
			// W is a multiple of 8: 8-byte pieces, eight in flight
			const unsigned long long *src =
			    (const unsigned long long *)(sp.nx + (cb + jb) * W);
			unsigned long long *dst = (unsigned long long *)nxl;
			const uint32_t cnt = ns * W / 8;
			for (uint32_t i = tid; i < cnt; i += 8 * SBLOCK) {
				unsigned long long v[8];
#pragma unroll
				for (int q = 0; q < 8; ++q)
					if (i + q * SBLOCK < cnt)
						v[q] = src[i + q * SBLOCK];
#pragma unroll
				for (int q = 0; q < 8; ++q)
					if (i + q * SBLOCK < cnt)
						dst[i + q * SBLOCK] = v[q];
			}
		}
		__syncthreads();
		// links over 16 chunks: block b maps an offset at chunk jb + 16b
		// to one at jb + 16b + 16 (that chunk exists, its links in LDS)
		// (with wide windows four entries per thread step together, so
		// their dependent LDS reads overlap)
		constexpr int IL = W >= 64 ? 4 : 1;
		const uint32_t nb1 = min(ns / 16, last / 16);
		const uint32_t n1 = nb1 * W;
		for (uint32_t it0 = tid; it0 < n1; it0 += IL * SBLOCK) {
			uint32_t x[IL], b[IL];
#pragma unroll
			for (int q = 0; q < IL; ++q) {
				const uint32_t it = it0 + q * SBLOCK;
				b[q] = it / W;
				x[q] = it < n1 ? it % W : HEMK_SPEC_NOLINK;
			}
#pragma unroll
			for (int t = 0; t < 16; ++t)
#pragma unroll
				for (int q = 0; q < IL; ++q)
					if (x[q] != HEMK_SPEC_NOLINK)
						x[q] = nxl[(16 * b[q] + t) * W + x[q]];
#pragma unroll
			for (int q = 0; q < IL; ++q)
				if (it0 + q * SBLOCK < n1)
					c1[it0 + q * SBLOCK] = (uint8_t)x[q];
		}
		__syncthreads();
		const uint32_t nb2 = min(ns / 256, last / 256);
		for (uint32_t it = tid; it < nb2 * W; it += SBLOCK) {
			const uint32_t b = it / W;
			uint32_t x = it % W;
			for (int t = 0; t < 16 && x != HEMK_SPEC_NOLINK; ++t)
				x = c1[(16 * b + t) * W + x];
			c2[it] = (uint8_t)x;
		}
		__syncthreads();
		while (j < jb + ns) {
			if (!have) {
				const unsigned long long ks = ukey(s);
				uint32_t idx;
				if (in_window(ks, readfirst_ll(sp.base[cb + j]), W,
				              idx)) {
					e = ufirst(idx);
					have = true;
				} else if (rep && j - rj < RC &&
				           in_window(ks, readfirst_ll(rbl[j - rj]), RW,
				                     idx)) {
					s = readfirst_d(rtl[(j - rj) * RW + ufirst(idx)]);
					++j;
					continue;
				} else {
					repair<W, XF>(sp, row, cb, j, m, s, ys, rtl, rbl);
					rj = j;
					rep = true;
					s = readfirst_d(rtl[0]);
					++misses;
					++j;
					continue;
				}
			}
			if (j == P - 1)
				break;
			const uint32_t q = j - jb;
			uint32_t x = HEMK_SPEC_NOLINK;
			if ((q & 255) == 0 && (q >> 8) < nb2)
				x = ufirst(c2[(q >> 8) * W + e]);
			if (x != HEMK_SPEC_NOLINK) {
				e = x;
				j += 256;
				continue;
			}
			if ((q & 15) == 0 && (q >> 4) < nb1)
				x = ufirst(c1[(q >> 4) * W + e]);
			if (x != HEMK_SPEC_NOLINK) {
				e = x;
				j += 16;
				continue;
			}
			x = ufirst(nxl[q * W + e]);
			if (x != HEMK_SPEC_NOLINK) {
				e = x;
				++j;
				continue;
			}
			s = readfirst_d(sp.tab[(cb + j) * W + e]);
			have = false;
			++j;
		}
	}
	if (have)
		s = readfirst_d(sp.tab[(cb + P - 1) * W + e]);
	if (tid == 0) {
		sp.res[row] = s;
		if (misses && sp.miss)
			atomicAdd(sp.miss, misses);
	}
}

template <int XF>
int launch_xf(const hemk_spec &sp, int produce, int phases, const double *u,
              const double *u1, const double *u2, const hemk_ctl *ctl,
              hipStream_t st)
{
	dim3 grid(sp.P, sp.nchain);
	if (produce)
		k_spec_rv<XF><<<grid, SBLOCK, 0, st>>>(sp, u, u1, u2,
		                                       (double *)sp.x, ctl);
	else if ((phases & HEMK_SPEC_PH_CSUM) && !(sp.wa && sp.cpub))
		k_spec_csum<XF><<<grid, SBLOCK, 0, st>>>(sp, ctl);
	if (sp.wa) {
		// chunk programs and the walk need no drifts or windows
		if (!(phases & HEMK_SPEC_PH_FINISH))
			return (int)hipGetLastError();
		const int e = (int)hipGetLastError();
		return e ? e : hemk_ws_finish(&sp, ctl, st);
	}
	if (phases & HEMK_SPEC_PH_DRIFT) {
		if (XF == HEMK_SPEC_IDENT)
			k_spec_seqdrift<XF>
			    <<<dim3((sp.P + SBLOCK / WAVE - 1) / (SBLOCK / WAVE),
			            sp.nchain),
			       SBLOCK, 0, st>>>(sp, ctl);
		else
			k_spec_drift<XF><<<grid, SBLOCK, 0, st>>>(sp, ctl);
	}
	if (!(phases & HEMK_SPEC_PH_FINISH))
		return (int)hipGetLastError();
#define SPEC_W_CASE(WW)                                                    \
	case WW:                                                           \
		k_spec_window<WW, XF><<<dim3(sp.P, sp.nchain),               \
		                        WW > WAVE ? WW : WAVE, 0, st>>>(sp, ctl); \
		k_spec_resolve<WW, XF><<<sp.nchain, SBLOCK, 0, st>>>(sp, ctl); \
		break;
	switch (sp.W) {
		SPEC_W_CASE(8)
		SPEC_W_CASE(16)
		SPEC_W_CASE(32)
		SPEC_W_CASE(64)
		SPEC_W_CASE(128)
	default:
		return (int)hipErrorInvalidValue;
	}
#undef SPEC_W_CASE
	return (int)hipGetLastError();
}

// out[r] = sqrt(q[r] / (n - 1)): the end of Norm (src/hash_align.c:161-163)
// on the device, for the allocation-free sums.
__global__ void k_sd_finish(const double *__restrict__ q, uint32_t rows,
                            uint32_t n, double *__restrict__ out)
{
	const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
	if (r < rows)
		out[r] = sqrt(q[r] / (double)(n - 1u));
}

} // namespace

// Fixed device scratch of hem_exact_sums_noalloc: chunk sums, drifts and
// window bases of up to HEMK_NOALLOC_CHUNKS chunks, their candidate tables
// and links at the widest window, and two sums per row.  Every call on a
// device uses that device's copy; the host reaches it only through the
// address hipGetSymbolAddress returns for the current device.
static __device__ __align__(256) unsigned char scratch[HEMK_NOALLOC_BYTES];

static std::mutex scratch_mu;
static void *scratch_dev[HEMK_NOALLOC_DEVICES];

extern "C" {

int hemk_noalloc_scratch(void **ptr, size_t *bytes)
{
	int dev = 0;
	hipError_t e = hipGetDevice(&dev);
	if (e != hipSuccess)
		return (int)e;
	if (dev < 0 || dev >= HEMK_NOALLOC_DEVICES)
		return (int)hipErrorInvalidDevice;
	std::lock_guard<std::mutex> g(scratch_mu);
	if (!scratch_dev[dev]) {
		void *p = nullptr;
		e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
		if (e != hipSuccess)
			return (int)e;
		scratch_dev[dev] = p;
	}
	*ptr = scratch_dev[dev];
	*bytes = sizeof(scratch);
	return 0;
}

int hemk_sd_finish(const double *q, uint32_t rows, uint32_t n, double *out,
                   void *s)
{
	if (!rows)
		return 0;
	k_sd_finish<<<(rows + 255) / 256, 256, 0, (hipStream_t)s>>>(q, rows, n,
	                                                           out);
	return (int)hipGetLastError();
}

uint32_t hemk_spec_chunk(uint32_t n)
{
	// ~L dependent adds per window run against ~n/L dependent lookups;
	// long chunks also average out the guesses' errors
	return n <= 16384 ? 256 : HEMK_SPEC_LMAX;
}

uint32_t hemk_spec_width(uint32_t L, int xf, uint32_t chunks)
{
	// Sums of squares are guessed within ~2 ulps: 16 candidates.  Signed
	// sums (whose running sum can pass close to zero) miss more: 64
	// candidates, or 128 (two waves per chunk) when the window launch
	// has few enough chunks that the second wave finds an idle SIMD;
	// on SQUAREM's r and v rows of 200k values that turns about a third
	// of the resolve's reruns into hits.
	(void)L;
	if (xf != HEMK_SPEC_IDENT)
		return 16;
	return chunks <= 512 ? 128 : 64;
}

size_t hemk_spec_tab_entries(uint32_t L, uint32_t max_rows, uint32_t P)
{
	// the width only narrows as the launch grows, so a smaller launch
	// can need more entries than the largest one
	size_t best = 0;
	for (uint32_t rows = 1; rows <= max_rows; ++rows) {
		const size_t chunks = (size_t)rows * P;
		const size_t w = hemk_spec_width(
		    L, HEMK_SPEC_IDENT,
		    chunks > 0xFFFFFFFFu ? 0xFFFFFFFFu : (uint32_t)chunks);
		if (chunks * w > best)
			best = chunks * w;
	}
	return best;
}

static int spec_launch(const hemk_spec *sp, int produce, int phases,
                       const double *u, const double *u1, const double *u2,
                       hemk_ctl *ctl, void *s)
{
	hipStream_t st = (hipStream_t)s;
	const bool by_row = !sp->pairs && !sp->alist;
	if (sp->n == 0 || sp->nchain == 0 || sp->L == 0 ||
	    sp->L > HEMK_SPEC_LMAX || sp->L % SBLOCK != 0 ||
	    sp->n > 0xFFFFFFFFu - sp->L ||
	    sp->P != (sp->n + sp->L - 1) / sp->L ||
	    (!sp->wa && (sp->W < 8 || sp->W > HEMK_SPEC_W || sp->W % 8 != 0 ||
	                  !sp->nx || !sp->tab || !sp->base)) ||
	    (sp->wa && (!sp->wlh || !sp->wrg || !sp->wcnt)) ||
	    (produce && (sp->nchain != 2 || sp->pairs || sp->alist ||
	                 sp->xf == HEMK_SPEC_CENTRED)) ||
	    (sp->csq && (sp->xf != HEMK_SPEC_IDENT || (!by_row && !sp->wa) ||
	                 produce)) ||
	    (!sp->gsq != !sp->gsum) ||
	    (sp->gsq && (sp->xf != HEMK_SPEC_CENTRED || (!by_row && !sp->wa))))
		return (int)hipErrorInvalidValue;
	switch (sp->xf) {
	case HEMK_SPEC_IDENT:
		return launch_xf<HEMK_SPEC_IDENT>(*sp, produce, phases, u, u1,
		                                  u2, ctl, st);
	case HEMK_SPEC_CENTRED:
		return launch_xf<HEMK_SPEC_CENTRED>(*sp, produce, phases, u, u1,
		                                    u2, ctl, st);
	case HEMK_SPEC_SQUARE:
		return launch_xf<HEMK_SPEC_SQUARE>(*sp, produce, phases, u, u1,
		                                   u2, ctl, st);
	}
	return (int)hipErrorInvalidValue;
}

int hemk_spec_sum(const hemk_spec *sp, int produce, const double *u,
                  const double *u1, const double *u2, hemk_ctl *ctl,
                  void *s)
{
	return spec_launch(sp, produce,
	                   HEMK_SPEC_PH_CSUM | HEMK_SPEC_PH_DRIFT |
	                       HEMK_SPEC_PH_FINISH,
	                   u, u1, u2, ctl, s);
}

int hemk_spec_phases(const hemk_spec *sp, int phases, hemk_ctl *ctl, void *s)
{
	return spec_launch(sp, 0, phases, NULL, NULL, NULL, ctl, s);
}

} // extern "C"
