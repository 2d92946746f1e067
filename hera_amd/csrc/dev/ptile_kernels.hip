// ptile_kernels.hip -- one EM instance over component tiles (gfx950).
//
// The plain solve's maps without the two global passes: a 64-lane
// workgroup owns one tile (whole connected components of the class graph,
// host/tiles.c) and keeps its transcripts, classes and members in LDS.  One
// launch applies the map twice (u -> u1 -> u2, src/EM.c:330-331), the other
// extrapolates (src/EM.c:333-352) and applies it once more with the stop
// test (src/EM.c:354-355).  Every sum runs in the reference's order:
// members in class order for D_c (src/EM.c:277-278), then for each
// transcript its terms (lam_t n_c) / D_c in ascending class order
// (src/EM.c:282-285, the order hera's class loop adds into overlap[t]),
// then (u_t + overlap_t) / len_t (src/EM.c:289-290) -- so every iterate is
// bitwise the reference's.  State vectors are in tile order (a tile's
// transcripts are consecutive); r and v are written in transcript order
// for the exact sums.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "em_step.h"
#include "hem_dev.h"

#define PBLOCK 256

namespace {

struct PTile {
	uint32_t tx_base, nt, cls_base, nc, nz, rec;
};

__device__ __forceinline__ PTile ptile(const hemk_tiles &t, uint32_t i)
{
	const uint32_t *d = t.desc + (size_t)HEMK_TILE_DESC * i;
	PTile c;
	c.tx_base = d[0];
	c.nt = d[1];
	c.cls_base = d[2];
	c.nc = d[3];
	c.nz = d[5];
	c.rec = d[7];
	return c;
}

// Bytes of a tile's record: u16 class offsets (nc + 1), class-list offsets
// (nt + 1), members (c << 6 | t) in class order and in class-list order
// (nz each), then 4-byte aligned u32 lengths and transcript ids (nt each),
// then for a weighted table, 8-byte aligned, FP64 member weights in class
// order and in class-list order (nz each); 16-byte units.
template <bool W>
__device__ __forceinline__ uint32_t rec_bytes(uint32_t nt, uint32_t nc,
                                              uint32_t nz)
{
	uint32_t u16s = (nc + 1) + (nt + 1) + 2 * nz;
	u16s += u16s & 1;
	uint32_t b = 2 * u16s + 8 * nt;
	if (W)
		b = (b + 7) / 8 * 8 + 16 * nz;
	return (b + 15) / 16 * 16;
}

// The tile in LDS: its record (room for the largest), then the FP64 work
// arrays and unique counts sized by the tiling's largest tile.
struct PLds {
	double *l0, *l1, *l2, *l3;  // densities in, after one, two, three maps
	double *D, *X, *n;     // class sums, member terms, class counts
	uint32_t *u;
	uint4 *rec;
	// record fields of the current tile
	const uint16_t *coff, *toff, *kmem, *kocc;
	const uint32_t *len, *tx;
	const double *wc, *wo;  // weighted tables: weights by member, by entry
};

__device__ __forceinline__ PLds pcarve(void *smem, const hemk_tiles &tl)
{
	const uint32_t mt = tl.max_tx, mc = tl.max_cls, mz = tl.max_nnz;
	PLds L;
	L.rec = (uint4 *)smem;
	L.l0 = (double *)((char *)smem + tl.max_rec);
	L.l1 = L.l0 + mt;
	L.l2 = L.l1 + mt;
	L.l3 = L.l2 + mt;
	L.D = L.l3 + mt;
	L.X = L.D + mc;
	L.n = L.X + mz;
	L.u = (uint32_t *)(L.n + mc);
	return L;
}

template <bool W>
__device__ __forceinline__ void pfields(PLds &L, const PTile &tc)
{
	const uint16_t *r = (const uint16_t *)L.rec;
	L.coff = r;
	L.toff = L.coff + tc.nc + 1;
	L.kmem = L.toff + tc.nt + 1;
	L.kocc = L.kmem + tc.nz;
	uint32_t k = (tc.nc + 1) + (tc.nt + 1) + 2 * tc.nz;
	k += k & 1;
	L.len = (const uint32_t *)(r + k);
	L.tx = L.len + tc.nt;
	L.wc = L.wo = nullptr;
	if (W) {
		const uint32_t b = (2 * k + 8 * tc.nt + 7) / 8 * 8;
		L.wc = (const double *)((const char *)r + b);
		L.wo = L.wc + tc.nz;
	}
}

// The tile's record, class counts and unique counts into LDS.  The loads
// of a typical tile (a record up to 4 KB, 128 classes, 64 transcripts)
// are all issued before the first LDS store, so the tile arrives in one
// round of memory latency; larger tiles finish in plain loops.
#define PK_R 4   // 16-byte record pieces per lane in the first round
template <bool W>
__device__ __forceinline__ void pstage(const hemk_tiles &tl, const PTile &tc,
                                       const PLds &L,
                                       const uint32_t *__restrict__ uq,
                                       const double *__restrict__ cntq)
{
	const uint32_t tid = threadIdx.x;
	const uint32_t n16 = rec_bytes<W>(tc.nt, tc.nc, tc.nz) / 16;
	const uint4 *src = (const uint4 *)tl.rec + tc.rec;
	uint4 rv[PK_R];
	double nn[2];
	uint32_t u = 0;
#pragma unroll
	for (int q = 0; q < PK_R; ++q)
		rv[q] = tid + q * PBLOCK < n16 ? src[tid + q * PBLOCK]
		                               : make_uint4(0, 0, 0, 0);
#pragma unroll
	for (int q = 0; q < 2; ++q)
		nn[q] = tid + q * PBLOCK < tc.nc
		            ? cntq[tc.cls_base + tid + q * PBLOCK]
		            : 0.0;
	if (tid < tc.nt)
		u = uq[tc.tx_base + tid];
#pragma unroll
	for (int q = 0; q < PK_R; ++q)
		if (tid + q * PBLOCK < n16)
			L.rec[tid + q * PBLOCK] = rv[q];
#pragma unroll
	for (int q = 0; q < 2; ++q)
		if (tid + q * PBLOCK < tc.nc)
			L.n[tid + q * PBLOCK] = nn[q];
	if (tid < tc.nt)
		L.u[tid] = u;
	for (uint32_t i = tid + PK_R * PBLOCK; i < n16; i += PBLOCK)
		L.rec[i] = src[i];
	for (uint32_t i = tid + 2 * PBLOCK; i < tc.nc; i += PBLOCK)
		L.n[i] = cntq[tc.cls_base + i];
	for (uint32_t i = tid + PBLOCK; i < tc.nt; i += PBLOCK)
		L.u[i] = uq[tc.tx_base + i];
}

// One map lin -> lout of the tile (src/EM.c:275-291) in three passes;
// with stop the test of src/EM.c:292-297 against lin sets bad.  The term
// pass writes each transcript's terms contiguously in the order its sum
// takes them, so the ordered sums read LDS without an indirection.  With
// member weights (W) a member's share is lam_t w_k: D_c sums lam_t w_k and
// the term is ((lam_t w_k) n_c) / D_c.
template <bool W>
__device__ __forceinline__ void pmap(const PLds &L, const PTile &tc,
                                     const double *lin, double *lout,
                                     bool stop, double smin, double srel,
                                     int &bad)
{
#pragma clang fp contract(off)
	const int tid = threadIdx.x;
	for (uint32_t c = tid; c < tc.nc; c += PBLOCK) {
		uint32_t k = L.coff[c];
		const uint32_t k1 = L.coff[c + 1];
		double acc = 0.0;
		for (; k + 4 <= k1; k += 4) {
			double x0 = lin[L.kmem[k] & 0x3Fu];
			double x1 = lin[L.kmem[k + 1] & 0x3Fu];
			double x2 = lin[L.kmem[k + 2] & 0x3Fu];
			double x3 = lin[L.kmem[k + 3] & 0x3Fu];
			if constexpr (W) {
				x0 *= L.wc[k];
				x1 *= L.wc[k + 1];
				x2 *= L.wc[k + 2];
				x3 *= L.wc[k + 3];
			}
			acc += x0;
			acc += x1;
			acc += x2;
			acc += x3;
		}
		for (; k < k1; ++k) {
			double x = lin[L.kmem[k] & 0x3Fu];
			if constexpr (W)
				x *= L.wc[k];
			acc += x;
		}
		// 1 in place of 0: densities (and weights) are >= 0, so D_c == 0
		// means every member's lam (times its weight) is 0 and the term
		// below is the exact 0 that the reference's skip of the class
		// leaves (src/EM.c:279-280)
		L.D[c] = acc == 0.0 ? 1.0 : acc;
	}
	__syncthreads();
	// (lam_t n_c) / D_c for occurrence j, class c of transcript t
	for (uint32_t j = tid; j < tc.nz; j += PBLOCK) {
		const uint32_t m = L.kocc[j];
		const uint32_t c = m >> 6, t = m & 0x3Fu;
		double lt = lin[t];
		if constexpr (W)
			lt *= L.wo[j];
		L.X[j] = lt * L.n[c] / L.D[c];
	}
	__syncthreads();
	for (uint32_t t = tid; t < tc.nt; t += PBLOCK) {
		uint32_t j = L.toff[t];
		const uint32_t j1 = L.toff[t + 1];
		double acc = 0.0;
		for (; j + 8 <= j1; j += 8) {
			double x[8];
#pragma unroll
			for (int q = 0; q < 8; ++q)
				x[q] = L.X[j + q];
#pragma unroll
			for (int q = 0; q < 8; ++q)
				acc += x[q];
		}
		for (; j < j1; ++j)
			acc += L.X[j];
		const double nw = ((double)L.u[t] + acc) / (double)L.len[t];
		lout[t] = nw;
		if (stop && nw > smin && fabs(nw - lin[t]) / nw > srel)
			bad = 1;
	}
	__syncthreads();
}

// The map in the FP32 bulk form: class sums of the densities rounded to
// FP32 (member order), one FP32 n_c / D_c per class, FP32 terms lam_t
// (n_c / D_c); each transcript's terms are summed in FP64 and the M-step
// (u_t + sum) / len_t, the state and the stop test stay FP64.  The class
// factors and terms use the FP64 slots of D and X as FP32 arrays.
__device__ __forceinline__ void pmap32(const PLds &L, const PTile &tc,
                                       const double *lin, double *lout,
                                       bool stop, double smin, double srel,
                                       int &bad)
{
#pragma clang fp contract(off)
	const int tid = threadIdx.x;
	float *rho = (float *)L.D;
	float *X = (float *)L.X;
	for (uint32_t c = tid; c < tc.nc; c += PBLOCK) {
		uint32_t k = L.coff[c];
		const uint32_t k1 = L.coff[c + 1];
		float acc = 0.0f;
		for (; k + 4 <= k1; k += 4) {
			const float x0 = (float)lin[L.kmem[k] & 0x3Fu];
			const float x1 = (float)lin[L.kmem[k + 1] & 0x3Fu];
			const float x2 = (float)lin[L.kmem[k + 2] & 0x3Fu];
			const float x3 = (float)lin[L.kmem[k + 3] & 0x3Fu];
			acc += x0;
			acc += x1;
			acc += x2;
			acc += x3;
		}
		for (; k < k1; ++k)
			acc += (float)lin[L.kmem[k] & 0x3Fu];
		rho[c] = acc == 0.0f ? 0.0f : (float)L.n[c] / acc;
	}
	__syncthreads();
	for (uint32_t j = tid; j < tc.nz; j += PBLOCK) {
		const uint32_t m = L.kocc[j];
		X[j] = (float)lin[m & 0x3Fu] * rho[m >> 6];
	}
	__syncthreads();
	for (uint32_t t = tid; t < tc.nt; t += PBLOCK) {
		uint32_t j = L.toff[t];
		const uint32_t j1 = L.toff[t + 1];
		double acc = 0.0;
		for (; j + 8 <= j1; j += 8) {
			float x[8];
#pragma unroll
			for (int q = 0; q < 8; ++q)
				x[q] = X[j + q];
#pragma unroll
			for (int q = 0; q < 8; ++q)
				acc += (double)x[q];
		}
		for (; j < j1; ++j)
			acc += (double)X[j];
		const double nw = ((double)L.u[t] + acc) / (double)L.len[t];
		lout[t] = nw;
		if (stop && nw > smin && fabs(nw - lin[t]) / nw > srel)
			bad = 1;
	}
	__syncthreads();
}

// One map in the kernels' precision: FP64 in the reference's order, or
// the FP32 bulk form (F).
template <bool W, bool F>
__device__ __forceinline__ void pmapx(const PLds &L, const PTile &tc,
                                      const double *lin, double *lout,
                                      bool stop, double smin, double srel,
                                      int &bad)
{
	if constexpr (F)
		pmap32(L, tc, lin, lout, stop, smin, srel, bad);
	else
		pmap<W>(L, tc, lin, lout, stop, smin, srel, bad);
}

// Sum over the 64 lanes of wave 0 (xor butterflies: every lane gets the
// same bits).
__device__ __forceinline__ double wave_sum(double x)
{
#pragma clang fp contract(off)
#pragma unroll
	for (int m = 32; m >= 1; m >>= 1)
		x += __shfl_xor(x, m, 64);
	return x;
}

// The tree norms' per-tile partials of r = u1 - u and v = (u2 - u1) - r
// (src/EM.c:334-335) for the next step length, by wave 0 (a tile has at
// most 64 transcripts): the sd form {mean r, centred sum of squares of r,
// mean v, centred sum of squares of v}, merged across tiles by Chan's rule;
// the L2 form {0, sum r^2, 0, sum v^2}.
__device__ __forceinline__ void ptile_parts(const PTile &tc, const double *u,
                                            const double *u1,
                                            const double *u2, int step,
                                            double *out)
{
#pragma clang fp contract(off)
	const uint32_t t = threadIdx.x;
	if (t >= 64)
		return;
	const bool in = t < tc.nt;
	double x = 0.0, y = 0.0;
	if (in) {
		x = u1[t] - u[t];
		y = (u2[t] - u1[t]) - x;
	}
	double a0 = 0.0, b0 = 0.0, qx, qy;
	if (step == 0) {
		a0 = wave_sum(x) / tc.nt;
		b0 = wave_sum(y) / tc.nt;
		qx = in ? (x - a0) * (x - a0) : 0.0;
		qy = in ? (y - b0) * (y - b0) : 0.0;
	} else {
		qx = x * x;
		qy = y * y;
	}
	qx = wave_sum(qx);
	qy = wave_sum(qy);
	if (t == 0) {
		out[0] = a0;
		out[1] = qx;
		out[2] = b0;
		out[3] = qy;
	}
}

// u1 = M(A) -> B, u2 = M(u1) -> C; r = u1 - u, v = (u2 - u1) - r into the
// rows xr, xv (transcript order) when xr != NULL.  FB (monotonicity
// fallback): the second map's stop test is ORed into ctl->stopbad2.  F:
// the FP32 bulk form.
template <bool W, bool FB, bool F>
__global__ __launch_bounds__(PBLOCK) void k_ptile_map12(
    hemk_tiles tl, const double *__restrict__ A, double *__restrict__ B,
    double *__restrict__ C, const uint32_t *__restrict__ uq,
    const double *__restrict__ cntq, double *__restrict__ xr,
    double *__restrict__ xv, double *__restrict__ part, int step,
    hemk_ctl *__restrict__ ctl)
{
#pragma clang fp contract(off)
	extern __shared__ double smem[];
	if (ctl->done)
		return;
	const PTile tc = ptile(tl, blockIdx.x);
	PLds L = pcarve(smem, tl);
	pfields<W>(L, tc);
	pstage<W>(tl, tc, L, uq, cntq);
	for (uint32_t i = threadIdx.x; i < tc.nt; i += PBLOCK)
		L.l0[i] = A[tc.tx_base + i];
	__syncthreads();
	int bad = 0;
	pmapx<W, F>(L, tc, L.l0, L.l1, false, 0.0, 0.0, bad);
	pmapx<W, F>(L, tc, L.l1, L.l2, FB, FB ? ctl->stop_min : 0.0,
	            FB ? ctl->stop_rel : 0.0, bad);
	if (FB) {
		bad = __syncthreads_or(bad);
		if (threadIdx.x == 0 && bad)
			em_mark_bad(&ctl->stopbad2);
	}
	for (uint32_t i = threadIdx.x; i < tc.nt; i += PBLOCK) {
		const double a = L.l0[i], b = L.l1[i], c = L.l2[i];
		B[tc.tx_base + i] = b;
		C[tc.tx_base + i] = c;
		if (xr) {
			const double r = b - a;
			xr[L.tx[i]] = r;
			xv[L.tx[i]] = (c - b) - r;
		}
	}
	if (part)
		ptile_parts(tc, L.l0, L.l1, L.l2, step,
		            part + 4 * (size_t)blockIdx.x);
}

// The end of one SQUAREM iteration and the start of the next in one pass
// over each tile: extrapolate u' from A, B, C with alpha (st.alpha_given 1:
// ctl->alpha; 2: from the two sums part[0], part[1]; src/EM.c:338-352),
// lam = M(u') -> A with the stop test (src/EM.c:354-355, ORed into
// ctl->stopbad), then the next iteration's u1 = M(lam) -> B, u2 = M(u1) -> C
// and, when xr != NULL, its r and v rows (src/EM.c:330-336).  If the stop
// rule ends the loop here, the extra u1, u2, r, v are never read.  F: the
// FP32 bulk form of every map.
template <bool W, bool F>
__global__ __launch_bounds__(PBLOCK) void k_ptile_ex3_map12(
    hemk_tiles tl, double *__restrict__ A, double *__restrict__ B,
    double *__restrict__ C, const uint32_t *__restrict__ uq,
    const double *__restrict__ cntq, const double *__restrict__ part,
    uint32_t T, hemk_step st, double *__restrict__ xr,
    double *__restrict__ xv, double *__restrict__ tpart,
    hemk_ctl *__restrict__ ctl)
{
#pragma clang fp contract(off)
	extern __shared__ double smem[];
	if (ctl->done)
		return;
	const double alpha = st.alpha_given == 2
	                         ? em_step_length(part[0], part[1], T, st)
	                         : ctl->alpha;
	const double smin = ctl->stop_min, srel = ctl->stop_rel;
	const PTile tc = ptile(tl, blockIdx.x);
	PLds L = pcarve(smem, tl);
	pfields<W>(L, tc);
	pstage<W>(tl, tc, L, uq, cntq);
	for (uint32_t i = threadIdx.x; i < tc.nt; i += PBLOCK) {
		const size_t g = tc.tx_base + i;
		L.l0[i] = em_extrapolate(A[g], B[g], C[g], alpha, st);
	}
	__syncthreads();
	int bad = 0;
	pmapx<W, F>(L, tc, L.l0, L.l1, true, smin, srel, bad);
	for (uint32_t i = threadIdx.x; i < tc.nt; i += PBLOCK)
		A[tc.tx_base + i] = L.l1[i];
	bad = __syncthreads_or(bad);
	if (threadIdx.x == 0) {
		if (bad)
			em_mark_bad(&ctl->stopbad);
		if (blockIdx.x == 0 && st.alpha_given == 2)
			ctl->alpha = alpha;
	}
	pmapx<W, F>(L, tc, L.l1, L.l2, false, 0.0, 0.0, bad);
	pmapx<W, F>(L, tc, L.l2, L.l3, false, 0.0, 0.0, bad);
	for (uint32_t i = threadIdx.x; i < tc.nt; i += PBLOCK) {
		const double a = L.l1[i], b = L.l2[i], c = L.l3[i];
		B[tc.tx_base + i] = b;
		C[tc.tx_base + i] = c;
		if (xr) {
			const double r = b - a;
			xr[L.tx[i]] = r;
			xv[L.tx[i]] = (c - b) - r;
		}
	}
	if (tpart)
		ptile_parts(tc, L.l1, L.l2, L.l3, st.step,
		            tpart + 4 * (size_t)blockIdx.x);
}

// Class sums of lin in member order (weights applied), stored as they are
// (a zero stays zero), for the fallback's log-likelihood terms.
template <bool W>
__device__ __forceinline__ void pclass_sums(const PLds &L, const PTile &tc,
                                            const double *lin, double *D)
{
#pragma clang fp contract(off)
	for (uint32_t c = threadIdx.x; c < tc.nc; c += PBLOCK) {
		double acc = 0.0;
		for (uint32_t k = L.coff[c]; k < L.coff[c + 1]; ++k) {
			double x = lin[L.kmem[k] & 0x3Fu];
			if constexpr (W)
				x *= L.wc[k];
			acc += x;
		}
		D[c] = acc;
	}
}

// Sum over the block's 256 threads in a fixed order (wave butterflies,
// then the four waves in turn); the result is valid in thread 0.
__device__ __forceinline__ double pblock_sum(double x, double *red)
{
#pragma unroll
	for (int m = 32; m >= 1; m >>= 1)
		x += __shfl_xor(x, m, 64);
	if (threadIdx.x % 64 == 0)
		red[threadIdx.x / 64] = x;
	__syncthreads();
	double t = 0.0;
	if (threadIdx.x == 0)
		t = ((red[0] + red[1]) + red[2]) + red[3];
	__syncthreads();
	return t;
}

// Monotonicity fallback, first half (no counterpart in the reference): the
// extrapolation and stabilising map of k_ptile_ex3_map12 without the next
// iteration's maps, plus this tile's terms of ell(lam) - ell(u2): per class
// n_c log1p((D_c(lam) - D_c(u2)) / D_c(u2)) (a zero sum at either point:
// that point's n_c log D_c as it stands), per transcript u_t log1p((lam_t -
// u2_t) / u2_t) and -len_t (lam_t - u2_t); dpart[2 tile] their sum and
// dpart[2 tile + 1] the sum of their magnitudes.  D2 (class sums at u2)
// lives d2 doubles into the dynamic LDS.
template <bool W>
__global__ __launch_bounds__(PBLOCK) void k_ptile_ex3_fb(
    hemk_tiles tl, double *__restrict__ A, const double *__restrict__ B,
    const double *__restrict__ C, const uint32_t *__restrict__ uq,
    const double *__restrict__ cntq, const double *__restrict__ part,
    uint32_t T, hemk_step st, double *__restrict__ dpart, size_t d2,
    hemk_ctl *__restrict__ ctl)
{
#pragma clang fp contract(off)
	extern __shared__ double smem[];
	__shared__ double red[4];
	if (ctl->done)
		return;
	const double alpha = st.alpha_given == 2
	                         ? em_step_length(part[0], part[1], T, st)
	                         : ctl->alpha;
	const double smin = ctl->stop_min, srel = ctl->stop_rel;
	const PTile tc = ptile(tl, blockIdx.x);
	PLds L = pcarve(smem, tl);
	double *D2 = smem + d2;
	pfields<W>(L, tc);
	pstage<W>(tl, tc, L, uq, cntq);
	for (uint32_t i = threadIdx.x; i < tc.nt; i += PBLOCK) {
		const size_t g = tc.tx_base + i;
		const double c = C[g];
		L.l2[i] = c;
		L.l0[i] = em_extrapolate(A[g], B[g], c, alpha, st);
	}
	__syncthreads();
	pclass_sums<W>(L, tc, L.l2, D2);
	int bad = 0;
	pmap<W>(L, tc, L.l0, L.l1, true, smin, srel, bad);
	for (uint32_t i = threadIdx.x; i < tc.nt; i += PBLOCK)
		A[tc.tx_base + i] = L.l1[i];
	bad = __syncthreads_or(bad);
	if (threadIdx.x == 0) {
		if (bad)
			em_mark_bad(&ctl->stopbad);
		if (blockIdx.x == 0 && st.alpha_given == 2)
			ctl->alpha = alpha;
	}
	pclass_sums<W>(L, tc, L.l1, L.D);
	__syncthreads();
	double d = 0.0, m = 0.0;
	for (uint32_t c = threadIdx.x; c < tc.nc; c += PBLOCK) {
		const double n = L.n[c], dl = L.D[c], dm = D2[c];
		double x = 0.0;
		if (n == 0.0)
			continue;
		if (dl > 0 && dm > 0)
			x = n * log1p((dl - dm) / dm);
		else if (dl > 0)
			x = n * log(dl);
		else if (dm > 0)
			x = -(n * log(dm));
		d += x;
		m += fabs(x);
	}
	for (uint32_t t = threadIdx.x; t < tc.nt; t += PBLOCK) {
		const double lam = L.l1[t], mu = L.l2[t];
		if (L.u[t] != 0) {
			const double x = (double)L.u[t] * log1p((lam - mu) / mu);
			d += x;
			m += fabs(x);
		}
		const double y = (double)L.len[t] * (lam - mu);
		d -= y;
		m += fabs(y);
	}
	d = pblock_sum(d, red);
	m = pblock_sum(m, red);
	if (threadIdx.x == 0) {
		dpart[2 * (size_t)blockIdx.x] = d;
		dpart[2 * (size_t)blockIdx.x + 1] = m;
	}
}

// Monotonicity fallback, the decision: totals of the tiles' differences and
// magnitudes (fixed order); the iteration keeps u2 when the difference is
// below -tol times the magnitudes, with the second map's stop test.
__global__ __launch_bounds__(256) void k_fb_decide(
    const double *__restrict__ dpart, uint32_t n_tiles, double tol,
    hemk_ctl *__restrict__ ctl)
{
	__shared__ double red[4];
	if (ctl->done)
		return;
	double d = 0.0, m = 0.0;
	for (uint32_t i = threadIdx.x; i < n_tiles; i += 256) {
		d += dpart[2 * (size_t)i];
		m += dpart[2 * (size_t)i + 1];
	}
	d = pblock_sum(d, red);
	m = pblock_sum(m, red);
	if (threadIdx.x == 0) {
		const uint32_t fb = d < -tol * m ? 1u : 0u;
		ctl->fb = fb;
		if (fb) {
			ctl->fallbacks += 1;
			ctl->stopbad = ctl->stopbad2;
		}
		ctl->stopbad2 = 0;
	}
}

// A = C (u2) when the decision kept u2.
__global__ __launch_bounds__(256) void k_fb_apply(
    uint32_t T, const double *__restrict__ C, double *__restrict__ A,
    const hemk_ctl *__restrict__ ctl)
{
	if (ctl->done || !ctl->fb)
		return;
	for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < T;
	     i += gridDim.x * 256)
		A[i] = C[i];
}

// Fixed-order sums over the workgroup of N values per thread: an xor
// butterfly within each wave (every lane ends with the same bits, since
// each level adds the same two operands in either order), then the waves'
// sums in wave order.  Every thread returns the totals.
#define TA_THREADS 1024
#define TA_WAVES (TA_THREADS / 64)
#define TA_PER 5
template <int N>
__device__ __forceinline__ void ta_block_sums(double (&x)[N],
                                              double (*sh)[TA_WAVES])
{
#pragma clang fp contract(off)
	const uint32_t w = threadIdx.x / 64;
#pragma unroll
	for (int k = 0; k < N; ++k)
		x[k] = wave_sum(x[k]);
	if (threadIdx.x % 64 == 0) {
#pragma unroll
		for (int k = 0; k < N; ++k)
			sh[k][w] = x[k];
	}
	__syncthreads();
#pragma unroll
	for (int k = 0; k < N; ++k) {
		double s = sh[k][0];
		for (int j = 1; j < TA_WAVES; ++j)
			s += sh[k][j];
		x[k] = s;
	}
	__syncthreads();
}

// One round of the fold's loads: tile base + threadIdx.x + TA_THREADS k
// for k < TA_PER, its transcript count and four partials (zeros past the
// last tile).
#define TA_STEP (TA_THREADS * TA_PER)
__device__ __forceinline__ void ta_load(const hemk_tiles &tl,
                                        const double *__restrict__ part,
                                        uint32_t base, double (&n)[TA_PER],
                                        double (&p)[4][TA_PER])
{
#pragma unroll
	for (int k = 0; k < TA_PER; ++k) {
		const uint32_t i = base + threadIdx.x + TA_THREADS * k;
		const bool in = i < tl.n_tiles;
		n[k] = in ? (double)tl.desc[(size_t)HEMK_TILE_DESC * i + 1] : 0.0;
#pragma unroll
		for (int c = 0; c < 4; ++c)
			p[c][k] = in ? part[4 * (size_t)i + c] : 0.0;
	}
}

// Tree norms on the tiles: fold the per-tile partials into alpha.  The L2
// form adds the tiles' sums of squares; the sd form first finds the global
// means m = sum n_i m_i / sum n_i, then the centred sums of squares as
// sum_i M2_i + n_i (m_i - m)^2 (the tiles' moments moved to the global
// mean), so no division sits on a dependent chain.  Thread i takes tiles i,
// i + 1024, ... in rounds of TA_STEP tiles; the first round's loads are
// issued before the control block is read, and the last round loaded stays
// in registers for the second sum, which takes it first.  With end the
// kernel first ends the iteration whose extrapolated map wrote the
// partials (as k_iter_end: k, the loop condition of src/EM.c:328 from the
// stop test) and folds only if the loop goes on.
__global__ __launch_bounds__(TA_THREADS) void k_ptile_tree_alpha(
    hemk_tiles tl, const double *__restrict__ part, uint32_t T,
    hemk_step st, int end, hemk_ctl *__restrict__ ctl)
{
#pragma clang fp contract(off)
	__shared__ double sh[3][TA_WAVES];
	__shared__ uint32_t fin;
	const uint32_t tid = threadIdx.x, nt = tl.n_tiles;
	double n[TA_PER], p[4][TA_PER];
	ta_load(tl, part, 0, n, p);
	if (ctl->done)
		return;
	if (end) {
		if (tid == 0) {
			const uint32_t stop = ctl->stopbad ? 0u : 1u;
			const uint32_t k = ctl->k + 1;
			ctl->stopbad = 0;
			em_set_k(ctl, k);
			ctl->maps += 3;
			fin = !(k < ctl->min_iter ||
			        (k < ctl->max_iter && !stop));
			if (fin)
				em_set_done(ctl, stop);
		}
		__syncthreads();
		if (fin)
			return;
	}
	uint32_t last = 0;
	if (st.step != 0) {
		double q[2] = {0.0, 0.0};
		for (uint32_t base = 0;;) {
#pragma unroll
			for (int k = 0; k < TA_PER; ++k) {
				q[0] += p[1][k];
				q[1] += p[3][k];
			}
			base += TA_STEP;
			if (base >= nt)
				break;
			ta_load(tl, part, base, n, p);
		}
		ta_block_sums(q, sh);
		if (tid == 0)
			ctl->alpha = em_step_length(q[0], q[1], T, st);
		return;
	}
	double s[3] = {0.0, 0.0, 0.0};   // sum n_i, sum n_i m_r,i, sum n_i m_v,i
	for (uint32_t base = 0;;) {
#pragma unroll
		for (int k = 0; k < TA_PER; ++k) {
			s[0] += n[k];
			s[1] += n[k] * p[0][k];
			s[2] += n[k] * p[2][k];
		}
		last = base;
		base += TA_STEP;
		if (base >= nt)
			break;
		ta_load(tl, part, base, n, p);
	}
	ta_block_sums(s, sh);
	const double mr = s[1] / s[0], mv = s[2] / s[0];
	// the round in registers (the last) first, then rounds 0 .. last - 1
	double q[2] = {0.0, 0.0};
	for (uint32_t base = 0;; base += TA_STEP) {
#pragma unroll
		for (int k = 0; k < TA_PER; ++k) {
			const double dr = p[0][k] - mr, dv = p[2][k] - mv;
			q[0] += p[1][k] + n[k] * dr * dr;
			q[1] += p[3][k] + n[k] * dv * dv;
		}
		if (base >= last)
			break;
		ta_load(tl, part, base, n, p);
	}
	ta_block_sums(q, sh);
	if (tid == 0)
		ctl->alpha = em_step_length(q[0], q[1], T, st);
}

// Tree norms: alpha from the block partials of the norm passes.
__global__ __launch_bounds__(256) void k_ptile_alpha(
    const double *__restrict__ part, uint32_t nparts, uint32_t T,
    hemk_step st, hemk_ctl *__restrict__ ctl)
{
	__shared__ double sa[4], sb[4];
	if (ctl->done)
		return;
	double a = 0.0, b = 0.0;
	for (uint32_t i = threadIdx.x; i < nparts; i += 256) {
		a += part[2 * i];
		b += part[2 * i + 1];
	}
#pragma unroll
	for (int m = 32; m >= 1; m >>= 1) {
		a += __shfl_xor(a, m, 64);
		b += __shfl_xor(b, m, 64);
	}
	if (threadIdx.x % 64 == 0) {
		sa[threadIdx.x / 64] = a;
		sb[threadIdx.x / 64] = b;
	}
	__syncthreads();
	if (threadIdx.x == 0) {
		a = sa[0];
		b = sb[0];
		for (int w = 1; w < 4; ++w) {
			a += sa[w];
			b += sb[w];
		}
		ctl->alpha = em_step_length(a, b, T, st);
	}
}

// ---- derivatives of the log-likelihood over the tiles ---------------------
// L = sum_c n_c log D_c + sum_t u_t log x_t, D_c = sum_{t in c} x_t (the
// objective of grad_kernels.hip).  The tile's record and its classes sit in
// LDS, so the class pass and the transcript pass read D_c, n_c / D_c and
// n_c without the scattered global loads of the two-pass form.

// The record of tile tc into LDS (no counts: the derivative kernels read
// them by class id).
__device__ __forceinline__ void pstage_rec(const hemk_tiles &tl,
                                           const PTile &tc, const PLds &L)
{
	const uint32_t n16 = rec_bytes<false>(tc.nt, tc.nc, tc.nz) / 16;
	const uint4 *src = (const uint4 *)tl.rec + tc.rec;
	uint4 rv[PK_R];
#pragma unroll
	for (int q = 0; q < PK_R; ++q)
		rv[q] = threadIdx.x + q * PBLOCK < n16 ? src[threadIdx.x + q * PBLOCK]
		                                       : make_uint4(0, 0, 0, 0);
#pragma unroll
	for (int q = 0; q < PK_R; ++q)
		if (threadIdx.x + q * PBLOCK < n16)
			L.rec[threadIdx.x + q * PBLOCK] = rv[q];
	for (uint32_t i = threadIdx.x + PK_R * PBLOCK; i < n16; i += PBLOCK)
		L.rec[i] = src[i];
}

// Workgroup sum (PBLOCK threads) in a fixed order; thread 0 has it.
__device__ __forceinline__ double pblock_total(double a, double *sh)
{
#pragma clang fp contract(off)
	a = wave_sum(a);
	if (threadIdx.x % 64 == 0)
		sh[threadIdx.x / 64] = a;
	__syncthreads();
	double t = sh[0];
	for (int w = 1; w < PBLOCK / 64; ++w)
		t += sh[w];
	__syncthreads();
	return t;
}

// The adjoint's passes over one tile: D_c (members in class order, in TX),
// rho_c = n_c / D_c, kappa_c = rho_c / D_c (0 when n_c or D_c is 0, kept
// by class tile position for the HVP) and n_c log D_c; then per transcript
// E_t = x_t sum rho_c + u_t and the class part of the Fisher diagonal
// fd_t = -sum n_c q (1 - q), q = x_t / D_c; tile partials of
// sum n_c log D_c, N = sum E_t and sum u_t log x_t.
template <typename TX>
__global__ __launch_bounds__(PBLOCK) void k_ptile_grad(
    hemk_tiles tl, const double *__restrict__ cntq,
    const TX *__restrict__ x, const uint32_t *__restrict__ unique,
    double *__restrict__ kappa_p, double *__restrict__ E,
    double *__restrict__ fd, double *__restrict__ part)
{
	extern __shared__ double smem[];
	__shared__ double sh[PBLOCK / 64];
	const uint32_t tid = threadIdx.x, tile = blockIdx.x;
	const PTile tc = ptile(tl, tile);
	PLds L = pcarve(smem, tl);
	pfields<false>(L, tc);
	pstage_rec(tl, tc, L);
	__syncthreads();
	double *xl = L.l0;     // x_t of the tile's transcripts
	for (uint32_t i = tid; i < tc.nt; i += PBLOCK) {
		const uint32_t id = L.tx[i];
		xl[i] = (double)x[id];
		L.u[i] = unique[id];
	}
	for (uint32_t c = tid; c < tc.nc; c += PBLOCK)
		L.n[c] = cntq[tc.cls_base + c];
	__syncthreads();
	double Lc = 0.0;
	for (uint32_t c = tid; c < tc.nc; c += PBLOCK) {
		uint32_t k = L.coff[c];
		const uint32_t k1 = L.coff[c + 1];
		TX acc = 0;
		for (; k + 4 <= k1; k += 4) {
			const TX x0 = (TX)xl[L.kmem[k] & 0x3Fu];
			const TX x1 = (TX)xl[L.kmem[k + 1] & 0x3Fu];
			const TX x2 = (TX)xl[L.kmem[k + 2] & 0x3Fu];
			const TX x3 = (TX)xl[L.kmem[k + 3] & 0x3Fu];
			acc += x0;
			acc += x1;
			acc += x2;
			acc += x3;
		}
		for (; k < k1; ++k)
			acc += (TX)xl[L.kmem[k] & 0x3Fu];
		const double D = (double)acc, n = L.n[c];
		double r = 0.0, q = 0.0;
		if (n != 0 && D != 0) {
			r = n / D;
			q = r / D;
			Lc += n * log(D);
		}
		L.D[c] = D;
		L.n[c] = r;    // rho_c from here on (each thread its own c)
		kappa_p[tc.cls_base + c] = q;
	}
	__syncthreads();
	// four lanes per transcript, each taking every fourth occurrence, then
	// combined across the four (n_c q (1 - q) = rho_c x_t (1 - x_t / D_c))
	double Nn = 0.0, Lt = 0.0;
	const uint32_t sub = tid & 3u;
	for (uint32_t t0 = 0; t0 < tc.nt; t0 += PBLOCK / 4) {
		const uint32_t t = t0 + tid / 4;
		const bool in = t < tc.nt;
		const double xt = in ? xl[t] : 0.0;
		double sr = 0.0, h = 0.0;
		if (in) {
			const uint32_t j1 = L.toff[t + 1];
			uint32_t j = L.toff[t] + sub;
			for (; j + 4 < j1; j += 8) {
				const uint32_t c0 = L.kocc[j] >> 6;
				const uint32_t c1 = L.kocc[j + 4] >> 6;
				const double r0 = L.n[c0], d0 = L.D[c0];
				const double r1 = L.n[c1], d1 = L.D[c1];
				sr += r0;
				sr += r1;
				if (r0 != 0)
					h += r0 * xt * (1 - xt / d0);
				if (r1 != 0)
					h += r1 * xt * (1 - xt / d1);
			}
			for (; j < j1; j += 4) {
				const uint32_t c = L.kocc[j] >> 6;
				const double r = L.n[c];
				sr += r;
				if (r != 0)
					h += r * xt * (1 - xt / L.D[c]);
			}
		}
		sr += __shfl_xor(sr, 1, 64);
		h += __shfl_xor(h, 1, 64);
		sr += __shfl_xor(sr, 2, 64);
		h += __shfl_xor(h, 2, 64);
		if (in && sub == 0) {
			const uint32_t u = L.u[t], id = L.tx[t];
			const double e = xt * sr + (double)u;
			E[id] = e;
			fd[id] = -h;
			Nn += e;
			if (u)
				Lt += (double)u * log(xt);
		}
	}
	Lc = pblock_total(Lc, sh);
	Nn = pblock_total(Nn, sh);
	Lt = pblock_total(Lt, sh);
	if (tid == 0) {
		part[tile] = Lc;
		part[tl.n_tiles + tile] = Nn;
		part[2 * (size_t)tl.n_tiles + tile] = Lt;
	}
}

// H v over one tile: kd_c = kappa_c sum_{s in c} x_s v_s (members in class
// order), hv_t = v_t a_t - x_t sum_{c ni t} kd_c + N (pi . v) pi_t.
__global__ __launch_bounds__(PBLOCK) void k_ptile_hvp(
    hemk_tiles tl, const double *__restrict__ x, const double *__restrict__ v,
    const double *__restrict__ kappa_p, const double *__restrict__ a,
    const double *__restrict__ pi, const double *__restrict__ pv_part,
    const double *__restrict__ scal, double *__restrict__ hv)
{
	extern __shared__ double smem[];
	__shared__ double sh[PBLOCK / 64];
	__shared__ double pvs;
	const uint32_t tid = threadIdx.x;
	const PTile tc = ptile(tl, blockIdx.x);
	PLds L = pcarve(smem, tl);
	pfields<false>(L, tc);
	pstage_rec(tl, tc, L);
	double pv = 0.0;
	for (uint32_t i = tid; i < HEMK_NORM_BLOCKS; i += PBLOCK)
		pv += pv_part[i];
	pv = pblock_total(pv, sh);   // also orders the record's stores
	if (tid == 0)
		pvs = pv;
	double *xv = L.l0;
	for (uint32_t i = tid; i < tc.nt; i += PBLOCK) {
		const uint32_t id = L.tx[i];
		xv[i] = x[id] * v[id];
	}
	for (uint32_t c = tid; c < tc.nc; c += PBLOCK)
		L.n[c] = kappa_p[tc.cls_base + c];
	__syncthreads();
	for (uint32_t c = tid; c < tc.nc; c += PBLOCK) {
		uint32_t k = L.coff[c];
		const uint32_t k1 = L.coff[c + 1];
		double acc = 0.0;
		for (; k + 4 <= k1; k += 4) {
			const double x0 = xv[L.kmem[k] & 0x3Fu];
			const double x1 = xv[L.kmem[k + 1] & 0x3Fu];
			const double x2 = xv[L.kmem[k + 2] & 0x3Fu];
			const double x3 = xv[L.kmem[k + 3] & 0x3Fu];
			acc += x0;
			acc += x1;
			acc += x2;
			acc += x3;
		}
		for (; k < k1; ++k)
			acc += xv[L.kmem[k] & 0x3Fu];
		L.D[c] = L.n[c] * acc;
	}
	__syncthreads();
	// four lanes per transcript, combined across the four
	const double Npv = scal[1] * pvs;
	const uint32_t sub = tid & 3u;
	for (uint32_t t0 = 0; t0 < tc.nt; t0 += PBLOCK / 4) {
		const uint32_t t = t0 + tid / 4;
		const bool in = t < tc.nt;
		double sk = 0.0;
		if (in) {
			const uint32_t j1 = L.toff[t + 1];
			uint32_t j = L.toff[t] + sub;
			for (; j + 4 < j1; j += 8) {
				const double k0 = L.D[L.kocc[j] >> 6];
				const double k1 = L.D[L.kocc[j + 4] >> 6];
				sk += k0;
				sk += k1;
			}
			for (; j < j1; j += 4)
				sk += L.D[L.kocc[j] >> 6];
		}
		sk += __shfl_xor(sk, 1, 64);
		sk += __shfl_xor(sk, 2, 64);
		if (in && sub == 0) {
			const uint32_t id = L.tx[t];
			hv[id] = v[id] * a[id] - x[id] * sk + Npv * pi[id];
		}
	}
}

// out[t] = A[txq[t]]: the state in transcript order.
__global__ __launch_bounds__(256) void k_ptile_rows(
    uint32_t T, const uint32_t *__restrict__ txq, const double *__restrict__ A,
    double *__restrict__ out)
{
	for (uint32_t t = blockIdx.x * 256 + threadIdx.x; t < T;
	     t += gridDim.x * 256)
		out[t] = A[txq[t]];
}

} // namespace

extern "C" {

size_t hemk_ptile_lds(const hemk_tiles *t)
{
	const size_t mt = t->max_tx, mc = t->max_cls, mz = t->max_nnz;
	return (size_t)t->max_rec + 8 * (4 * mt + 2 * mc + mz) + 4 * mt;
}

int hemk_ptile_map12(const hemk_tiles *t, const double *A, double *B,
                     double *C, const uint32_t *uq, const double *cntq,
                     double *xr, double *xv, double *part, int step,
                     int f32, const hemk_ctl *ctl, void *s)
{
	if (!t->n_tiles || t->max_tx > 64 || (f32 && t->weighted))
		return (int)hipErrorInvalidValue;
	// the kernel writes no control word without FB
	hemk_ctl *c = (hemk_ctl *)ctl;
	const size_t lds = hemk_ptile_lds(t);
	hipStream_t st = (hipStream_t)s;
	if (f32)
		k_ptile_map12<false, false, true><<<t->n_tiles, PBLOCK, lds, st>>>(
		    *t, A, B, C, uq, cntq, xr, xv, part, step, c);
	else if (t->weighted)
		k_ptile_map12<true, false, false><<<t->n_tiles, PBLOCK, lds, st>>>(
		    *t, A, B, C, uq, cntq, xr, xv, part, step, c);
	else
		k_ptile_map12<false, false, false><<<t->n_tiles, PBLOCK, lds, st>>>(
		    *t, A, B, C, uq, cntq, xr, xv, part, step, c);
	return (int)hipGetLastError();
}

int hemk_ptile_map12_fb(const hemk_tiles *t, const double *A, double *B,
                        double *C, const uint32_t *uq, const double *cntq,
                        double *xr, double *xv, double *part, int step,
                        hemk_ctl *ctl, void *s)
{
	if (!t->n_tiles || t->max_tx > 64)
		return (int)hipErrorInvalidValue;
	const size_t lds = hemk_ptile_lds(t);
	hipStream_t st = (hipStream_t)s;
	if (t->weighted)
		k_ptile_map12<true, true, false><<<t->n_tiles, PBLOCK, lds, st>>>(
		    *t, A, B, C, uq, cntq, xr, xv, part, step, ctl);
	else
		k_ptile_map12<false, true, false><<<t->n_tiles, PBLOCK, lds, st>>>(
		    *t, A, B, C, uq, cntq, xr, xv, part, step, ctl);
	return (int)hipGetLastError();
}

int hemk_ptile_grad(const hemk_tiles *t, const double *cntq, int fp32,
                    const double *x, const float *x32,
                    const uint32_t *unique, double *kappa_p, double *E,
                    double *fd, double *part, void *s)
{
	if (!t->n_tiles || t->max_tx > 64 || t->weighted)
		return (int)hipErrorInvalidValue;
	const size_t lds = hemk_ptile_lds(t);
	hipStream_t st = (hipStream_t)s;
	if (fp32)
		k_ptile_grad<float><<<t->n_tiles, PBLOCK, lds, st>>>(
		    *t, cntq, x32, unique, kappa_p, E, fd, part);
	else
		k_ptile_grad<double><<<t->n_tiles, PBLOCK, lds, st>>>(
		    *t, cntq, x, unique, kappa_p, E, fd, part);
	return (int)hipGetLastError();
}

int hemk_ptile_hvp(const hemk_tiles *t, const double *x, const double *v,
                   const double *kappa_p, const double *a, const double *pi,
                   const double *pv_part, const double *scal, double *hv,
                   void *s)
{
	if (!t->n_tiles || t->max_tx > 64 || t->weighted)
		return (int)hipErrorInvalidValue;
	k_ptile_hvp<<<t->n_tiles, PBLOCK, hemk_ptile_lds(t), (hipStream_t)s>>>(
	    *t, x, v, kappa_p, a, pi, pv_part, scal, hv);
	return (int)hipGetLastError();
}

int hemk_ptile_tree_alpha(const hemk_tiles *t, const double *part,
                          uint32_t T, hemk_step st, int end, hemk_ctl *ctl,
                          void *s)
{
	k_ptile_tree_alpha<<<1, TA_THREADS, 0, (hipStream_t)s>>>(*t, part, T,
	                                                        st, end, ctl);
	return (int)hipGetLastError();
}

// LDS of the fallback's extrapolated map: the tile kernels' layout, then
// the class sums at u2 (8-byte aligned).
static size_t ptile_fb_lds(const hemk_tiles *t)
{
	return (hemk_ptile_lds(t) + 7) / 8 * 8 + 8 * (size_t)t->max_cls;
}

int hemk_ptile_ex3_fb(const hemk_tiles *t, double *A, const double *B,
                      const double *C, const uint32_t *uq,
                      const double *cntq, const double *part, uint32_t T,
                      hemk_step st, double *dpart, hemk_ctl *ctl, void *s)
{
	if (!t->n_tiles || t->max_tx > 256)
		return (int)hipErrorInvalidValue;
	const size_t d2 = (hemk_ptile_lds(t) + 7) / 8;   // in doubles
	if (t->weighted)
		k_ptile_ex3_fb<true><<<t->n_tiles, PBLOCK, ptile_fb_lds(t),
		                       (hipStream_t)s>>>(*t, A, B, C, uq, cntq,
		                                         part, T, st, dpart, d2,
		                                         ctl);
	else
		k_ptile_ex3_fb<false><<<t->n_tiles, PBLOCK, ptile_fb_lds(t),
		                        (hipStream_t)s>>>(*t, A, B, C, uq, cntq,
		                                          part, T, st, dpart, d2,
		                                          ctl);
	return (int)hipGetLastError();
}

int hemk_fb_decide(const double *dpart, uint32_t n_tiles, double tol,
                   hemk_ctl *ctl, void *s)
{
	k_fb_decide<<<1, 256, 0, (hipStream_t)s>>>(dpart, n_tiles, tol, ctl);
	return (int)hipGetLastError();
}

int hemk_fb_apply(uint32_t T, const double *C, double *A,
                  const hemk_ctl *ctl, void *s)
{
	const uint32_t g = (T + 255) / 256 < 2048 ? (T + 255) / 256 : 2048;
	k_fb_apply<<<g ? g : 1, 256, 0, (hipStream_t)s>>>(T, C, A, ctl);
	return (int)hipGetLastError();
}

int hemk_ptile_ex3_map12(const hemk_tiles *t, double *A, double *B,
                         double *C, const uint32_t *uq, const double *cntq,
                         const double *part, uint32_t T, hemk_step st,
                         double *xr, double *xv, double *tpart, int f32,
                         hemk_ctl *ctl, void *s)
{
	if (!t->n_tiles || t->max_tx > 64 || (f32 && t->weighted))
		return (int)hipErrorInvalidValue;
	const size_t lds = hemk_ptile_lds(t);
	hipStream_t sm = (hipStream_t)s;
	if (f32)
		k_ptile_ex3_map12<false, true><<<t->n_tiles, PBLOCK, lds, sm>>>(
		    *t, A, B, C, uq, cntq, part, T, st, xr, xv, tpart, ctl);
	else if (t->weighted)
		k_ptile_ex3_map12<true, false><<<t->n_tiles, PBLOCK, lds, sm>>>(
		    *t, A, B, C, uq, cntq, part, T, st, xr, xv, tpart, ctl);
	else
		k_ptile_ex3_map12<false, false><<<t->n_tiles, PBLOCK, lds, sm>>>(
		    *t, A, B, C, uq, cntq, part, T, st, xr, xv, tpart, ctl);
	return (int)hipGetLastError();
}

int hemk_ptile_alpha(const double *part, uint32_t nparts, uint32_t T,
                     hemk_step st, hemk_ctl *ctl, void *s)
{
	k_ptile_alpha<<<1, 256, 0, (hipStream_t)s>>>(part, nparts, T, st, ctl);
	return (int)hipGetLastError();
}

int hemk_ptile_rows(uint32_t T, const uint32_t *txq, const double *A,
                    double *out, void *s)
{
	k_ptile_rows<<<(T + 255) / 256 < 2048 ? (T + 255) / 256 : 2048, 256, 0,
	               (hipStream_t)s>>>(T, txq, A, out);
	return (int)hipGetLastError();
}

} // extern "C"
