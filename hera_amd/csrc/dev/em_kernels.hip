// em_kernels.hip -- gfx950 kernels of one SQUAREM EM instance.
//
// One EM map (src/EM.c:271-302) is two deterministic passes with no
// atomics:
//   k_denom   CSR pass over classes: rho_c = n_c / D_c, D_c = sum_t lam_t,
//             a group of G lanes per class, butterfly reduction;
//   k_gather  CSC pass over transcripts: s_t = sum_{c ni t} rho_c, then the
//             M-step lam'_t = (u_t + lam_t s_t) / len_t in FP64, with the
//             stop test of src/EM.c:292-297 folded in for the third map.
// The SQUAREM step (src/EM.c:333-352) is three grid-fixed kernels: two
// passes of the sample-sd norms (src/hash_align.c:150-164) and the
// extrapolation.  Every reduction has a fixed shape, so results are bitwise
// reproducible run to run.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "em_step.h"
#include "hem_dev.h"

#define BLOCK 256
#define WAVE 64
#define NWAVE (BLOCK / WAVE)

namespace {

template <int G, typename T>
__device__ __forceinline__ T group_sum(T v)
{
#pragma unroll
	for (int m = G / 2; m >= 1; m >>= 1)
		v += __shfl_xor(v, m, G);
	return v;
}

// Sum of two doubles over the block; every thread gets the same bits.
__device__ __forceinline__ void block_sum2(double &a, double &b)
{
	__shared__ double sa[NWAVE], sb[NWAVE];
	a = group_sum<WAVE>(a);
	b = group_sum<WAVE>(b);
	int w = threadIdx.x / WAVE;
	if ((threadIdx.x & (WAVE - 1)) == 0) {
		sa[w] = a;
		sb[w] = b;
	}
	__syncthreads();
	a = sa[0];
	b = sb[0];
#pragma unroll
	for (int i = 1; i < NWAVE; ++i) {
		a += sa[i];
		b += sb[i];
	}
	__syncthreads();
}

// Every block reduces the same HEMK_NORM_BLOCKS partial pairs in the same
// order, so every block holds identical totals.
__device__ __forceinline__ void reduce_parts(const double *__restrict__ part,
                                             double &a, double &b)
{
	a = 0;
	b = 0;
	for (int i = threadIdx.x; i < HEMK_NORM_BLOCKS; i += BLOCK) {
		a += part[2 * i];
		b += part[2 * i + 1];
	}
	block_sum2(a, b);
}

// Sum over a group's lanes in lane order, chunk by chunk: x holds element
// base+lane of a row; every lane ends with the same left-to-right sum, so
// the result is bitwise the sequential loop of the reference while the
// loads that produce x run in parallel.
template <int G>
__device__ __forceinline__ void ordered_add(double &acc, double x,
                                            uint32_t n_here)
{
	for (uint32_t j = 0; j < n_here; ++j)
		acc += __shfl(x, (int)j, G);
}

// FP64 class pass, order-faithful: D_c is the member-order sequential sum
// of src/EM.c:277-278.  nd[c] = {n_c, D_c}; only D_c is written.  With
// member weights w (CSR order) the summand is lam * w, so w == 1 gives the
// unweighted sum bit for bit.
template <int G>
__global__ __launch_bounds__(BLOCK) void k_denom_exact(
    const uint32_t *__restrict__ off, const uint32_t *__restrict__ ref,
    const double *__restrict__ w, const double *__restrict__ lam,
    double2 *__restrict__ nd, uint32_t n_cls,
    const hemk_ctl *__restrict__ ctl)
{
	if (ctl->done)
		return;
	const uint32_t c = blockIdx.x * (BLOCK / G) + threadIdx.x / G;
	const uint32_t lane = threadIdx.x % G;
	uint32_t beg = 0, end = 0;
	if (c < n_cls) {
		beg = off[c];
		end = off[c + 1];
	}
	double acc = 0.0;
	for (uint32_t base = beg; base < end; base += G) {
		const uint32_t k = base + lane;
		double x = 0.0;
		if (k < end) {
			x = lam[ref[k]];
			if (w)
				x *= w[k];
		}
		ordered_add<G>(acc, x, min((uint32_t)G, end - base));
	}
	if (c < n_cls && lane == 0)
		nd[c].y = acc;
}

// FP64 transcript pass + M-step, order-faithful: overlap_t accumulates
// (lam_t * n_c) / D_c over the classes of t in ascending class order (the
// order src/EM.c:276-285 visits them), skipping D_c == 0; then
// lam'_t = ((double)u_t + overlap_t) / len_t (src/EM.c:289-290) and the
// stop test of src/EM.c:292-297 when flags != NULL.  With member weights w
// (CSC order) the term is ((lam_t * w) * n_c) / D_c.
template <int G>
__global__ __launch_bounds__(BLOCK) void k_gather_exact(
    const uint32_t *__restrict__ off, const uint32_t *__restrict__ ref,
    const double *__restrict__ w, const double2 *__restrict__ nd, const uint32_t *__restrict__ unique,
    const uint32_t *__restrict__ len, const double *__restrict__ lam,
    double *__restrict__ out64, uint32_t *__restrict__ flags, uint32_t n_tx,
    const hemk_ctl *__restrict__ ctl)
{
	if (ctl->done)
		return;
	const uint32_t t = blockIdx.x * (BLOCK / G) + threadIdx.x / G;
	const uint32_t lane = threadIdx.x % G;
	uint32_t beg = 0, end = 0;
	double l = 0.0;
	if (t < n_tx) {
		beg = off[t];
		end = off[t + 1];
		l = lam[t];
	}
	double acc = 0.0;
	for (uint32_t base = beg; base < end; base += G) {
		const uint32_t k = base + lane;
		double x = 0.0;
		if (k < end) {
			const double2 v = nd[ref[k]];
			const double lw = w ? l * w[k] : l;
			x = v.y == 0.0 ? 0.0 : lw * v.x / v.y;
		}
		ordered_add<G>(acc, x, min((uint32_t)G, end - base));
	}
	int bad = 0;
	if (t < n_tx && lane == 0) {
		const double nw = ((double)unique[t] + acc) / (double)len[t];
		out64[t] = nw;
		if (flags && nw > ctl->stop_min &&
		    fabs(nw - l) / nw > ctl->stop_rel)
			bad = 1;
	}
	if (flags) {
		bad = __syncthreads_or(bad);
		if (threadIdx.x == 0 && bad)
			em_mark_bad(flags);
	}
}

__device__ __forceinline__ double step_length(double A, double B, uint32_t n,
                                              hemk_step st)
{
	return em_step_length(A, B, n, st);
}

template <int G, typename TL>
__global__ __launch_bounds__(BLOCK) void k_denom(
    const uint32_t *__restrict__ off, const uint32_t *__restrict__ ref,
    const TL *__restrict__ cnt, const TL *__restrict__ lam,
    TL *__restrict__ rho, uint32_t n_cls, const hemk_ctl *__restrict__ ctl)
{
	if (ctl->done)
		return;
	const uint32_t c = blockIdx.x * (BLOCK / G) + threadIdx.x / G;
	const uint32_t lane = threadIdx.x % G;
	uint32_t beg = 0, end = 0;
	if (c < n_cls) {
		beg = off[c];
		end = off[c + 1];
	}
	TL acc = 0;
	for (uint32_t k = beg + lane; k < end; k += G)
		acc += lam[ref[k]];
	acc = group_sum<G>(acc);
	if (c < n_cls && lane == 0)
		rho[c] = acc == (TL)0 ? (TL)0 : cnt[c] / acc;
}

template <int G, typename TL>
__global__ __launch_bounds__(BLOCK) void k_gather(
    const uint32_t *__restrict__ off, const uint32_t *__restrict__ ref,
    const TL *__restrict__ rho, const uint32_t *__restrict__ unique,
    const uint32_t *__restrict__ len, const double *__restrict__ lam,
    double *__restrict__ out64, float *__restrict__ out32,
    uint32_t *__restrict__ flags, uint32_t n_tx,
    const hemk_ctl *__restrict__ ctl)
{
	if (ctl->done)
		return;
	const uint32_t t = blockIdx.x * (BLOCK / G) + threadIdx.x / G;
	const uint32_t lane = threadIdx.x % G;
	uint32_t beg = 0, end = 0;
	if (t < n_tx) {
		beg = off[t];
		end = off[t + 1];
	}
	TL acc = 0;
	for (uint32_t k = beg + lane; k < end; k += G)
		acc += rho[ref[k]];
	acc = group_sum<G>(acc);
	int bad = 0;
	if (t < n_tx && lane == 0) {
		const double l = lam[t];
		const double nw = ((double)unique[t] + l * (double)acc) /
		                  (double)len[t];
		out64[t] = nw;
		if (out32)
			out32[t] = (float)nw;
		if (flags && nw > ctl->stop_min &&
		    fabs(nw - l) / nw > ctl->stop_rel)
			bad = 1;
	}
	if (flags) {
		bad = __syncthreads_or(bad);
		if (threadIdx.x == 0 && bad)
			em_mark_bad(flags);
	}
}

__global__ __launch_bounds__(BLOCK) void k_norm1(
    uint32_t n, const double *__restrict__ u, const double *__restrict__ u1,
    const double *__restrict__ u2, double *__restrict__ part, int step,
    const hemk_ctl *__restrict__ ctl)
{
#pragma clang fp contract(off)
	if (ctl->done)
		return;
	double a = 0, b = 0;
	for (uint32_t i = blockIdx.x * BLOCK + threadIdx.x; i < n;
	     i += gridDim.x * BLOCK) {
		const double r = u1[i] - u[i];
		const double v = (u2[i] - u1[i]) - r;
		if (step == 0) {
			a += r;
			b += v;
		} else {
			a += r * r;
			b += v * v;
		}
	}
	block_sum2(a, b);
	if (threadIdx.x == 0) {
		part[2 * blockIdx.x] = a;
		part[2 * blockIdx.x + 1] = b;
	}
}

__global__ __launch_bounds__(BLOCK) void k_norm2(
    uint32_t n, const double *__restrict__ u, const double *__restrict__ u1,
    const double *__restrict__ u2, const double *__restrict__ part1,
    double *__restrict__ part2, const hemk_ctl *__restrict__ ctl)
{
#pragma clang fp contract(off)
	if (ctl->done)
		return;
	double mr, mv;
	reduce_parts(part1, mr, mv);
	mr /= n;
	mv /= n;
	double a = 0, b = 0;
	for (uint32_t i = blockIdx.x * BLOCK + threadIdx.x; i < n;
	     i += gridDim.x * BLOCK) {
		const double r = u1[i] - u[i];
		const double v = (u2[i] - u1[i]) - r;
		a += (r - mr) * (r - mr);
		b += (v - mv) * (v - mv);
	}
	block_sum2(a, b);
	if (threadIdx.x == 0) {
		part2[2 * blockIdx.x] = a;
		part2[2 * blockIdx.x + 1] = b;
	}
}

__global__ __launch_bounds__(BLOCK) void k_extrap(
    uint32_t n, const double *__restrict__ u, const double *__restrict__ u1,
    const double *__restrict__ u2, const double *__restrict__ part,
    double *__restrict__ out64, float *__restrict__ out32, hemk_step st,
    hemk_ctl *__restrict__ ctl)
{
	// No contraction: u - 2*alpha*r + alpha*alpha*v rounds exactly as
	// src/EM.c:350 (u - (2*alpha)*r, then + (alpha*alpha)*v).
#pragma clang fp contract(off)
	if (ctl->done)
		return;
	double alpha;
	if (st.alpha_given == 1) {
		alpha = ctl->alpha;
	} else if (st.alpha_given == 2) {
		alpha = step_length(part[0], part[1], n, st);
	} else {
		double A, B;
		reduce_parts(part, A, B);
		alpha = step_length(A, B, n, st);
	}
	for (uint32_t i = blockIdx.x * BLOCK + threadIdx.x; i < n;
	     i += gridDim.x * BLOCK) {
		const double x = em_extrapolate(u[i], u1[i], u2[i], alpha, st);
		out64[i] = x;
		if (out32)
			out32[i] = (float)x;
	}
	if (blockIdx.x == 0 && threadIdx.x == 0 && st.alpha_given != 1)
		ctl->alpha = alpha;
}

// acc + x[0] + ... + x[cnt-1] left to right, reading the next eight values
// from LDS while the current eight are added.
__device__ __forceinline__ double seq_chain(double acc, const double *x,
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

// Sequential sums on one lane per chain (see hemk_seq_norm).  Waves 2-3
// stage chunk c (the element values, computed in parallel) in LDS while
// lane 0 of waves 0-1 adds chunk c-1 left to right.
#define SEQ_TB 1024
__global__ __launch_bounds__(BLOCK) void k_seq_norm(
    uint32_t n, const double *__restrict__ u, const double *__restrict__ u1,
    const double *__restrict__ u2, int phase, hemk_step st,
    hemk_ctl *__restrict__ ctl)
{
#pragma clang fp contract(off)
	__shared__ double buf[2][2][SEQ_TB];
	if (phase != 0 && ctl->done)
		return;
	const int w = threadIdx.x / WAVE;
	const int lane = threadIdx.x % WAVE;
	const uint32_t nch = (n + SEQ_TB - 1) / SEQ_TB;
	const double mr = phase == 2 ? ctl->mean_r : 0.0;
	const double mv = phase == 2 ? ctl->mean_v : 0.0;
	double acc = 0;
	for (uint32_t c = 0; c <= nch; ++c) {
		if (w >= 2 && c < nch) {
			const int slot = c & 1;
			const uint32_t base = c * SEQ_TB;
			for (int j = threadIdx.x - 2 * WAVE; j < SEQ_TB;
			     j += 2 * WAVE) {
				const uint32_t i = base + j;
				double a = 0, b = 0;
				if (i < n) {
					if (phase == 0) {
						a = u[i];
					} else {
						const double r = u1[i] - u[i];
						const double v = (u2[i] - u1[i]) - r;
						if (phase == 1) {
							a = r;
							b = v;
						} else if (phase == 2) {
							a = (r - mr) * (r - mr);
							b = (v - mv) * (v - mv);
						} else {
							a = r * r;
							b = v * v;
						}
					}
				}
				buf[slot][0][j] = a;
				buf[slot][1][j] = b;
			}
		}
		if (w < 2 && lane == 0 && c >= 1) {
			const int slot = (c - 1) & 1;
			const uint32_t cnt = min((uint32_t)SEQ_TB,
			                         n - (c - 1) * SEQ_TB);
			acc = seq_chain(acc, buf[slot][w], cnt);
		}
		__syncthreads();
	}
	__shared__ double res[2];
	if (w < 2 && lane == 0)
		res[w] = acc;
	__syncthreads();
	if (threadIdx.x == 0) {
		if (phase == 0) {
			ctl->lam_sum = res[0];
		} else if (phase == 1) {
			ctl->mean_r = res[0] / n;
			ctl->mean_v = res[1] / n;
		} else {
			ctl->alpha = step_length(res[0], res[1], n, st);
		}
	}
}

// The third map's kernels OR a failure of the stop test into ctl->stopbad.
__global__ void k_iter_end(hemk_ctl *__restrict__ ctl)
{
	if (ctl->done)
		return;
	em_iter_end(ctl);
}

__global__ void k_ctl_reset(hemk_ctl *ctl, uint32_t min_iter,
                            uint32_t max_iter, double stop_min,
                            double stop_rel, uint32_t *hdone,
                            uint32_t *hk)
{
	ctl->hdone = hdone;
	ctl->hk = hk;
	ctl->k = 0;
	ctl->maps = 0;
	ctl->converged = 0;
	ctl->fallbacks = 0;
	ctl->reruns = 0;
	ctl->stopbad = 0;
	ctl->stopbad2 = 0;
	ctl->fb = 0;
	ctl->min_iter = min_iter;
	ctl->max_iter = max_iter;
	ctl->alpha = 0;
	ctl->stop_min = stop_min;
	ctl->stop_rel = stop_rel;
	// Loop condition of src/EM.c:328 at k = 0 with stop = 0.
	ctl->done = 0;
	if (!(0 < min_iter || 0 < max_iter))
		em_set_done(ctl, 0);
}

__global__ __launch_bounds__(BLOCK) void k_init_state(
    uint32_t n, double *__restrict__ lam64, float *__restrict__ lam32,
    double value, const double *__restrict__ init)
{
	for (uint32_t i = blockIdx.x * BLOCK + threadIdx.x; i < n;
	     i += gridDim.x * BLOCK) {
		const double x = init ? init[i] : value;
		lam64[i] = x;
		if (lam32)
			lam32[i] = (float)x;
	}
}

__global__ __launch_bounds__(BLOCK) void k_fin1(uint32_t n,
                                                const double *__restrict__ lam,
                                                double *__restrict__ part)
{
	double a = 0, b = 0;
	for (uint32_t i = blockIdx.x * BLOCK + threadIdx.x; i < n;
	     i += gridDim.x * BLOCK)
		a += lam[i];
	block_sum2(a, b);
	if (threadIdx.x == 0) {
		part[2 * blockIdx.x] = a;
		part[2 * blockIdx.x + 1] = 0;
	}
}

__global__ __launch_bounds__(BLOCK) void k_fin2(
    uint32_t n, const double *__restrict__ lam,
    const uint32_t *__restrict__ len, const double *__restrict__ part,
    double *__restrict__ tpm, double *__restrict__ est)
{
	double sum, zero;
	reduce_parts(part, sum, zero);
	for (uint32_t i = blockIdx.x * BLOCK + threadIdx.x; i < n;
	     i += gridDim.x * BLOCK) {
		const double l = lam[i];
		double e = 0, g = 0;
		if (sum != 0) {
			e = l * (double)len[i];
			g = l * 1000000 / sum;
		}
		if (est)
			est[i] = e;
		if (tpm)
			tpm[i] = g;
	}
}

__global__ __launch_bounds__(BLOCK) void k_fin_seq(
    uint32_t n, const double *__restrict__ lam,
    const uint32_t *__restrict__ len, const hemk_ctl *__restrict__ ctl,
    double *__restrict__ tpm, double *__restrict__ est)
{
	const double sum = ctl->lam_sum;
	for (uint32_t i = blockIdx.x * BLOCK + threadIdx.x; i < n;
	     i += gridDim.x * BLOCK) {
		const double l = lam[i];
		double e = 0, g = 0;
		if (sum != 0) {
			e = l * (double)len[i];
			g = l * 1000000 / sum;
		}
		if (est)
			est[i] = e;
		if (tpm)
			tpm[i] = g;
	}
}

// FP64 log-likelihood partials: class terms in a, transcript terms in b.
// Classes with D_c == 0 are skipped, as the map skips them.
template <int G>
__global__ __launch_bounds__(BLOCK) void k_loglik(
    const uint32_t *__restrict__ off, const uint32_t *__restrict__ ref,
    const double *__restrict__ cnt, const uint32_t *__restrict__ unique,
    const uint32_t *__restrict__ len, const double *__restrict__ lam,
    uint32_t n_cls, uint32_t n_tx, double *__restrict__ part)
{
	const uint32_t lane = threadIdx.x % G;
	const uint32_t groups = gridDim.x * (BLOCK / G);
	double a = 0, b = 0;
	for (uint32_t c = blockIdx.x * (BLOCK / G) + threadIdx.x / G;
	     c < n_cls; c += groups) {
		double d = 0;
		for (uint32_t k = off[c] + lane; k < off[c + 1]; k += G)
			d += lam[ref[k]];
		d = group_sum<G>(d);
		if (lane == 0 && cnt[c] != 0 && d > 0)
			a += cnt[c] * log(d);
	}
	for (uint32_t i = blockIdx.x * BLOCK + threadIdx.x; i < n_tx;
	     i += gridDim.x * BLOCK) {
		const double l = lam[i];
		if (unique[i] != 0)
			b += (double)unique[i] * log(l);
		b -= (double)len[i] * l;
	}
	block_sum2(a, b);
	if (threadIdx.x == 0) {
		part[2 * blockIdx.x] = a;
		part[2 * blockIdx.x + 1] = b;
	}
}

__global__ __launch_bounds__(BLOCK) void k_loglik_fin(
    const double *__restrict__ part, double *__restrict__ out)
{
	double a, b;
	reduce_parts(part, a, b);
	if (threadIdx.x == 0)
		out[0] = a + b;
}

inline uint32_t cdiv(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

#define HEM_GROUP_SWITCH(group, CASE)                                      \
	switch (group) {                                                    \
		CASE(1)                                                     \
		CASE(2)                                                     \
		CASE(4)                                                     \
		CASE(8)                                                     \
		CASE(16)                                                    \
		CASE(32)                                                    \
		CASE(64)                                                    \
	default:                                                            \
		return (int)hipErrorInvalidValue;                           \
	}

int launch_denom(const hemk_prob *p, int prec, const void *lam, void *rho,
                 int group, const hemk_ctl *ctl, hipStream_t s)
{
	if (p->n_cls == 0)
		return 0;
	const uint32_t grid = cdiv(p->n_cls, BLOCK / group);
	if (prec == 0) {
#define HEM_DENOM32(G)                                                      \
	case G:                                                             \
		k_denom<G, float><<<grid, BLOCK, 0, s>>>(                   \
		    p->cls_off, p->cls_ref, p->cnt32, (const float *)lam,     \
		    (float *)rho, p->n_cls, ctl);                             \
		break;
		HEM_GROUP_SWITCH(group, HEM_DENOM32)
#undef HEM_DENOM32
	} else {
#define HEM_DENOM64(G)                                                      \
	case G:                                                             \
		k_denom_exact<G><<<grid, BLOCK, 0, s>>>(                    \
		    p->cls_off, p->cls_ref, p->cls_w, (const double *)lam,    \
		    (double2 *)rho, p->n_cls, ctl);                           \
		break;
		HEM_GROUP_SWITCH(group, HEM_DENOM64)
#undef HEM_DENOM64
	}
	return (int)hipGetLastError();
}

int launch_gather(const hemk_prob *p, int prec, const void *rho,
                  const uint32_t *unique, const double *lam, double *out64,
                  float *out32, uint32_t *flags, int group,
                  const hemk_ctl *ctl, hipStream_t s)
{
	const uint32_t grid = cdiv(p->n_tx, BLOCK / group);
	if (prec == 0) {
#define HEM_GATHER32(G)                                                     \
	case G:                                                             \
		k_gather<G, float><<<grid, BLOCK, 0, s>>>(                  \
		    p->tx_off, p->tx_ref, (const float *)rho, unique, p->len, \
		    lam, out64, out32, flags, p->n_tx, ctl);                  \
		break;
		HEM_GROUP_SWITCH(group, HEM_GATHER32)
#undef HEM_GATHER32
	} else {
#define HEM_GATHER64(G)                                                     \
	case G:                                                             \
		k_gather_exact<G><<<grid, BLOCK, 0, s>>>(                   \
		    p->tx_off, p->tx_ref, p->tx_w, (const double2 *)rho,      \
		    unique,                                                   \
		    p->len, lam, out64, flags, p->n_tx, ctl);                 \
		break;
		HEM_GROUP_SWITCH(group, HEM_GATHER64)
#undef HEM_GATHER64
	}
	return (int)hipGetLastError();
}

} // namespace

extern "C" {

int hemk_init_state(uint32_t n, double *lam64, float *lam32, double value,
                    const double *init64, void *s)
{
	const uint32_t grid = cdiv(n, BLOCK) < 4096 ? cdiv(n, BLOCK) : 4096;
	k_init_state<<<grid ? grid : 1, BLOCK, 0, (hipStream_t)s>>>(
	    n, lam64, lam32, value, init64);
	return (int)hipGetLastError();
}

int hemk_ctl_reset(hemk_ctl *ctl, uint32_t min_iter, uint32_t max_iter,
                   double stop_min, double stop_rel, uint32_t *hdone,
                   uint32_t *hk, void *s)
{
	k_ctl_reset<<<1, 1, 0, (hipStream_t)s>>>(ctl, min_iter, max_iter,
	                                         stop_min, stop_rel, hdone, hk);
	return (int)hipGetLastError();
}

int hemk_denom(const hemk_prob *p, int prec, const void *lam, void *rho,
               int group, const hemk_ctl *ctl, void *s)
{
	return launch_denom(p, prec, lam, rho, group, ctl, (hipStream_t)s);
}

uint32_t hemk_gather_blocks(const hemk_prob *p, int group)
{
	return cdiv(p->n_tx, BLOCK / group);
}

int hemk_gather_mstep(const hemk_prob *p, int prec, const void *rho,
                      const uint32_t *unique, const double *lam64,
                      double *out64, float *out32, uint32_t *stop_flags,
                      int group, const hemk_ctl *ctl, void *s)
{
	return launch_gather(p, prec, rho, unique, lam64, out64, out32,
	                     stop_flags, group, ctl, (hipStream_t)s);
}

int hemk_norm_pass1(uint32_t n, const double *u, const double *u1,
                    const double *u2, double *part1, int step,
                    const hemk_ctl *ctl, void *s)
{
	k_norm1<<<HEMK_NORM_BLOCKS, BLOCK, 0, (hipStream_t)s>>>(n, u, u1, u2,
	                                                        part1, step, ctl);
	return (int)hipGetLastError();
}

int hemk_norm_pass2(uint32_t n, const double *u, const double *u1,
                    const double *u2, const double *part1, double *part2,
                    int step, const hemk_ctl *ctl, void *s)
{
	(void)step;
	k_norm2<<<HEMK_NORM_BLOCKS, BLOCK, 0, (hipStream_t)s>>>(
	    n, u, u1, u2, part1, part2, ctl);
	return (int)hipGetLastError();
}

int hemk_extrapolate(uint32_t n, const double *u, const double *u1,
                     const double *u2, const double *part2, double *out64,
                     float *out32, hemk_step st, hemk_ctl *ctl, void *s)
{
	k_extrap<<<HEMK_NORM_BLOCKS, BLOCK, 0, (hipStream_t)s>>>(
	    n, u, u1, u2, part2, out64, out32, st, ctl);
	return (int)hipGetLastError();
}

int hemk_iter_end(hemk_ctl *ctl, void *s)
{
	k_iter_end<<<1, 1, 0, (hipStream_t)s>>>(ctl);
	return (int)hipGetLastError();
}

int hemk_finalize(uint32_t n, const double *lam, const uint32_t *len,
                  double *part, double *tpm, double *est, void *s)
{
	k_fin1<<<HEMK_NORM_BLOCKS, BLOCK, 0, (hipStream_t)s>>>(n, lam, part);
	k_fin2<<<HEMK_NORM_BLOCKS, BLOCK, 0, (hipStream_t)s>>>(n, lam, len,
	                                                       part, tpm, est);
	return (int)hipGetLastError();
}

int hemk_seq_norm(uint32_t n, const double *u, const double *u1,
                  const double *u2, int phase, hemk_step st, hemk_ctl *ctl,
                  void *s)
{
	k_seq_norm<<<1, BLOCK, 0, (hipStream_t)s>>>(n, u, u1, u2, phase, st, ctl);
	return (int)hipGetLastError();
}

int hemk_finalize_seq(uint32_t n, const double *lam, const uint32_t *len,
                      const hemk_ctl *ctl, double *tpm, double *est, void *s)
{
	k_fin_seq<<<HEMK_NORM_BLOCKS, BLOCK, 0, (hipStream_t)s>>>(n, lam, len,
	                                                          ctl, tpm, est);
	return (int)hipGetLastError();
}

int hemk_loglik(const hemk_prob *p, const uint32_t *unique, const double *lam,
                double *part, double *out, void *s)
{
	k_loglik<4><<<HEMK_NORM_BLOCKS, BLOCK, 0, (hipStream_t)s>>>(
	    p->cls_off, p->cls_ref, p->cnt64, unique, p->len, lam, p->n_cls,
	    p->n_tx, part);
	k_loglik_fin<<<1, BLOCK, 0, (hipStream_t)s>>>(part, out);
	return (int)hipGetLastError();
}

} // extern "C"
