// hera_em_kernels.hip -- hera-EM's simplified-model SQUAREM (gfx950).
//
// The simple model's E-step (the e_step macro of hera-EM/em.c:75-157 with
// s_joint_prob / s_noise_prob / s_model_update / s_model_update_noise,
// hera-EM/smodel.h:15-45) on a class table whose classes are groups of
// identical reads, its M-step (trans_master_init + lock_normalize,
// hera-EM/transdis.c:20-30, 54-61, utility.h:104-122), the locked
// extrapolation and locking rule (hera-EM/utility.c:64-107,
// transdis.c:97-115) and the stop test (arr_diff / is_stop, utility.c:
// 109-128, transdis.c:123-130).  Filtering state persists on the device:
// a live flag per member, the locked mass P0 and a removed flag and noise
// bank per class, a lock flag per transcript.  Every sum runs in one fixed
// order (classes ascending, members in class order, rows in index order),
// the order of the CPU restatement, so the results are bitwise equal.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "hem_dev.h"

#define HBLOCK 256
#define HCHUNK 2048

namespace {

// out = (((+0 + x[0]) + x[1]) + ...): the block stages chunks of the row in
// LDS and one lane adds them in order.
__global__ __launch_bounds__(HBLOCK) void k_he_rowsum(
    const double *__restrict__ x, uint32_t n, double *__restrict__ out)
{
#pragma clang fp contract(off)
	__shared__ double buf[HCHUNK];
	double s = 0.0;
	for (uint32_t b = 0; b < n; b += HCHUNK) {
		const uint32_t m = n - b < HCHUNK ? n - b : HCHUNK;
		for (uint32_t i = threadIdx.x; i < m; i += HBLOCK)
			buf[i] = x[b + i];
		__syncthreads();
		if (threadIdx.x == 0)
			for (uint32_t i = 0; i < m; ++i)
				s += buf[i];
		__syncthreads();
	}
	if (threadIdx.x == 0)
		*out = s;
}

// Member k of class c is occurrence pos of its transcript's class list
// (classes ascending, a class listed once per occurrence): occ2mem[pos] = k.
__global__ __launch_bounds__(HBLOCK) void k_he_occ2mem(
    uint32_t C, const uint32_t *__restrict__ off,
    const uint32_t *__restrict__ ref, const uint32_t *__restrict__ tx_off,
    const uint32_t *__restrict__ tx_ref, uint32_t *__restrict__ occ2mem)
{
	const uint32_t c = blockIdx.x * HBLOCK + threadIdx.x;
	if (c >= C)
		return;
	for (uint32_t k = off[c]; k < off[c + 1]; ++k) {
		const uint32_t t = ref[k];
		uint32_t dup = 0;
		for (uint32_t q = off[c]; q < k; ++q)
			dup += ref[q] == t;
		uint32_t lo = tx_off[t], hi = tx_off[t + 1];
		while (lo < hi) {
			const uint32_t mid = (lo + hi) / 2;
			if (tx_ref[mid] < c)
				lo = mid + 1;
			else
				hi = mid;
		}
		occ2mem[lo + dup] = k;
	}
}

// row[c] = bank[c] for a removed class, else 0.
__global__ __launch_bounds__(HBLOCK) void k_he_bank_row(
    uint32_t C, const uint8_t *__restrict__ dead,
    const double *__restrict__ bank, double *__restrict__ row)
{
	for (uint32_t c = blockIdx.x * HBLOCK + threadIdx.x; c < C;
	     c += gridDim.x * HBLOCK)
		row[c] = dead[c] ? bank[c] : 0.0;
}

// zrow[0]: the banked noise count when it is added (not below the
// threshold, noise not locked).
__global__ void k_he_bank_term(const hemk_he h, double *__restrict__ zrow)
{
	const double B = h.sums[0];
	zrow[0] = (h.z < h.T && !h.lock[h.z] && B >= h.thr) ? B : 0.0;
}

// One thread per class: its total (P0 and live members' p_t w_k in member
// order), its noise term into zrow[c + 1], and with filt the removals.
__global__ __launch_bounds__(HBLOCK) void k_he_class(
    const hemk_he h, const double *__restrict__ p, int filt,
    double *__restrict__ zrow)
{
#pragma clang fp contract(off)
	const uint32_t c = blockIdx.x * HBLOCK + threadIdx.x;
	if (c >= h.C)
		return;
	double zterm = 0.0;
	if (!h.dead[c]) {
		const uint32_t k0 = h.off[c], k1 = h.off[c + 1];
		const double n = h.cnt[c];
		double total = h.P0[c];
		for (uint32_t k = k0; k < k1; ++k)
			if (h.alive[k])
				total += p[h.ref[k]] * (h.w ? h.w[k] : 1.0);
		h.total[c] = total;
		if (total < 1e-300) {
			if (filt) {
				h.dead[c] = 1;
				h.bank[c] = n;
				atomicAdd(&h.counts[1], 1u);
			}
		} else {
			uint32_t live = 0, gone = 0;
			double zpost = 0.0;
			for (uint32_t k = k0; k < k1; ++k) {
				if (!h.alive[k])
					continue;
				const uint32_t t = h.ref[k];
				const double x = p[t] * (h.w ? h.w[k] : 1.0);
				const double post = x / total;
				if (t == h.z) {
					zpost = post;
					if (post >= h.thr && !h.lock[t])
						zterm = n * post;
					continue;
				}
				if (post < h.thr) {
					if (filt) {
						h.alive[k] = 0;
						++gone;
					}
				} else if (h.lock[t]) {
					if (filt) {
						h.alive[k] = 0;
						++gone;
						h.P0[c] += x;
					}
				} else {
					++live;
				}
			}
			if (gone)
				atomicAdd(&h.counts[0], gone);
			if (filt && live == 0) {
				h.dead[c] = 1;
				h.bank[c] = n * zpost;
				atomicAdd(&h.counts[1], 1u);
			}
		}
	}
	zrow[c + 1] = zterm;
}

// One thread per transcript (the noise's sum is a row sum): the terms
// n_c ((p_t w_k) / total_c) of its live class occurrences in ascending
// class order, skipping posteriors below the threshold and locked
// transcripts.
__global__ __launch_bounds__(HBLOCK) void k_he_tx(
    const hemk_he h, const double *__restrict__ p,
    double *__restrict__ acc)
{
#pragma clang fp contract(off)
	const uint32_t t = blockIdx.x * HBLOCK + threadIdx.x;
	if (t >= h.T || t == h.z)
		return;
	double s = 0.0;
	if (!h.lock[t]) {
		const double pt = p[t];
		for (uint32_t j = h.tx_off[t]; j < h.tx_off[t + 1]; ++j) {
			const uint32_t c = h.tx_ref[j];
			const uint32_t k = h.occ2mem[j];
			if (h.dead[c] || !h.alive[k])
				continue;
			const double total = h.total[c];
			if (total < 1e-300)
				continue;
			const double post =
			    (pt * (h.w ? h.tx_w[j] : 1.0)) / total;
			if (post < h.thr)
				continue;
			s += h.cnt[c] * post;
		}
	}
	acc[t] = s;
}

// Rows of the M-step's two sums: unlocked acc, locked p.
__global__ __launch_bounds__(HBLOCK) void k_he_mrows(
    const hemk_he h, const double *__restrict__ p,
    const double *__restrict__ acc, double *__restrict__ urow,
    double *__restrict__ lrow)
{
	for (uint32_t t = blockIdx.x * HBLOCK + threadIdx.x; t < h.T;
	     t += gridDim.x * HBLOCK) {
		const bool l = h.lock[t];
		urow[t] = l ? 0.0 : acc[t];
		lrow[t] = l ? p[t] : 0.0;
	}
}

// out = locked ? p : acc * ((1 / sum) * (1 - l_sum)).
__global__ __launch_bounds__(HBLOCK) void k_he_mapply(
    const hemk_he h, const double *__restrict__ p,
    const double *__restrict__ acc, double *__restrict__ out)
{
#pragma clang fp contract(off)
	const double f = 1 / h.sums[1] * (1 - h.sums[2]);
	for (uint32_t t = blockIdx.x * HBLOCK + threadIdx.x; t < h.T;
	     t += gridDim.x * HBLOCK)
		out[t] = h.lock[t] ? p[t] : acc[t] * f;
}

// Rows of r^2 and v^2 over unlocked entries (r = p2 - p1, v = p3 - p2 - r).
__global__ __launch_bounds__(HBLOCK) void k_he_rvrows(
    const hemk_he h, const double *__restrict__ p1,
    const double *__restrict__ p2, const double *__restrict__ p3,
    double *__restrict__ rrow, double *__restrict__ vrow)
{
#pragma clang fp contract(off)
	for (uint32_t t = blockIdx.x * HBLOCK + threadIdx.x; t < h.T;
	     t += gridDim.x * HBLOCK) {
		double a = 0.0, b = 0.0;
		if (!h.lock[t]) {
			const double r = p2[t] - p1[t];
			const double v = p3[t] - p2[t] - r;
			a = r * r;
			b = v * v;
		}
		rrow[t] = a;
		vrow[t] = b;
	}
}

// The locked extrapolation of lock_accelerate and the locking rule.
__global__ __launch_bounds__(HBLOCK) void k_he_extrap(
    const hemk_he h, const double *__restrict__ p1,
    const double *__restrict__ p2, double *__restrict__ p3, int locking,
    double lock_thr)
{
#pragma clang fp contract(off)
	const double r2 = h.sums[3], v2 = h.sums[4];
	const double g2 = v2 <= 1e-300 ? 1 : r2 / v2;
	const double g = -sqrt(g2);
	for (uint32_t t = blockIdx.x * HBLOCK + threadIdx.x; t < h.T;
	     t += gridDim.x * HBLOCK) {
		if (h.lock[t])
			continue;
		const double r = p2[t] - p1[t];
		const double v = p3[t] - p2[t] - r;
		double nv = p1[t] - 2 * g * r + g2 * v;
		if (nv <= 1e-300)
			nv = p3[t] / 10;
		p3[t] = nv;
		if (locking) {
			const double diff = fabs(nv - p1[t]);
			h.lock[t] = nv < 1e-300 || diff < lock_thr * nv;
		}
	}
}

// is_stop: counts entries with p3 >= stop_min whose relative change is
// >= stop_rel.
__global__ __launch_bounds__(HBLOCK) void k_he_stop(
    const hemk_he h, const double *__restrict__ p3,
    const double *__restrict__ p1, double stop_min, double stop_rel)
{
	for (uint32_t t = blockIdx.x * HBLOCK + threadIdx.x; t < h.T;
	     t += gridDim.x * HBLOCK) {
		double diff = fabs(p3[t] - p1[t]);
		if (p3[t] >= stop_min) {
			diff /= p3[t];
			if (diff >= stop_rel)
				atomicAdd(&h.counts[2], 1u);
		}
	}
}

__global__ __launch_bounds__(HBLOCK) void k_he_fill(double *__restrict__ x,
                                                    uint32_t n, double v)
{
	for (uint32_t i = blockIdx.x * HBLOCK + threadIdx.x; i < n;
	     i += gridDim.x * HBLOCK)
		x[i] = v;
}

uint32_t grid_for(uint32_t n)
{
	const uint32_t g = (n + HBLOCK - 1) / HBLOCK;
	return g ? (g < 4096 ? g : 4096) : 1;
}

} // namespace

extern "C" {

int hemk_he_occ2mem(uint32_t C, const uint32_t *off, const uint32_t *ref,
                    const uint32_t *tx_off, const uint32_t *tx_ref,
                    uint32_t *occ2mem, void *s)
{
	if (!C)
		return 0;
	k_he_occ2mem<<<(C + HBLOCK - 1) / HBLOCK, HBLOCK, 0, (hipStream_t)s>>>(
	    C, off, ref, tx_off, tx_ref, occ2mem);
	return (int)hipGetLastError();
}

int hemk_he_fill(double *x, uint32_t n, double v, void *s)
{
	if (!n)
		return 0;
	k_he_fill<<<grid_for(n), HBLOCK, 0, (hipStream_t)s>>>(x, n, v);
	return (int)hipGetLastError();
}

int hemk_he_rowsum(const double *x, uint32_t n, double *out, void *s)
{
	k_he_rowsum<<<1, HBLOCK, 0, (hipStream_t)s>>>(x, n, out);
	return (int)hipGetLastError();
}

int hemk_he_estep(const hemk_he *h, const double *p, int filt, double *acc,
                  void *s)
{
	hipStream_t st = (hipStream_t)s;
	// the banked noise count B (removed classes, class order)
	if (h->C) {
		k_he_bank_row<<<grid_for(h->C), HBLOCK, 0, st>>>(
		    h->C, h->dead, h->bank, h->brow);
		k_he_rowsum<<<1, HBLOCK, 0, st>>>(h->brow, h->C, &h->sums[0]);
	} else {
		k_he_fill<<<1, HBLOCK, 0, st>>>(&h->sums[0], 1, 0.0);
	}
	k_he_bank_term<<<1, 1, 0, st>>>(*h, h->zrow);
	if (h->C)
		k_he_class<<<(h->C + HBLOCK - 1) / HBLOCK, HBLOCK, 0, st>>>(
		    *h, p, filt, h->zrow);
	k_he_tx<<<(h->T + HBLOCK - 1) / HBLOCK, HBLOCK, 0, st>>>(*h, p, acc);
	if (h->z < h->T)
		k_he_rowsum<<<1, HBLOCK, 0, st>>>(h->zrow, h->C + 1, acc + h->z);
	return (int)hipGetLastError();
}

int hemk_he_mstep(const hemk_he *h, const double *p, const double *acc,
                  double *out, void *s)
{
	hipStream_t st = (hipStream_t)s;
	k_he_mrows<<<grid_for(h->T), HBLOCK, 0, st>>>(*h, p, acc, h->urow,
	                                              h->lrow);
	k_he_rowsum<<<1, HBLOCK, 0, st>>>(h->urow, h->T, &h->sums[1]);
	k_he_rowsum<<<1, HBLOCK, 0, st>>>(h->lrow, h->T, &h->sums[2]);
	k_he_mapply<<<grid_for(h->T), HBLOCK, 0, st>>>(*h, p, acc, out);
	return (int)hipGetLastError();
}

int hemk_he_accelerate(const hemk_he *h, const double *p1, const double *p2,
                       double *p3, int locking, double lock_thr, void *s)
{
	hipStream_t st = (hipStream_t)s;
	k_he_rvrows<<<grid_for(h->T), HBLOCK, 0, st>>>(*h, p1, p2, p3, h->urow,
	                                               h->lrow);
	k_he_rowsum<<<1, HBLOCK, 0, st>>>(h->urow, h->T, &h->sums[3]);
	k_he_rowsum<<<1, HBLOCK, 0, st>>>(h->lrow, h->T, &h->sums[4]);
	k_he_extrap<<<grid_for(h->T), HBLOCK, 0, st>>>(*h, p1, p2, p3, locking,
	                                               lock_thr);
	return (int)hipGetLastError();
}

int hemk_he_stop(const hemk_he *h, const double *p3, const double *p1,
                 double stop_min, double stop_rel, void *s)
{
	k_he_stop<<<grid_for(h->T), HBLOCK, 0, (hipStream_t)s>>>(
	    *h, p3, p1, stop_min, stop_rel);
	return (int)hipGetLastError();
}

} // extern "C"
