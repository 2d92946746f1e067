// tile_kernels.hip -- batched SQUAREM EM over component tiles (gfx950).
//
// A workgroup owns one tile (whole connected components) and a chunk of RC
// replicates.  The batch's state, unique counts and class counts are kept
// in tile position order, replicate fastest, so a workgroup's rows are one
// contiguous block each and every global read it needs goes out in a single
// round once the tile descriptor is known.  It stages them in LDS with the
// tile's structure and runs hera's class pass and transcript pass
// (src/EM.c:276-291) on LDS only, in the reference's order: members in
// class order for D_c, classes in ascending original id (then member
// order) for each transcript's sum of (lam_t n_c) / D_c.  Replicates are
// independent lanes (lane r of an entity group).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "em_step.h"
#include "hem_dev.h"

#define TBLOCK 256
// dynamic LDS per workgroup: four workgroups share a CU's 160 KB with
// room for the kernels' static arrays
#define TLDS_BUDGET (40 * 1024 - 512)

namespace {

struct TileCtx {
	uint32_t tx_base, nt, cls_base, nc, mb, nz, jb;
};

__device__ __forceinline__ TileCtx load_tile(const hemk_tiles &t, uint32_t i)
{
	const uint32_t *d = t.desc + (size_t)HEMK_TILE_DESC * i;
	TileCtx c;
	c.tx_base = d[0];
	c.nt = d[1];
	c.cls_base = d[2];
	c.nc = d[3];
	c.mb = d[4];
	c.nz = d[5];
	c.jb = d[6];
	return c;
}

// The tile's structure staged in LDS, offsets relative to the tile: member
// lists (each member as (local class << 8) | local transcript), per
// transcript the local classes of its occurrences in class order, and
// the transcripts' ids and lengths.
struct TileStruct {
	uint16_t *coff;  // nc + 1
	uint16_t *toff;  // nt + 1
	uint32_t *kmem;  // members
	uint16_t *tcls;  // class-list entries: local class
	uint32_t *tx;    // nt
	uint32_t *len;   // nt
};

__device__ __forceinline__ void stage_struct(const hemk_tiles &tl,
                                             const TileCtx &tc,
                                             const TileStruct &ts)
{
	for (uint32_t i = threadIdx.x; i <= tc.nc; i += TBLOCK)
		ts.coff[i] = (uint16_t)(tl.coff[tc.cls_base + i] - tc.mb);
	for (uint32_t i = threadIdx.x; i <= tc.nt; i += TBLOCK)
		ts.toff[i] = (uint16_t)(tl.toff[tc.tx_base + i] - tc.jb);
	for (uint32_t i = threadIdx.x; i < tc.nz; i += TBLOCK) {
		ts.kmem[i] = tl.kmem[tc.mb + i];
		ts.tcls[i] = tl.tcls[tc.jb + i];
	}
	for (uint32_t i = threadIdx.x; i < tc.nt; i += TBLOCK) {
		ts.tx[i] = tl.tx[tc.tx_base + i];
		ts.len[i] = tl.len[tc.tx_base + i];
	}
}

// One EM map of the tile for the block's replicate chunk, lin -> lout in
// LDS (and to column qlist[r], the replicate of slot r, of the global
// [q][R] vector gout when given), in two passes separated by barriers.
// Slot r of an entity's RC slots holds the chunk's r-th replicate; rcl
// (a power of two, at most RC, at least the chunk's live replicates)
// lanes serve each entity, so a chunk with few live replicates spreads
// over more entities at once instead of idling lanes:
//   class pass       D_c = sum of lam over the members, in member order
//   transcript pass  lam'_t = (u_t + sum of (lam_t n_c) / D_c over its
//                    occurrences in class order) / len_t
// which is src/EM.c:275-291 with every sum in the reference's order.  With
// stopf the stop test of src/EM.c:292-297 marks stopf[r].  F: the FP32 bulk
// form, as the single-problem tile kernels compute it: FP32 class sums of
// the densities rounded to FP32, rho_c = n_c / D_c in FP32 (0 for a zero
// sum) in the D slots, FP32 terms lam_t rho_c summed per transcript in FP64
// in the same order.
//
// The class pass stores 1 in place of a D_c that is 0: densities are >= 0,
// so D_c == 0 means every member, the transcript whose term it is
// included, has lam == 0, and (0 n_c) / 1 adds the exact 0 that the
// reference's skip of that class leaves (src/EM.c:279-280).
template <typename TC, int RC>
__device__ __forceinline__ void tile_map32(
    const TileCtx &tc, const TileStruct &ts, const double *lin, double *lout,
    double *D, const TC *n, const uint32_t *u, const int *alive,
    double *gout, uint32_t R, const uint32_t *qlist, uint32_t rcl,
    int *stopf, double smin, double srel)
{
#pragma clang fp contract(off)
	const int r = threadIdx.x & (rcl - 1);
	const uint32_t e0 = threadIdx.x / rcl;
	const uint32_t ES = TBLOCK / rcl;
	const bool live = alive[r] != 0;
	const uint32_t qr = live ? qlist[r] : 0u;
	float *rho = (float *)D;
	if (live) {
		for (uint32_t c = e0; c < tc.nc; c += ES) {
			uint32_t k = ts.coff[c];
			const uint32_t k1 = ts.coff[c + 1];
			float acc = 0.0f;
			for (; k + 4 <= k1; k += 4) {
				const float x0 = (float)lin[(ts.kmem[k] & 0xFFu) * RC + r];
				const float x1 = (float)lin[(ts.kmem[k + 1] & 0xFFu) * RC + r];
				const float x2 = (float)lin[(ts.kmem[k + 2] & 0xFFu) * RC + r];
				const float x3 = (float)lin[(ts.kmem[k + 3] & 0xFFu) * RC + r];
				acc += x0;
				acc += x1;
				acc += x2;
				acc += x3;
			}
			for (; k < k1; ++k)
				acc += (float)lin[(ts.kmem[k] & 0xFFu) * RC + r];
			rho[c * RC + r] =
			    acc == 0.0f ? 0.0f : (float)n[c * RC + r] / acc;
		}
	}
	__syncthreads();
	if (live) {
		for (uint32_t t = e0; t < tc.nt; t += ES) {
			const double l = lin[t * RC + r];
			const float lf = (float)l;
			uint32_t j = ts.toff[t];
			const uint32_t j1 = ts.toff[t + 1];
			double acc = 0.0;
			for (; j + 4 <= j1; j += 4) {
				float x[4];
#pragma unroll
				for (int q = 0; q < 4; ++q)
					x[q] = lf * rho[ts.tcls[j + q] * RC + r];
#pragma unroll
				for (int q = 0; q < 4; ++q)
					acc += (double)x[q];
			}
			for (; j < j1; ++j)
				acc += (double)(lf * rho[ts.tcls[j] * RC + r]);
			const double nw = ((double)u[t * RC + r] + acc) /
			                  (double)ts.len[t];
			lout[t * RC + r] = nw;
			if (gout)
				gout[(size_t)(tc.tx_base + t) * R + qr] = nw;
			if (stopf && nw > smin && fabs(nw - l) / nw > srel)
				stopf[r] = 1;
		}
	}
	__syncthreads();
}

template <typename TC, int RC, bool F>
__device__ __forceinline__ void tile_map(
    const hemk_tiles &tl, const TileCtx &tc, const TileStruct &ts,
    const double *lin, double *lout, double *D, const TC *n,
    const uint32_t *u, const int *alive, double *gout, uint32_t R,
    const uint32_t *qlist, uint32_t rcl, int *stopf, double smin,
    double srel)
{
	if (F) {
		tile_map32<TC, RC>(tc, ts, lin, lout, D, n, u, alive, gout, R,
		                   qlist, rcl, stopf, smin, srel);
		return;
	}
	const int r = threadIdx.x & (rcl - 1);
	const uint32_t e0 = threadIdx.x / rcl;
	const uint32_t ES = TBLOCK / rcl;
	const bool live = alive[r] != 0;
	const uint32_t qr = live ? qlist[r] : 0u;
	(void)tl;
	int out = 0;   // a value outside em_quot's range
	if (live) {
		for (uint32_t c = e0; c < tc.nc; c += ES) {
			uint32_t k = ts.coff[c];
			const uint32_t k1 = ts.coff[c + 1];
			double acc = 0.0;
			for (; k + 4 <= k1; k += 4) {
				const double x0 = lin[(ts.kmem[k] & 0xFFu) * RC + r];
				const double x1 = lin[(ts.kmem[k + 1] & 0xFFu) * RC + r];
				const double x2 = lin[(ts.kmem[k + 2] & 0xFFu) * RC + r];
				const double x3 = lin[(ts.kmem[k + 3] & 0xFFu) * RC + r];
				acc += x0;
				acc += x1;
				acc += x2;
				acc += x3;
			}
			for (; k < k1; ++k)
				acc += lin[(ts.kmem[k] & 0xFFu) * RC + r];
			const double d = acc == 0.0 ? 1.0 : acc;
			D[c * RC + r] = d;
			out |= !(d >= EM_QUOT_LO && d <= EM_QUOT_HI);
		}
		// lam_t n_c with n_c < 2^32 stays in em_quot's range when lam_t
		// is 0 or in [2^-299, 2^260]
		for (uint32_t t = e0; t < tc.nt; t += ES) {
			const double l = lin[t * RC + r];
			out |= l != 0.0 && !(l >= 0x1p-299 && l <= 0x1p260);
		}
	}
	const bool fast = !__syncthreads_or(out);
	if (live && fast) {
		// every divisor and dividend in em_quot's range (the usual
		// case): the same quotients with fewer instructions, no masked
		// slots, and eight occurrences at a time with all their LDS
		// loads issued before the first quotient (the pass waits on
		// dependent LDS loads more than on arithmetic)
		for (uint32_t t = e0; t < tc.nt; t += ES) {
			const double l = lin[t * RC + r];
			uint32_t j = ts.toff[t];
			const uint32_t j1 = ts.toff[t + 1];
			double acc = 0.0;
			for (; j + 8 <= j1; j += 8) {
				uint32_t c[8];
				double x[8];
#pragma unroll
				for (int q = 0; q < 8; ++q)
					c[q] = ts.tcls[j + q];
				double nn[8], dd[8];
#pragma unroll
				for (int q = 0; q < 8; ++q) {
					nn[q] = (double)n[c[q] * RC + r];
					dd[q] = D[c[q] * RC + r];
				}
#pragma unroll
				for (int q = 0; q < 8; ++q)
					x[q] = em_quot(l * nn[q], dd[q]);
#pragma unroll
				for (int q = 0; q < 8; ++q)
					acc += x[q];
			}
			for (; j + 4 <= j1; j += 4) {
				uint32_t c[4];
				double x[4];
#pragma unroll
				for (int q = 0; q < 4; ++q)
					c[q] = ts.tcls[j + q];
#pragma unroll
				for (int q = 0; q < 4; ++q)
					x[q] = em_quot(l * (double)n[c[q] * RC + r],
					               D[c[q] * RC + r]);
#pragma unroll
				for (int q = 0; q < 4; ++q)
					acc += x[q];
			}
			for (; j < j1; ++j) {
				const uint32_t c = ts.tcls[j];
				acc += em_quot(l * (double)n[c * RC + r],
				               D[c * RC + r]);
			}
			const double nw = ((double)u[t * RC + r] + acc) /
			                  (double)ts.len[t];
			lout[t * RC + r] = nw;
			if (gout)
				gout[(size_t)(tc.tx_base + t) * R + qr] = nw;
			if (stopf && nw > smin && fabs(nw - l) / nw > srel)
				stopf[r] = 1;
		}
	} else if (live) {
		// each occurrence's term (lam_t n_c) / D_c computed where the
		// transcript's ordered sum takes it; four occurrences at a time,
		// their LDS reads and divisions straight-line (a slot past the
		// list computes a term it adds as +0, so the four divisions
		// overlap instead of waiting on branches)
		for (uint32_t t = e0; t < tc.nt; t += ES) {
			const double l = lin[t * RC + r];
			uint32_t j = ts.toff[t];
			const uint32_t j1 = ts.toff[t + 1];
			double acc = 0.0;
			for (; j < j1; j += 4) {
				uint32_t c[4];
				double x[4];
#pragma unroll
				for (int q = 0; q < 4; ++q)
					c[q] = ts.tcls[j + q < j1 ? j + q : j];
#pragma unroll
				for (int q = 0; q < 4; ++q)
					x[q] = l * (double)n[c[q] * RC + r] /
					       D[c[q] * RC + r];
#pragma unroll
				for (int q = 0; q < 4; ++q)
					acc += j + q < j1 ? x[q] : 0.0;
			}
			const double nw = ((double)u[t * RC + r] + acc) /
			                  (double)ts.len[t];
			lout[t * RC + r] = nw;
			if (gout)
				gout[(size_t)(tc.tx_base + t) * R + qr] = nw;
			if (stopf && nw > smin && fabs(nw - l) / nw > srel)
				stopf[r] = 1;
		}
	}
	__syncthreads();
}

template <typename TC, int RC>
__device__ __forceinline__ void carve(void *smem, const hemk_tiles &tl,
                                      double *&lam0, double *&lam1,
                                      double *&lam2, double *&D, TC *&n,
                                      uint32_t *&u, TileStruct &ts)
{
	const uint32_t mt = tl.max_tx, mc = tl.max_cls, mz = tl.max_nnz;
	lam0 = (double *)smem;
	lam1 = lam0 + mt * RC;
	lam2 = lam1 + mt * RC;
	D = lam2 + mt * RC;
	n = (TC *)(D + mc * RC);
	u = (uint32_t *)(n + mc * RC);
	ts.tx = u + mt * RC;
	ts.len = ts.tx + mt;
	ts.kmem = ts.len + mt;
	ts.coff = (uint16_t *)(ts.kmem + mz);
	ts.toff = ts.coff + (mc + 1);
	ts.tcls = ts.toff + (mt + 1);
}

// Rows of one replicate chunk held in registers from the prefetch, issued
// while the previous chunk computes, to the store into LDS.  Element e of a
// row block is (entity e / RC, replicate e % RC); thread i holds elements
// i + TBLOCK k for k below the slot count, so its replicate is i % RC.
// Elements beyond the slots (tiles larger than the slots cover) are read
// at the store.
#define PF_TX 2   // slots for transcript rows (n_tx * RC elements)
#define PF_CLS 4  // slots for class count rows (n_cls * RC elements)

template <typename TC, int NS>
struct ChunkRows {
	double s[NS][PF_TX];  // state rows: A, or A B C
	uint32_t u[PF_TX];    // unique counts
	TC n[PF_CLS];         // class counts
	uint32_t q;           // the thread's replicate, NOREP past the list
	double qa, qb;        // ex3: alpha, or the two sums it comes from
};

#define NOREP 0xFFFFFFFFu

// The sample of a tile, its active list, its transcript count, its first
// tile (a single table is sample 0 of ns == 1).
__device__ __forceinline__ uint32_t tile_sample(const hemk_rctl &rc,
                                                uint32_t tile)
{
	return rc.tsamp ? rc.tsamp[tile] : 0u;
}

__device__ __forceinline__ const uint32_t *sample_alist(const hemk_rctl &rc,
                                                        uint32_t smp)
{
	return rc.alist + (size_t)smp * (rc.R + 1);
}

__device__ __forceinline__ uint32_t sample_T(const hemk_rctl &rc,
                                             uint32_t smp, uint32_t T)
{
	return rc.sinfo ? rc.sinfo[4 * smp + 1] : T;
}

// Lanes per entity for chunk i of an active list of nrep replicates: the
// smallest power of two that holds the chunk's replicates, at most RC.
template <int RC>
__device__ __forceinline__ uint32_t chunk_lanes(uint32_t nrep, uint32_t i)
{
	const uint32_t left = nrep - i * RC;
	uint32_t l = 1;
	while (l < RC && l < left)
		l <<= 1;
	return l;
}

// The replicate of this thread's lane in chunk i of an active list.
template <int RC>
__device__ __forceinline__ uint32_t lane_rep(const uint32_t *al,
                                             uint32_t nrep, uint32_t i)
{
	const uint32_t j = i * RC + threadIdx.x % RC;
	return j < nrep ? al[1 + j] : NOREP;
}

template <typename TC, int RC, int NS>
__device__ __forceinline__ void load_rows(ChunkRows<TC, NS> &p,
                                          const TileCtx &tc,
                                          const hemk_rctl &rc,
                                          const uint32_t *al, uint32_t nrep,
                                          uint32_t i,
                                          const TC *__restrict__ cnt,
                                          const uint32_t *__restrict__ uniq,
                                          const double *s0, const double *s1,
                                          const double *s2)
{
	const uint32_t q = lane_rep<RC>(al, nrep, i);
	const bool rin = q != NOREP;
	p.q = q;
#pragma unroll
	for (int k = 0; k < PF_TX; ++k) {
		const uint32_t t = (threadIdx.x + TBLOCK * k) / RC;
		double v0 = 0, v1 = 0, v2 = 0;
		uint32_t w = 0;
		if (rin && t < tc.nt) {
			const size_t g = (size_t)(tc.tx_base + t) * rc.R + q;
			v0 = s0[g];
			if (NS == 3) {
				v1 = s1[g];
				v2 = s2[g];
			}
			w = uniq[g];
		}
		p.s[0][k] = v0;
		if (NS == 3) {
			p.s[1 % NS][k] = v1;
			p.s[2 % NS][k] = v2;
		}
		p.u[k] = w;
	}
#pragma unroll
	for (int k = 0; k < PF_CLS; ++k) {
		const uint32_t c = (threadIdx.x + TBLOCK * k) / RC;
		TC v = 0;
		if (rin && c < tc.nc)
			v = cnt[(size_t)(tc.cls_base + c) * rc.R + q];
		p.n[k] = v;
	}
}

// Class and unique count rows of the chunk into LDS (n, u); q is the
// lane's replicate.
template <typename TC, int RC, int NS>
__device__ __forceinline__ void store_counts(const ChunkRows<TC, NS> &p,
                                             const TileCtx &tc,
                                             const hemk_rctl &rc, uint32_t q,
                                             const TC *__restrict__ cnt,
                                             const uint32_t *__restrict__ uniq,
                                             TC *n, uint32_t *u)
{
	const bool rin = q != NOREP;
#pragma unroll
	for (int k = 0; k < PF_TX; ++k) {
		const uint32_t e = threadIdx.x + TBLOCK * k;
		if (e / RC < tc.nt)
			u[e] = p.u[k];
	}
	for (uint32_t e = threadIdx.x + TBLOCK * PF_TX; e / RC < tc.nt;
	     e += TBLOCK)
		u[e] = rin ? uniq[(size_t)(tc.tx_base + e / RC) * rc.R + q] : 0u;
#pragma unroll
	for (int k = 0; k < PF_CLS; ++k) {
		const uint32_t e = threadIdx.x + TBLOCK * k;
		if (e / RC < tc.nc)
			n[e] = p.n[k];
	}
	for (uint32_t e = threadIdx.x + TBLOCK * PF_CLS; e / RC < tc.nc;
	     e += TBLOCK)
		n[e] = rin ? cnt[(size_t)(tc.cls_base + e / RC) * rc.R + q] : (TC)0;
}

// The r and v of one replicate chunk from u, u1, u2 in LDS (src/EM.c:
// 334-335): as replicate-major rows for the exact or sequential norms
// (xr, xv), or, with part, as per (tile, replicate) partials for the tree
// norms: the local two-pass mean and centred sum of squares of r and v
// (combined across tiles by Chan's rule), or plain sums of squares for the
// L2 step.
template <typename TC, int RC>
__device__ __forceinline__ void rv_parts(const hemk_tiles &tl,
                                         const TileCtx &tc,
                                         const TileStruct &ts, uint32_t tile,
                                         const double *lam0,
                                         const double *lam1,
                                         const double *lam2, uint32_t q,
                                         double *xr, double *xv, double *part,
                                         uint32_t T, int step)
{
#pragma clang fp contract(off)
	const int r = threadIdx.x % RC;
	const uint32_t e0 = threadIdx.x / RC;
	constexpr uint32_t ES = TBLOCK / RC;
	const bool rin = q != NOREP;
	if (xr && rin) {
		// r_t and v_t (src/EM.c:334-335) in replicate-major rows
		// for the sequential norms
		for (uint32_t t = e0; t < tc.nt; t += ES) {
			const double x = lam1[t * RC + r] - lam0[t * RC + r];
			const double y = (lam2[t * RC + r] - lam1[t * RC + r]) - x;
			const size_t g = (size_t)q * T + ts.tx[t];
			xr[g] = x;
			xv[g] = y;
		}
	}
	if (part && threadIdx.x < RC && rin) {
		// per (tile, replicate): local two-pass mean and centred
		// sum of squares of r and v (combined across tiles by
		// Chan's rule), or plain sums of squares for the L2 step
		const int lq = threadIdx.x;
		double a0, a1, b0, b1;
		if (step == 0) {
			double sr = 0, sv = 0;
			for (uint32_t t = 0; t < tc.nt; ++t) {
				const double x = lam1[t * RC + lq] - lam0[t * RC + lq];
				const double y =
				    (lam2[t * RC + lq] - lam1[t * RC + lq]) - x;
				sr += x;
				sv += y;
			}
			const double mr = sr / tc.nt, mv = sv / tc.nt;
			double qr = 0, qv = 0;
			for (uint32_t t = 0; t < tc.nt; ++t) {
				const double x = lam1[t * RC + lq] - lam0[t * RC + lq];
				const double y =
				    (lam2[t * RC + lq] - lam1[t * RC + lq]) - x;
				qr += (x - mr) * (x - mr);
				qv += (y - mv) * (y - mv);
			}
			a0 = mr;
			a1 = qr;
			b0 = mv;
			b1 = qv;
		} else {
			double qr = 0, qv = 0;
			for (uint32_t t = 0; t < tc.nt; ++t) {
				const double x = lam1[t * RC + lq] - lam0[t * RC + lq];
				const double y =
				    (lam2[t * RC + lq] - lam1[t * RC + lq]) - x;
				qr += x * x;
				qv += y * y;
			}
			a0 = 0;
			a1 = qr;
			b0 = 0;
			b1 = qv;
		}
		double *pp = part + ((size_t)q * tl.n_tiles + tile) * 4;
		pp[0] = a0;
		pp[1] = a1;
		pp[2] = b0;
		pp[3] = b1;
	}
}

template <typename TC, int RC, bool F>
__global__ __launch_bounds__(TBLOCK) void k_tile_map12(
    hemk_tiles tl, hemk_rctl rc, const TC *__restrict__ cnt,
    const uint32_t *__restrict__ uniq, const double *__restrict__ A,
    double *__restrict__ B, double *__restrict__ C, double *__restrict__ part,
    double *__restrict__ xr, double *__restrict__ xv, uint32_t T, int step)
{
#pragma clang fp contract(off)
	extern __shared__ double smem[];
	__shared__ int alive[RC];
	const uint32_t tile = blockIdx.x;
	const uint32_t *al = sample_alist(rc, tile_sample(rc, tile));
	const uint32_t nrep = al[0];
	if (nrep == 0)
		return;
	const uint32_t nch = (nrep + RC - 1) / RC;
	const TileCtx tc = load_tile(tl, tile);
	double *lam0, *lam1, *lam2, *D;
	TC *n;
	uint32_t *u;
	TileStruct ts;
	carve<TC, RC>(smem, tl, lam0, lam1, lam2, D, n, u, ts);
	ChunkRows<TC, 1> p;
	load_rows<TC, RC, 1>(p, tc, rc, al, nrep, 0, cnt, uniq, A, nullptr,
	                     nullptr);
	stage_struct(tl, tc, ts);
	for (uint32_t i = 0; i < nch; ++i) {
		const uint32_t q = p.q;
		const bool rin = q != NOREP;
#pragma unroll
		for (int k = 0; k < PF_TX; ++k) {
			const uint32_t e = threadIdx.x + TBLOCK * k;
			if (e / RC < tc.nt)
				lam0[e] = p.s[0][k];
		}
		for (uint32_t e = threadIdx.x + TBLOCK * PF_TX; e / RC < tc.nt;
		     e += TBLOCK)
			lam0[e] = rin ? A[(size_t)(tc.tx_base + e / RC) * rc.R + q]
			              : 0.0;
		store_counts<TC, RC, 1>(p, tc, rc, q, cnt, uniq, n, u);
		if (threadIdx.x < RC)
			alive[threadIdx.x] = rin ? 1 : 0;
		__syncthreads();
		if (i + 1 < nch)
			load_rows<TC, RC, 1>(p, tc, rc, al, nrep, i + 1, cnt, uniq,
			                     A, nullptr, nullptr);
		const uint32_t *ql = al + 1 + (size_t)i * RC;
		const uint32_t rcl = chunk_lanes<RC>(nrep, i);
		tile_map<TC, RC, F>(tl, tc, ts, lam0, lam1, D, n, u, alive, B,
		                    rc.R, ql, rcl, nullptr, 0, 0);
		tile_map<TC, RC, F>(tl, tc, ts, lam1, lam2, D, n, u, alive, C,
		                    rc.R, ql, rcl, nullptr, 0, 0);
		rv_parts<TC, RC>(tl, tc, ts, tile, lam0, lam1, lam2, q, xr, xv,
		                 part, T, step);
		__syncthreads();
	}
}

__device__ __forceinline__ double step_len(double A, double B, uint32_t n,
                                           hemk_step st)
{
#pragma clang fp contract(off)
	double n_r, n_v, alpha;
	if (st.step == 0) {
		n_r = sqrt(A / (double)(n - 1u));
		n_v = sqrt(B / (double)(n - 1u));
	} else {
		n_r = sqrt(A);
		n_v = B <= 1e-300 ? 0.0 : sqrt(B);
	}
	alpha = n_v != 0 ? -n_r / n_v : -1.0;
	if (st.alpha_clamp && alpha > -1)
		alpha = -1;
	return alpha;
}

// u' = u - 2 alpha r + alpha^2 v, clamped (src/EM.c:338-360); rounds as the
// reference (no contraction).
__device__ __forceinline__ double extrapolate(double a, double b, double c2,
                                              double alpha, hemk_step st)
{
#pragma clang fp contract(off)
	const double rr = b - a;
	const double vv = (c2 - b) - rr;
	double x = a - 2 * alpha * rr + alpha * alpha * vv;
	if (st.clamp == 0) {
		if (x < 0)
			x = 0;
	} else if (x <= 1e-300) {
		x = c2 / 10;
	}
	return x;
}

// The step length inputs of the lane's replicate (after load_rows).
template <typename TC>
__device__ __forceinline__ void load_alpha(ChunkRows<TC, 3> &p,
                                           const hemk_rctl &rc, uint32_t smp,
                                           hemk_step st)
{
	const uint32_t q = p.q;
	p.qa = 0;
	p.qb = 0;
	if (q != NOREP) {
		const uint32_t in = smp * rc.R + q;
		if (st.alpha_given) {
			p.qa = rc.alpha[in];
		} else {
			p.qa = rc.q_r[in];
			p.qb = rc.q_v[in];
		}
	}
}

template <typename TC, int RC, bool F>
// four waves per SIMD (at most 128 VGPRs), as many as the LDS budget allows
__global__ __launch_bounds__(TBLOCK, 4) void k_tile_ex3(
    hemk_tiles tl, hemk_rctl rc, const TC *__restrict__ cnt,
    const uint32_t *__restrict__ uniq, double *__restrict__ A,
    double *__restrict__ B, double *__restrict__ C, uint32_t T,
    hemk_step st, int next, double *__restrict__ part,
    double *__restrict__ xr, double *__restrict__ xv)
{
#pragma clang fp contract(off)
	extern __shared__ double smem[];
	__shared__ int alive[RC];
	__shared__ int stopf[RC];
	const uint32_t tile = blockIdx.x;
	const uint32_t smp = tile_sample(rc, tile);
	const uint32_t *al = sample_alist(rc, smp);
	const uint32_t nrep = al[0];
	if (nrep == 0)
		return;
	const uint32_t Ts = sample_T(rc, smp, T);
	const uint32_t tile0 = rc.sinfo ? rc.sinfo[4 * smp + 2] : 0u;
	const uint32_t nch = (nrep + RC - 1) / RC;
	const TileCtx tc = load_tile(tl, tile);
	double *lam0, *lam1, *lam2, *D;
	TC *n;
	uint32_t *u;
	TileStruct ts;
	carve<TC, RC>(smem, tl, lam0, lam1, lam2, D, n, u, ts);
	ChunkRows<TC, 3> p;
	load_rows<TC, RC, 3>(p, tc, rc, al, nrep, 0, cnt, uniq, A, B, C);
	load_alpha<TC>(p, rc, smp, st);
	stage_struct(tl, tc, ts);
	for (uint32_t i = 0; i < nch; ++i) {
		const uint32_t q = p.q;
		const bool rin = q != NOREP;
		// sequential norms leave the two sums; the parallel ones alpha
		const double alpha = st.alpha_given
		                         ? p.qa
		                         : step_len(p.qa, p.qb, Ts, st);
		if (!st.alpha_given && tile == tile0 && threadIdx.x < RC && rin)
			rc.alpha[smp * rc.R + q] = alpha;
#pragma unroll
		for (int k = 0; k < PF_TX; ++k) {
			const uint32_t e = threadIdx.x + TBLOCK * k;
			if (e / RC < tc.nt)
				lam0[e] = extrapolate(p.s[0][k], p.s[1][k], p.s[2][k],
				                      alpha, st);
		}
		for (uint32_t e = threadIdx.x + TBLOCK * PF_TX; e / RC < tc.nt;
		     e += TBLOCK) {
			double x = 0.0;
			if (rin) {
				const size_t g = (size_t)(tc.tx_base + e / RC) * rc.R + q;
				x = extrapolate(A[g], B[g], C[g], alpha, st);
			}
			lam0[e] = x;
		}
		store_counts<TC, RC, 3>(p, tc, rc, q, cnt, uniq, n, u);
		if (threadIdx.x < RC) {
			alive[threadIdx.x] = rin ? 1 : 0;
			stopf[threadIdx.x] = 0;
		}
		__syncthreads();
		if (i + 1 < nch) {
			load_rows<TC, RC, 3>(p, tc, rc, al, nrep, i + 1, cnt, uniq,
			                     A, B, C);
			load_alpha<TC>(p, rc, smp, st);
		}
		const uint32_t *ql = al + 1 + (size_t)i * RC;
		const uint32_t rcl = chunk_lanes<RC>(nrep, i);
		tile_map<TC, RC, F>(tl, tc, ts, lam0, lam1, D, n, u, alive, A,
		                    rc.R, ql, rcl, stopf, rc.stop_min,
		                    rc.stop_rel);
		if (threadIdx.x < RC && rin && stopf[threadIdx.x])
			rc.notconv[smp * rc.R + q] = 1;
		if (next) {
			// the next iteration's u1 -> B, u2 -> C (into lam0, whose
			// extrapolated values are spent) and its r, v
			tile_map<TC, RC, F>(tl, tc, ts, lam1, lam2, D, n, u, alive, B,
			                    rc.R, ql, rcl, nullptr, 0, 0);
			tile_map<TC, RC, F>(tl, tc, ts, lam2, lam0, D, n, u, alive, C,
			                    rc.R, ql, rcl, nullptr, 0, 0);
			rv_parts<TC, RC>(tl, tc, ts, tile, lam1, lam2, lam0, q, xr,
			                 xv, part, T, st.step);
		}
		__syncthreads();
	}
}

// (n, mean, M2) folding of Chan, Golub & LeVeque (1979).
// One block per instance: fold the tile partials of its sample in tile
// order.  Partials are replicate-major (part[(q * n_tiles + tile) * 4]), so
// each thread reads its run of tiles as one contiguous stream.
__global__ __launch_bounds__(TBLOCK) void k_batch_alpha(
    hemk_tiles tl, hemk_rctl rc, const double *__restrict__ part, uint32_t T,
    hemk_step st)
{
	const uint32_t in = blockIdx.x;
	const uint32_t smp = in / rc.R, r = in % rc.R;
	if (rc.done[in])
		return;
	__shared__ Moments sr[TBLOCK], sv[TBLOCK];
	const uint32_t tb = rc.sinfo ? rc.sinfo[4 * smp + 2] : 0u;
	const uint32_t nt = rc.sinfo ? rc.sinfo[4 * smp + 3] : tl.n_tiles;
	const uint32_t per = (nt + TBLOCK - 1) / TBLOCK;
	const uint32_t i0 = tb + threadIdx.x * per;
	const uint32_t i1 = min(tb + nt, i0 + per);
	Moments a = {0, 0, 0}, b = {0, 0, 0};
	for (uint32_t i = i0; i < i1; ++i) {
		const double nt = (double)tl.desc[(size_t)HEMK_TILE_DESC * i + 1];
		const double *pp = part + ((size_t)r * tl.n_tiles + i) * 4;
		if (st.step == 0) {
			a = mom_merge(a, Moments{nt, pp[0], pp[1]});
			b = mom_merge(b, Moments{nt, pp[2], pp[3]});
		} else {
			a.q += pp[1];
			b.q += pp[3];
		}
	}
	sr[threadIdx.x] = a;
	sv[threadIdx.x] = b;
	__syncthreads();
	for (uint32_t w = TBLOCK / 2; w >= 1; w >>= 1) {
		if (threadIdx.x < w) {
			if (st.step == 0) {
				sr[threadIdx.x] = mom_merge(sr[threadIdx.x],
				                            sr[threadIdx.x + w]);
				sv[threadIdx.x] = mom_merge(sv[threadIdx.x],
				                            sv[threadIdx.x + w]);
			} else {
				sr[threadIdx.x].q += sr[threadIdx.x + w].q;
				sv[threadIdx.x].q += sv[threadIdx.x + w].q;
			}
		}
		__syncthreads();
	}
	if (threadIdx.x == 0)
		rc.alpha[in] = step_len(sr[0].q, sv[0].q, sample_T(rc, smp, T), st);
}

// acc + b[0] + b[1] + ... + b[cnt-1], left to right.  Two register banks of
// sixteen values alternate: one is read from LDS while the other is added,
// so the chain runs at the latency of a dependent FP64 add rather than of
// an LDS read, with no register copies in the loop.
__device__ __forceinline__ double chain_add(double acc, const double *b,
                                            uint32_t cnt)
{
#pragma clang fp contract(off)
	uint32_t j = 0;
	if (cnt >= 16) {
		double a[16], c[16];
#pragma unroll
		for (int q = 0; q < 16; ++q)
			a[q] = b[q];
		// a holds b[j-16, j)
		for (j = 16; j + 32 <= cnt; j += 32) {
#pragma unroll
			for (int q = 0; q < 16; ++q)
				c[q] = b[j + q];
#pragma unroll
			for (int q = 0; q < 16; ++q)
				acc += a[q];
#pragma unroll
			for (int q = 0; q < 16; ++q)
				a[q] = b[j + 16 + q];
#pragma unroll
			for (int q = 0; q < 16; ++q)
				acc += c[q];
		}
		if (j + 16 <= cnt) {
#pragma unroll
			for (int q = 0; q < 16; ++q)
				c[q] = b[j + q];
#pragma unroll
			for (int q = 0; q < 16; ++q)
				acc += a[q];
#pragma unroll
			for (int q = 0; q < 16; ++q)
				acc += c[q];
			j += 16;
		} else {
#pragma unroll
			for (int q = 0; q < 16; ++q)
				acc += a[q];
		}
	}
	for (; j < cnt; ++j)
		acc += b[j];
	return acc;
}

// Sequential sums of a batch over replicate-major rows: block (replicate,
// row) runs one left-to-right chain on lane 0 of wave 0 while waves 1-3
// stage the next chunk of element values.
//   phase 0: lam_sum = sum x      phase 1: mean = (sum x)/T
//   phase 2: q = sum (x - mean)^2  phase 3: q = sum x^2 (L2 step)
#define CHAIN_TB 2048
__global__ __launch_bounds__(TBLOCK) void k_batch_chain(
    hemk_rctl rc, uint32_t T, const double *__restrict__ xr,
    const double *__restrict__ xv, int phase)
{
#pragma clang fp contract(off)
	__shared__ double buf[2][CHAIN_TB];
	const uint32_t in = blockIdx.x;
	const uint32_t smp = in / rc.R, rr = in % rc.R;
	const int which = blockIdx.y;
	if (phase != 0 && rc.done[in])
		return;
	const uint32_t t0 = rc.sinfo ? rc.sinfo[4 * smp] : 0u;
	const uint32_t Tr = T;
	T = sample_T(rc, smp, T);
	const double *x = (which ? xv : xr) + (size_t)rr * Tr + t0;
	const uint32_t r = in;
	const double m = phase == 2 ? (which ? rc.mean_v[r] : rc.mean_r[r]) : 0.0;
	const int w = threadIdx.x / 64;
	const uint32_t nch = (T + CHAIN_TB - 1) / CHAIN_TB;
	double acc = 0;
	for (uint32_t ch = 0; ch <= nch; ++ch) {
		if (w >= 1 && ch < nch) {
			const int slot = ch & 1;
			const uint32_t base = ch * CHAIN_TB;
			for (int j = threadIdx.x - 64; j < CHAIN_TB; j += TBLOCK - 64) {
				const uint32_t t = base + j;
				double a = 0;
				if (t < T) {
					const double v = x[t];
					a = phase <= 1 ? v : phase == 2 ? (v - m) * (v - m)
					                                : v * v;
				}
				buf[slot][j] = a;
			}
		}
		if (threadIdx.x == 0 && ch >= 1) {
			const double *b = buf[(ch - 1) & 1];
			const uint32_t cnt = min((uint32_t)CHAIN_TB,
			                         T - (ch - 1) * CHAIN_TB);
			acc = chain_add(acc, b, cnt);
		}
		__syncthreads();
	}
	if (threadIdx.x == 0) {
		if (phase == 0) {
			rc.lam_sum[r] = acc;
		} else if (phase == 1) {
			if (which)
				rc.mean_v[r] = acc / T;
			else
				rc.mean_r[r] = acc / T;
		} else {
			if (which)
				rc.q_v[r] = acc;
			else
				rc.q_r[r] = acc;
		}
	}
}

// alist = the replicates not done, ascending, after alist[0] = their count;
// the tile kernels deal them to chunks of rc.chunk lanes, so finished
// replicates cost nothing.  One block; done[] written by this block before
// the call.
__device__ void build_active_list(const hemk_rctl &rc, uint32_t smp)
{
	__shared__ uint32_t sc[TBLOCK];
	uint32_t *al = rc.alist + (size_t)smp * (rc.R + 1);
	const uint32_t *done = rc.done + (size_t)smp * rc.R;
	uint32_t total = 0;
	for (uint32_t base = 0; base < rc.R; base += TBLOCK) {
		const uint32_t r = base + threadIdx.x;
		const uint32_t f = (r < rc.R && !done[r]) ? 1u : 0u;
		sc[threadIdx.x] = f;
		__syncthreads();
		// inclusive scan of the flags
		for (uint32_t w = 1; w < TBLOCK; w <<= 1) {
			const uint32_t v = threadIdx.x >= w ? sc[threadIdx.x - w] : 0u;
			__syncthreads();
			sc[threadIdx.x] += v;
			__syncthreads();
		}
		if (f)
			al[1 + total + sc[threadIdx.x] - 1] = r;
		total += sc[TBLOCK - 1];
		__syncthreads();
	}
	if (threadIdx.x == 0)
		al[0] = total;
}

// One block per sample: end the iteration of its instances.
__global__ __launch_bounds__(TBLOCK) void k_batch_iter_end(hemk_rctl rc)
{
	__shared__ uint32_t cnt[TBLOCK];
	const uint32_t smp = blockIdx.x;
	uint32_t act = 0;
	for (uint32_t rr = threadIdx.x; rr < rc.R; rr += TBLOCK) {
		const uint32_t r = smp * rc.R + rr;
		if (!rc.done[r]) {
			const uint32_t k = rc.k[r] + 1;
			const uint32_t stop = rc.notconv[r] ? 0u : 1u;
			rc.k[r] = k;
			rc.notconv[r] = 0;
			if (!(k < rc.min_iter || (k < rc.max_iter && !stop))) {
				rc.done[r] = 1;
				rc.conv[r] = stop;
			} else {
				++act;
			}
		}
	}
	cnt[threadIdx.x] = act;
	__syncthreads();
	for (uint32_t w = TBLOCK / 2; w >= 1; w >>= 1) {
		if (threadIdx.x < w)
			cnt[threadIdx.x] += cnt[threadIdx.x + w];
		__syncthreads();
	}
	if (threadIdx.x == 0)
		rc.active[smp] = cnt[0];
	build_active_list(rc, smp);
}

__global__ __launch_bounds__(TBLOCK) void k_batch_reset(hemk_rctl rc,
                                                        uint32_t T,
                                                        double *__restrict__ A)
{
	const size_t n = (size_t)T * rc.R;
	// the uniform start 1 / n_tx of each sample (src/EM.c:311-312)
	for (size_t i = (size_t)blockIdx.x * TBLOCK + threadIdx.x; i < n;
	     i += (size_t)gridDim.x * TBLOCK) {
		const uint32_t q = (uint32_t)(i / rc.R);
		const uint32_t smp = rc.txsamp ? rc.txsamp[q] : 0u;
		A[i] = (double)1 / sample_T(rc, smp, T);
	}
	if (blockIdx.x == 0) {
		// the loop condition of src/EM.c:328 at k = 0, stop = 0
		const uint32_t run = (0 < rc.min_iter || 0 < rc.max_iter) ? 1u : 0u;
		const uint32_t ni = rc.ns * rc.R;
		for (uint32_t r = threadIdx.x; r < ni; r += TBLOCK) {
			rc.k[r] = 0;
			rc.conv[r] = 0;
			rc.notconv[r] = 0;
			rc.alpha[r] = 0;
			rc.lam_sum[r] = 0;
			rc.done[r] = run ? 0u : 1u;
		}
		for (uint32_t smp = threadIdx.x; smp < rc.ns; smp += TBLOCK)
			rc.active[smp] = run ? rc.R : 0;
		__syncthreads();
		for (uint32_t smp = 0; smp < rc.ns; ++smp)
			build_active_list(rc, smp);
	}
}

// X[r][t] = A[txq[t]][r] through a 32x32 LDS transpose.
__global__ __launch_bounds__(TBLOCK) void k_batch_rows(
    uint32_t R, uint32_t T, const uint32_t *__restrict__ txq,
    const double *__restrict__ A, double *__restrict__ X)
{
	__shared__ double tile[32][33];
	const uint32_t t0 = blockIdx.x * 32, q0 = blockIdx.y * 32;
	// load A[txq[t0+i]][q0+j]: j fastest
	for (uint32_t k = threadIdx.x; k < 32 * 32; k += TBLOCK) {
		const uint32_t i = k / 32, j = k % 32;
		const uint32_t t = t0 + i, q = q0 + j;
		tile[i][j] = (t < T && q < R) ? A[(size_t)txq[t] * R + q] : 0.0;
	}
	__syncthreads();
	// store X[q0+j][t0+i]: i fastest
	for (uint32_t k = threadIdx.x; k < 32 * 32; k += TBLOCK) {
		const uint32_t j = k / 32, i = k % 32;
		const uint32_t t = t0 + i, q = q0 + j;
		if (t < T && q < R)
			X[(size_t)q * T + t] = tile[i][j];
	}
}

// est and tpm of src/EM.c:381-387 for row r = blockIdx.y.
__global__ __launch_bounds__(TBLOCK) void k_batch_out(
    hemk_rctl rc, uint32_t T, const double *__restrict__ X,
    const uint32_t *__restrict__ len, double *__restrict__ est,
    double *__restrict__ tpm)
{
	const uint32_t r = blockIdx.y;
	for (uint32_t t = blockIdx.x * TBLOCK + threadIdx.x; t < T;
	     t += gridDim.x * TBLOCK) {
		const size_t i = (size_t)r * T + t;
		const uint32_t smp = rc.txsamp ? rc.txsamp[t] : 0u;
		const double sum = rc.lam_sum[smp * rc.R + r];
		const double l = X[i];
		double e = 0, g = 0;
		if (sum != 0) {
			e = l * (double)len[t];
			g = l * 1000000 / sum;
		}
		if (est)
			est[i] = e;
		if (tpm)
			tpm[i] = g;
	}
}

__global__ __launch_bounds__(TBLOCK) void k_transpose_u32(
    const uint32_t *__restrict__ in, uint32_t rows, uint32_t cols,
    uint32_t *__restrict__ out)
{
	// in: rows x cols (row-major) -> out: cols x rows
	__shared__ uint32_t tile[32][33];
	const uint32_t c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
	for (uint32_t k = threadIdx.x; k < 32 * 32; k += TBLOCK) {
		const uint32_t i = k / 32, j = k % 32;
		const uint32_t rr = r0 + i, cc = c0 + j;
		tile[i][j] = (rr < rows && cc < cols) ? in[(size_t)rr * cols + cc]
		                                      : 0u;
	}
	__syncthreads();
	for (uint32_t k = threadIdx.x; k < 32 * 32; k += TBLOCK) {
		const uint32_t j = k / 32, i = k % 32;
		const uint32_t rr = r0 + i, cc = c0 + j;
		if (rr < rows && cc < cols)
			out[(size_t)cc * rows + rr] = tile[i][j];
	}
}

inline uint32_t cdiv32(uint64_t a, uint64_t b)
{
	return (uint32_t)((a + b - 1) / b);
}

size_t tile_lds(const hemk_tiles *t, int rc, int cnt_bytes)
{
	return (size_t)rc * ((size_t)t->max_tx * (3 * 8 + 4) +
	                     (size_t)t->max_cls * (8 + cnt_bytes)) +
	       (size_t)t->max_nnz * (4 + 2) + (size_t)t->max_tx * 8 +
	       2 * ((size_t)t->max_cls + 1) + 2 * ((size_t)t->max_tx + 1) + 16;
}

template <int RC, bool F>
int launch_map12(const hemk_tiles *t, hemk_rctl rc, const uint32_t *cnt,
                 const uint32_t *uniq, const double *A, double *B, double *C,
                 double *part, double *xr, double *xv, uint32_t T, int step,
                 hipStream_t s)
{
	const size_t lds = tile_lds(t, RC, sizeof(uint32_t));
	if (rc.chunk != (uint32_t)RC)
		return (int)hipErrorInvalidValue;
	k_tile_map12<uint32_t, RC, F><<<t->n_tiles, TBLOCK, lds, s>>>(
	    *t, rc, cnt, uniq, A, B, C, part, xr, xv, T, step);
	return (int)hipGetLastError();
}

template <int RC, bool F>
int launch_ex3(const hemk_tiles *t, hemk_rctl rc, const uint32_t *cnt,
               const uint32_t *uniq, double *A, double *B, double *C,
               uint32_t T, hemk_step st, int next, double *part, double *xr,
               double *xv, hipStream_t s)
{
	const size_t lds = tile_lds(t, RC, sizeof(uint32_t));
	if (rc.chunk != (uint32_t)RC)
		return (int)hipErrorInvalidValue;
	k_tile_ex3<uint32_t, RC, F><<<t->n_tiles, TBLOCK, lds, s>>>(
	    *t, rc, cnt, uniq, A, B, C, T, st, next, part, xr, xv);
	return (int)hipGetLastError();
}

#define HEM_RC_SWITCH(chunk, CALL)                                          \
	switch (chunk) {                                                    \
	case 1:                                                             \
		return CALL(1);                                             \
	case 2:                                                             \
		return CALL(2);                                             \
	case 4:                                                             \
		return CALL(4);                                             \
	case 8:                                                             \
		return CALL(8);                                             \
	case 16:                                                            \
		return CALL(16);                                            \
	case 32:                                                            \
		return CALL(32);                                            \
	default:                                                            \
		return (int)hipErrorInvalidValue;                           \
	}

} // namespace

extern "C" {

int hemk_tile_chunk(const hemk_tiles *t, uint32_t R, int cnt_bytes)
{
	int rc = 32;
	while (rc > 1 && ((uint32_t)rc / 2 >= R ||
	                  tile_lds(t, rc, cnt_bytes) > TLDS_BUDGET))
		rc /= 2;
	return tile_lds(t, rc, cnt_bytes) > TLDS_BUDGET ? 0 : rc;
}

int hemk_tile_map12(const hemk_tiles *t, hemk_rctl rc, int chunk,
                    const uint32_t *cnt, int f32, const uint32_t *uniq,
                    const double *A, double *B, double *C, double *part,
                    double *xr, double *xv, uint32_t T, int step, void *s)
{
	hipStream_t st = (hipStream_t)s;
	if (t->n_tiles == 0)
		return 0;
	if (f32) {
#define M12F(RC) launch_map12<RC, true>(t, rc, cnt, uniq, A, B, C, part, xr, xv, T, step, st)
		HEM_RC_SWITCH(chunk, M12F)
#undef M12F
	}
#define M12D(RC) launch_map12<RC, false>(t, rc, cnt, uniq, A, B, C, part, xr, xv, T, step, st)
	HEM_RC_SWITCH(chunk, M12D)
#undef M12D
}

int hemk_tile_ex3(const hemk_tiles *t, hemk_rctl rc, int chunk,
                  const uint32_t *cnt, int f32, const uint32_t *uniq,
                  double *A, double *B, double *C, uint32_t T,
                  hemk_step st, int next, double *part, double *xr,
                  double *xv, void *s)
{
	hipStream_t sm = (hipStream_t)s;
	if (t->n_tiles == 0)
		return 0;
	if (f32) {
#define EXF(RC) launch_ex3<RC, true>(t, rc, cnt, uniq, A, B, C, T, st, next, part, xr, xv, sm)
		HEM_RC_SWITCH(chunk, EXF)
#undef EXF
	}
#define EXD(RC) launch_ex3<RC, false>(t, rc, cnt, uniq, A, B, C, T, st, next, part, xr, xv, sm)
	HEM_RC_SWITCH(chunk, EXD)
#undef EXD
}

int hemk_batch_alpha(const hemk_tiles *t, hemk_rctl rc, const double *part,
                     uint32_t T, hemk_step st, void *s)
{
	k_batch_alpha<<<rc.ns * rc.R, TBLOCK, 0, (hipStream_t)s>>>(*t, rc, part,
	                                                           T, st);
	return (int)hipGetLastError();
}

int hemk_batch_chain(hemk_rctl rc, uint32_t T, const double *xr,
                     const double *xv, int phase, void *s)
{
	dim3 grid(rc.ns * rc.R, phase == 0 ? 1 : 2);
	k_batch_chain<<<grid, TBLOCK, 0, (hipStream_t)s>>>(rc, T, xr, xv, phase);
	return (int)hipGetLastError();
}

int hemk_batch_iter_end(hemk_rctl rc, void *s)
{
	k_batch_iter_end<<<rc.ns, TBLOCK, 0, (hipStream_t)s>>>(rc);
	return (int)hipGetLastError();
}

int hemk_batch_reset(hemk_rctl rc, uint32_t T, double *A, void *s)
{
	uint32_t grid = cdiv32((uint64_t)T * rc.R, TBLOCK);
	if (grid > 8192)
		grid = 8192;
	k_batch_reset<<<grid ? grid : 1, TBLOCK, 0, (hipStream_t)s>>>(rc, T, A);
	return (int)hipGetLastError();
}

int hemk_batch_rows(hemk_rctl rc, uint32_t T, const uint32_t *txq,
                    const double *A, double *X, void *s)
{
	if (!T || !rc.R)
		return 0;
	dim3 grid(cdiv32(T, 32), cdiv32(rc.R, 32));
	k_batch_rows<<<grid, TBLOCK, 0, (hipStream_t)s>>>(rc.R, T, txq, A, X);
	return (int)hipGetLastError();
}

int hemk_batch_out(hemk_rctl rc, uint32_t T, const double *X,
                   const uint32_t *len, double *est, double *tpm, void *s)
{
	if (!T || !rc.R)
		return 0;
	uint32_t gx = cdiv32(T, TBLOCK);
	if (gx > 1024)
		gx = 1024;
	dim3 grid(gx, rc.R);
	k_batch_out<<<grid, TBLOCK, 0, (hipStream_t)s>>>(rc, T, X, len, est, tpm);
	return (int)hipGetLastError();
}

int hemk_transpose_u32(const uint32_t *in, uint32_t rows, uint32_t cols,
                       uint32_t *out, void *s)
{
	if (!rows || !cols)
		return 0;
	dim3 grid(cdiv32(cols, 32), cdiv32(rows, 32));
	k_transpose_u32<<<grid, TBLOCK, 0, (hipStream_t)s>>>(in, rows, cols, out);
	return (int)hipGetLastError();
}

} // extern "C"
