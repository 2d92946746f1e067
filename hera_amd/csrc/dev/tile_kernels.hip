// tile_kernels.hip -- batched SQUAREM EM over component tiles (gfx950).
//
// A workgroup owns one tile (whole connected components) and a chunk of RC
// replicates.  It stages the tile's state rows ([t][r], replicate fastest,
// one coalesced row per transcript), unique counts and class counts in LDS,
// then runs hera's class pass and transcript pass (src/EM.c:276-291) on LDS
// only, in the reference's order: members in class order for D_c, classes
// in ascending original id for each transcript's sum of (lam_t n_c) / D_c.
// The structure is read once per tile and shared by the RC replicates of the
// workgroup; replicates are independent lanes (lane r of an entity group).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "hem_dev.h"

#define TBLOCK 256
#define TLDS_BUDGET (40 * 1024)

namespace {

struct TileCtx {
	uint32_t tx_base, nt, cls_base, nc;
};

__device__ __forceinline__ TileCtx load_tile(const hemk_tiles &t, uint32_t i)
{
	TileCtx c;
	c.tx_base = t.desc[4 * i + 0];
	c.nt = t.desc[4 * i + 1];
	c.cls_base = t.desc[4 * i + 2];
	c.nc = t.desc[4 * i + 3];
	return c;
}

// The tile's structure staged in LDS: member lists and class lists with
// offsets relative to the tile (u16; a tile holds at most max_nnz members).
struct TileStruct {
	uint16_t *coff;  // nc + 1
	uint16_t *toff;  // nt + 1
	uint8_t *cmem;   // members
	uint16_t *tcls;  // class-list entries
};

__device__ __forceinline__ void stage_struct(const hemk_tiles &tl,
                                             const TileCtx &tc,
                                             const TileStruct &ts)
{
	const uint32_t mb = tl.coff[tc.cls_base];
	const uint32_t me = tl.coff[tc.cls_base + tc.nc];
	const uint32_t jb = tl.toff[tc.tx_base];
	const uint32_t je = tl.toff[tc.tx_base + tc.nt];
	for (uint32_t i = threadIdx.x; i <= tc.nc; i += TBLOCK)
		ts.coff[i] = (uint16_t)(tl.coff[tc.cls_base + i] - mb);
	for (uint32_t i = threadIdx.x; i <= tc.nt; i += TBLOCK)
		ts.toff[i] = (uint16_t)(tl.toff[tc.tx_base + i] - jb);
	for (uint32_t i = threadIdx.x; i < me - mb; i += TBLOCK)
		ts.cmem[i] = tl.cmem[mb + i];
	for (uint32_t i = threadIdx.x; i < je - jb; i += TBLOCK)
		ts.tcls[i] = tl.tcls[jb + i];
}

// One EM map of the tile for the block's replicate chunk, lin -> lout in
// LDS (and to the global [T][R] vector gout when given).  With stopf the
// stop test of src/EM.c:292-297 marks stopf[r].  Terms are formed four at a
// time (independent loads and divisions in flight) and added in order.
template <typename TC, int RC>
__device__ __forceinline__ void tile_map(
    const hemk_tiles &tl, const TileCtx &tc, const TileStruct &ts,
    const double *lin, double *lout, double *D, const TC *n,
    const uint32_t *u, const int *alive, double *gout, uint32_t R,
    uint32_t r0, int *stopf, double smin, double srel)
{
	const int r = threadIdx.x % RC;
	const uint32_t e0 = threadIdx.x / RC;
	constexpr uint32_t ES = TBLOCK / RC;
	const bool live = alive[r] != 0;
	// class pass: D_c, left to right over the members
	for (uint32_t c = e0; c < tc.nc; c += ES) {
		if (!live)
			continue;
		uint32_t k = ts.coff[c];
		const uint32_t k1 = ts.coff[c + 1];
		double acc = 0.0;
		for (; k + 4 <= k1; k += 4) {
			const double x0 = lin[ts.cmem[k] * RC + r];
			const double x1 = lin[ts.cmem[k + 1] * RC + r];
			const double x2 = lin[ts.cmem[k + 2] * RC + r];
			const double x3 = lin[ts.cmem[k + 3] * RC + r];
			acc += x0;
			acc += x1;
			acc += x2;
			acc += x3;
		}
		for (; k < k1; ++k)
			acc += lin[ts.cmem[k] * RC + r];
		D[c * RC + r] = acc;
	}
	__syncthreads();
	// transcript pass + M-step
	for (uint32_t t = e0; t < tc.nt; t += ES) {
		if (!live)
			continue;
		const double l = lin[t * RC + r];
		uint32_t j = ts.toff[t];
		const uint32_t j1 = ts.toff[t + 1];
		double acc = 0.0;
		for (; j + 8 <= j1; j += 8) {
			double x[8];
#pragma unroll
			for (int q = 0; q < 8; ++q) {
				const uint32_t c = ts.tcls[j + q];
				const double d = D[c * RC + r];
				x[q] = d == 0.0 ? 0.0 : l * (double)n[c * RC + r] / d;
			}
#pragma unroll
			for (int q = 0; q < 8; ++q)
				acc += x[q];
		}
		for (; j < j1; ++j) {
			const uint32_t c = ts.tcls[j];
			const double d = D[c * RC + r];
			acc += d == 0.0 ? 0.0 : l * (double)n[c * RC + r] / d;
		}
		const double nw = ((double)u[t * RC + r] + acc) /
		                  (double)tl.len[tc.tx_base + t];
		lout[t * RC + r] = nw;
		if (gout)
			gout[(size_t)tl.tx[tc.tx_base + t] * R + r0 + r] = nw;
		if (stopf && nw > smin && fabs(nw - l) / nw > srel)
			stopf[r] = 1;
	}
	__syncthreads();
}

template <typename TC, int RC>
__device__ __forceinline__ void carve(void *smem, const hemk_tiles &tl,
                                      double *&lam0, double *&lam1,
                                      double *&lam2, double *&D, TC *&n,
                                      uint32_t *&u, TileStruct &ts)
{
	const uint32_t mt = tl.max_tx, mc = tl.max_cls;
	lam0 = (double *)smem;
	lam1 = lam0 + mt * RC;
	lam2 = lam1 + mt * RC;
	D = lam2 + mt * RC;
	n = (TC *)(D + mc * RC);
	u = (uint32_t *)(n + mc * RC);
	ts.coff = (uint16_t *)(u + mt * RC);
	ts.toff = ts.coff + (mc + 1);
	ts.tcls = ts.toff + (mt + 1);
	ts.cmem = (uint8_t *)(ts.tcls + tl.max_nnz);
}

// Marks the chunk's live replicates; returns false when none is live.
template <int RC>
__device__ __forceinline__ bool chunk_alive(const hemk_rctl &rc, uint32_t r0,
                                            int *alive)
{
	if (threadIdx.x < RC) {
		const uint32_t r = r0 + threadIdx.x;
		alive[threadIdx.x] = (r < rc.R && !rc.done[r]) ? 1 : 0;
	}
	__syncthreads();
	int any = 0;
	for (int i = 0; i < RC; ++i)
		any |= alive[i];
	return any != 0;
}

template <typename TC, int RC>
__device__ __forceinline__ void stage_counts(const hemk_tiles &tl,
                                             const TileCtx &tc,
                                             const hemk_rctl &rc, uint32_t r0,
                                             const int *alive,
                                             const TC *__restrict__ cnt,
                                             const uint32_t *__restrict__ uniq,
                                             TC *n, uint32_t *u)
{
	const int r = threadIdx.x % RC;
	const uint32_t e0 = threadIdx.x / RC;
	constexpr uint32_t ES = TBLOCK / RC;
	if (alive[r]) {
		for (uint32_t c = e0; c < tc.nc; c += ES)
			n[c * RC + r] = cnt[(size_t)(tc.cls_base + c) * rc.R + r0 + r];
		for (uint32_t t = e0; t < tc.nt; t += ES)
			u[t * RC + r] =
			    uniq[(size_t)tl.tx[tc.tx_base + t] * rc.R + r0 + r];
	}
}

template <typename TC, int RC>
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
	const uint32_t r0 = blockIdx.y * RC;
	if (!chunk_alive<RC>(rc, r0, alive))
		return;
	const TileCtx tc = load_tile(tl, tile);
	double *lam0, *lam1, *lam2, *D;
	TC *n;
	uint32_t *u;
	TileStruct ts;
	carve<TC, RC>(smem, tl, lam0, lam1, lam2, D, n, u, ts);
	const int r = threadIdx.x % RC;
	const uint32_t e0 = threadIdx.x / RC;
	constexpr uint32_t ES = TBLOCK / RC;
	if (alive[r])
		for (uint32_t t = e0; t < tc.nt; t += ES)
			lam0[t * RC + r] =
			    A[(size_t)tl.tx[tc.tx_base + t] * rc.R + r0 + r];
	stage_counts<TC, RC>(tl, tc, rc, r0, alive, cnt, uniq, n, u);
	stage_struct(tl, tc, ts);
	__syncthreads();
	tile_map<TC, RC>(tl, tc, ts, lam0, lam1, D, n, u, alive, B, rc.R, r0,
	                 nullptr, 0, 0);
	tile_map<TC, RC>(tl, tc, ts, lam1, lam2, D, n, u, alive, C, rc.R, r0,
	                 nullptr, 0, 0);
	if (xr && alive[r]) {
		// r_t and v_t (src/EM.c:334-335) in replicate-major rows for the
		// sequential norms
		for (uint32_t t = e0; t < tc.nt; t += ES) {
			const double x = lam1[t * RC + r] - lam0[t * RC + r];
			const double y = (lam2[t * RC + r] - lam1[t * RC + r]) - x;
			const size_t g = (size_t)(r0 + r) * T + tl.tx[tc.tx_base + t];
			xr[g] = x;
			xv[g] = y;
		}
	}
	if (part && threadIdx.x < RC && alive[threadIdx.x]) {
		// per (tile, replicate): local two-pass mean and centred sum of
		// squares of r and v (combined across tiles by Chan's rule), or
		// plain sums of squares for the L2 step
		const int q = threadIdx.x;
		double a0, a1, b0, b1;
		if (step == 0) {
			double sr = 0, sv = 0;
			for (uint32_t t = 0; t < tc.nt; ++t) {
				const double x = lam1[t * RC + q] - lam0[t * RC + q];
				const double y = (lam2[t * RC + q] - lam1[t * RC + q]) - x;
				sr += x;
				sv += y;
			}
			const double mr = sr / tc.nt, mv = sv / tc.nt;
			double qr = 0, qv = 0;
			for (uint32_t t = 0; t < tc.nt; ++t) {
				const double x = lam1[t * RC + q] - lam0[t * RC + q];
				const double y = (lam2[t * RC + q] - lam1[t * RC + q]) - x;
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
				const double x = lam1[t * RC + q] - lam0[t * RC + q];
				const double y = (lam2[t * RC + q] - lam1[t * RC + q]) - x;
				qr += x * x;
				qv += y * y;
			}
			a0 = 0;
			a1 = qr;
			b0 = 0;
			b1 = qv;
		}
		double *pp = part + ((size_t)tile * rc.R + r0 + q) * 4;
		pp[0] = a0;
		pp[1] = a1;
		pp[2] = b0;
		pp[3] = b1;
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

template <typename TC, int RC>
__global__ __launch_bounds__(TBLOCK) void k_tile_ex3(
    hemk_tiles tl, hemk_rctl rc, const TC *__restrict__ cnt,
    const uint32_t *__restrict__ uniq, double *__restrict__ A,
    const double *__restrict__ B, const double *__restrict__ C, uint32_t T,
    hemk_step st)
{
	// no contraction: the extrapolation rounds as src/EM.c:350
#pragma clang fp contract(off)
	extern __shared__ double smem[];
	__shared__ int alive[RC];
	__shared__ int stopf[RC];
	__shared__ double alph[RC];
	const uint32_t tile = blockIdx.x;
	const uint32_t r0 = blockIdx.y * RC;
	if (threadIdx.x < RC)
		stopf[threadIdx.x] = 0;
	if (!chunk_alive<RC>(rc, r0, alive))
		return;
	if (threadIdx.x < RC && alive[threadIdx.x]) {
		const uint32_t q = r0 + threadIdx.x;
		// sequential norms leave the two sums; the parallel ones alpha
		alph[threadIdx.x] = st.alpha_given
		                        ? rc.alpha[q]
		                        : step_len(rc.q_r[q], rc.q_v[q], T, st);
		if (!st.alpha_given && tile == 0)
			rc.alpha[q] = alph[threadIdx.x];
	}
	__syncthreads();
	const TileCtx tc = load_tile(tl, tile);
	double *lam0, *lam1, *lam2, *D;
	TC *n;
	uint32_t *u;
	TileStruct ts;
	carve<TC, RC>(smem, tl, lam0, lam1, lam2, D, n, u, ts);
	const int r = threadIdx.x % RC;
	const uint32_t e0 = threadIdx.x / RC;
	constexpr uint32_t ES = TBLOCK / RC;
	if (alive[r]) {
		const double alpha = alph[r];
		for (uint32_t t = e0; t < tc.nt; t += ES) {
			const size_t g = (size_t)tl.tx[tc.tx_base + t] * rc.R + r0 + r;
			const double a = A[g], b = B[g], c2 = C[g];
			const double rr = b - a;
			const double vv = (c2 - b) - rr;
			double x = a - 2 * alpha * rr + alpha * alpha * vv;
			if (st.clamp == 0) {
				if (x < 0)
					x = 0;
			} else if (x <= 1e-300) {
				x = c2 / 10;
			}
			lam0[t * RC + r] = x;
		}
	}
	stage_counts<TC, RC>(tl, tc, rc, r0, alive, cnt, uniq, n, u);
	stage_struct(tl, tc, ts);
	__syncthreads();
	tile_map<TC, RC>(tl, tc, ts, lam0, lam1, D, n, u, alive, A, rc.R, r0,
	                 stopf, rc.stop_min, rc.stop_rel);
	if (threadIdx.x < RC && alive[threadIdx.x] && stopf[threadIdx.x])
		rc.notconv[r0 + threadIdx.x] = 1;
}

// (n, mean, M2) folding of Chan, Golub & LeVeque (1979).
struct Moments {
	double n, m, q;
};

__device__ __forceinline__ Moments mom_merge(Moments a, Moments b)
{
#pragma clang fp contract(off)
	if (a.n == 0)
		return b;
	if (b.n == 0)
		return a;
	Moments o;
	o.n = a.n + b.n;
	const double d = b.m - a.m;
	o.m = a.m + d * (b.n / o.n);
	o.q = a.q + b.q + d * d * (a.n * b.n / o.n);
	return o;
}

// One block per replicate: fold the tile partials in tile order.
__global__ __launch_bounds__(TBLOCK) void k_batch_alpha(
    hemk_tiles tl, hemk_rctl rc, const double *__restrict__ part, uint32_t T,
    hemk_step st)
{
	const uint32_t r = blockIdx.x;
	if (rc.done[r])
		return;
	__shared__ Moments sr[TBLOCK], sv[TBLOCK];
	const uint32_t per = (tl.n_tiles + TBLOCK - 1) / TBLOCK;
	const uint32_t i0 = threadIdx.x * per;
	const uint32_t i1 = min(tl.n_tiles, i0 + per);
	Moments a = {0, 0, 0}, b = {0, 0, 0};
	for (uint32_t i = i0; i < i1; ++i) {
		const double nt = (double)tl.desc[4 * i + 1];
		const double *pp = part + ((size_t)i * rc.R + r) * 4;
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
		rc.alpha[r] = step_len(sr[0].q, sv[0].q, T, st);
}

// Sequential norms of a batch (see hemk_batch_seq): lanes 0..NR-1 of wave 0
// each run two left-to-right chains for one replicate while waves 1-3 stage
// rows [t][r0, r0+NR) of the element values in LDS.
#define BSEQ_NR 16
#define BSEQ_TB 64
__global__ __launch_bounds__(TBLOCK) void k_batch_seq(
    hemk_rctl rc, uint32_t T, const double *__restrict__ A,
    const double *__restrict__ B, const double *__restrict__ C, int phase,
    hemk_step st)
{
#pragma clang fp contract(off)
	__shared__ double buf[2][2][BSEQ_TB][BSEQ_NR];
	__shared__ int alive[BSEQ_NR];
	const uint32_t r0 = blockIdx.x * BSEQ_NR;
	const uint32_t R = rc.R;
	if (threadIdx.x < BSEQ_NR) {
		const uint32_t r = r0 + threadIdx.x;
		alive[threadIdx.x] =
		    (r < R && (phase == 0 || !rc.done[r])) ? 1 : 0;
	}
	__syncthreads();
	int any = 0;
	for (int i = 0; i < BSEQ_NR; ++i)
		any |= alive[i];
	if (!any)
		return;
	const int w = threadIdx.x / 64;
	const uint32_t nch = (T + BSEQ_TB - 1) / BSEQ_TB;
	double acc_a = 0, acc_b = 0, mr = 0, mv = 0;
	if (phase == 2 && threadIdx.x < BSEQ_NR && alive[threadIdx.x]) {
		mr = rc.mean_r[r0 + threadIdx.x];
		mv = rc.mean_v[r0 + threadIdx.x];
	}
	__shared__ double smr[BSEQ_NR], smv[BSEQ_NR];
	if (threadIdx.x < BSEQ_NR) {
		smr[threadIdx.x] = mr;
		smv[threadIdx.x] = mv;
	}
	__syncthreads();
	for (uint32_t ch = 0; ch <= nch; ++ch) {
		if (w >= 1 && ch < nch) {
			const int slot = ch & 1;
			const uint32_t t0 = ch * BSEQ_TB;
			for (int it = threadIdx.x - 64; it < BSEQ_TB * BSEQ_NR;
			     it += TBLOCK - 64) {
				const int q = it % BSEQ_NR;
				const int row = it / BSEQ_NR;
				const uint32_t t = t0 + row;
				double a = 0, b = 0;
				if (t < T && alive[q]) {
					const size_t g = (size_t)t * R + r0 + q;
					if (phase == 0) {
						a = A[g];
					} else {
						const double x = B[g] - A[g];
						const double y = (C[g] - B[g]) - x;
						if (phase == 1) {
							a = x;
							b = y;
						} else if (phase == 2) {
							a = (x - smr[q]) * (x - smr[q]);
							b = (y - smv[q]) * (y - smv[q]);
						} else {
							a = x * x;
							b = y * y;
						}
					}
				}
				buf[slot][0][row][q] = a;
				buf[slot][1][row][q] = b;
			}
		}
		if (w == 0 && threadIdx.x < BSEQ_NR && ch >= 1) {
			const int slot = (ch - 1) & 1;
			const uint32_t cnt = min((uint32_t)BSEQ_TB,
			                         T - (ch - 1) * BSEQ_TB);
			const int q = threadIdx.x;
#pragma unroll 8
			for (uint32_t j = 0; j < cnt; ++j) {
				acc_a += buf[slot][0][j][q];
				acc_b += buf[slot][1][j][q];
			}
		}
		__syncthreads();
	}
	if (threadIdx.x < BSEQ_NR && alive[threadIdx.x]) {
		const uint32_t r = r0 + threadIdx.x;
		if (phase == 0) {
			rc.lam_sum[r] = acc_a;
		} else if (phase == 1) {
			rc.mean_r[r] = acc_a / T;
			rc.mean_v[r] = acc_b / T;
		} else {
			rc.alpha[r] = step_len(acc_a, acc_b, T, st);
		}
	}
}

// acc + b[0] + b[1] + ... + b[cnt-1], left to right.  The next eight values
// are read from LDS while the current eight are added, so the chain runs at
// the latency of a dependent FP64 add rather than of an LDS read.
__device__ __forceinline__ double chain_add(double acc, const double *b,
                                            uint32_t cnt)
{
#pragma clang fp contract(off)
	uint32_t j = 0;
	if (cnt >= 8) {
		double x0 = b[0], x1 = b[1], x2 = b[2], x3 = b[3];
		double x4 = b[4], x5 = b[5], x6 = b[6], x7 = b[7];
		for (j = 8; j + 8 <= cnt; j += 8) {
			const double y0 = b[j], y1 = b[j + 1], y2 = b[j + 2],
			             y3 = b[j + 3], y4 = b[j + 4], y5 = b[j + 5],
			             y6 = b[j + 6], y7 = b[j + 7];
			acc += x0;
			acc += x1;
			acc += x2;
			acc += x3;
			acc += x4;
			acc += x5;
			acc += x6;
			acc += x7;
			x0 = y0;
			x1 = y1;
			x2 = y2;
			x3 = y3;
			x4 = y4;
			x5 = y5;
			x6 = y6;
			x7 = y7;
		}
		acc += x0;
		acc += x1;
		acc += x2;
		acc += x3;
		acc += x4;
		acc += x5;
		acc += x6;
		acc += x7;
	}
	for (; j < cnt; ++j)
		acc += b[j];
	return acc;
}

// Sequential norms of a batch from the replicate-major r/v rows written by
// the map kernel: block (replicate, which) runs one left-to-right chain on
// lane 0 of wave 0 while waves 1-3 stage the next chunk of element values.
//   phase 1: mean = (sum x)/T      phase 2: q = sum (x - mean)^2
//   phase 3: q = sum x^2 (L2 step)
#define CHAIN_TB 2048
__global__ __launch_bounds__(TBLOCK) void k_batch_chain(
    hemk_rctl rc, uint32_t T, const double *__restrict__ xr,
    const double *__restrict__ xv, int phase)
{
#pragma clang fp contract(off)
	__shared__ double buf[2][CHAIN_TB];
	const uint32_t r = blockIdx.x;
	const int which = blockIdx.y;
	if (rc.done[r])
		return;
	const double *x = (which ? xv : xr) + (size_t)r * T;
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
					a = phase == 1 ? v : phase == 2 ? (v - m) * (v - m)
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
		if (phase == 1) {
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

__global__ __launch_bounds__(TBLOCK) void k_batch_iter_end(hemk_rctl rc)
{
	__shared__ uint32_t cnt[TBLOCK];
	uint32_t act = 0;
	for (uint32_t r = threadIdx.x; r < rc.R; r += TBLOCK) {
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
		rc.active[0] = cnt[0];
}

__global__ __launch_bounds__(TBLOCK) void k_batch_reset(hemk_rctl rc,
                                                        uint32_t T,
                                                        double *__restrict__ A)
{
	const size_t n = (size_t)T * rc.R;
	const double x = (double)1 / T;
	for (size_t i = (size_t)blockIdx.x * TBLOCK + threadIdx.x; i < n;
	     i += (size_t)gridDim.x * TBLOCK)
		A[i] = x;
	if (blockIdx.x == 0) {
		uint32_t act = 0;
		for (uint32_t r = threadIdx.x; r < rc.R; r += TBLOCK) {
			rc.k[r] = 0;
			rc.conv[r] = 0;
			rc.notconv[r] = 0;
			rc.alpha[r] = 0;
			rc.lam_sum[r] = 0;
			// the loop condition of src/EM.c:328 at k = 0, stop = 0
			rc.done[r] = (0 < rc.min_iter || 0 < rc.max_iter) ? 0u : 1u;
		}
		(void)act;
		if (threadIdx.x == 0)
			rc.active[0] = (0 < rc.min_iter || 0 < rc.max_iter) ? rc.R : 0;
	}
}

// est[r][t], tpm[r][t] from lam[t][r] through a 32x32 LDS transpose.
__global__ __launch_bounds__(TBLOCK) void k_batch_finalize(
    hemk_rctl rc, uint32_t T, const double *__restrict__ A,
    const uint32_t *__restrict__ len, double *__restrict__ est,
    double *__restrict__ tpm)
{
	__shared__ double tile[32][33];
	const uint32_t t0 = blockIdx.x * 32, q0 = blockIdx.y * 32;
	const uint32_t R = rc.R;
	// load lam[t0+i][q0+j]: j fastest
	for (uint32_t k = threadIdx.x; k < 32 * 32; k += TBLOCK) {
		const uint32_t i = k / 32, j = k % 32;
		const uint32_t t = t0 + i, q = q0 + j;
		tile[i][j] = (t < T && q < R) ? A[(size_t)t * R + q] : 0.0;
	}
	__syncthreads();
	// store out[q0+j][t0+i]: i fastest
	for (uint32_t k = threadIdx.x; k < 32 * 32; k += TBLOCK) {
		const uint32_t j = k / 32, i = k % 32;
		const uint32_t t = t0 + i, q = q0 + j;
		if (t < T && q < R) {
			const double l = tile[i][j];
			const double sum = rc.lam_sum[q];
			double e = 0, g = 0;
			if (sum != 0) {
				e = l * (double)len[t];
				g = l * 1000000 / sum;
			}
			if (est)
				est[(size_t)q * T + t] = e;
			if (tpm)
				tpm[(size_t)q * T + t] = g;
		}
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
	       2 * ((size_t)t->max_cls + 1) + 2 * ((size_t)t->max_tx + 1) +
	       3 * (size_t)t->max_nnz + 16;
}

template <typename TC, int RC>
int launch_map12(const hemk_tiles *t, hemk_rctl rc, const void *cnt,
                 const uint32_t *uniq, const double *A, double *B, double *C,
                 double *part, double *xr, double *xv, uint32_t T, int step,
                 hipStream_t s)
{
	const size_t lds = tile_lds(t, RC, sizeof(TC));
	dim3 grid(t->n_tiles, cdiv32(rc.R, RC));
	k_tile_map12<TC, RC><<<grid, TBLOCK, lds, s>>>(
	    *t, rc, (const TC *)cnt, uniq, A, B, C, part, xr, xv, T, step);
	return (int)hipGetLastError();
}

template <typename TC, int RC>
int launch_ex3(const hemk_tiles *t, hemk_rctl rc, const void *cnt,
               const uint32_t *uniq, double *A, const double *B,
               const double *C, uint32_t T, hemk_step st, hipStream_t s)
{
	const size_t lds = tile_lds(t, RC, sizeof(TC));
	dim3 grid(t->n_tiles, cdiv32(rc.R, RC));
	k_tile_ex3<TC, RC><<<grid, TBLOCK, lds, s>>>(
	    *t, rc, (const TC *)cnt, uniq, A, B, C, T, st);
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
                    const void *cnt, int cnt_bytes, const uint32_t *uniq,
                    const double *A, double *B, double *C, double *part,
                    double *xr, double *xv, uint32_t T, int step, void *s)
{
	hipStream_t st = (hipStream_t)s;
	if (t->n_tiles == 0)
		return 0;
	if (cnt_bytes == 4) {
#define M12U(RC) launch_map12<uint32_t, RC>(t, rc, cnt, uniq, A, B, C, part, xr, xv, T, step, st)
		HEM_RC_SWITCH(chunk, M12U)
#undef M12U
	}
#define M12D(RC) launch_map12<double, RC>(t, rc, cnt, uniq, A, B, C, part, xr, xv, T, step, st)
	HEM_RC_SWITCH(chunk, M12D)
#undef M12D
}

int hemk_tile_ex3(const hemk_tiles *t, hemk_rctl rc, int chunk,
                  const void *cnt, int cnt_bytes, const uint32_t *uniq,
                  double *A, const double *B, const double *C, uint32_t T,
                  hemk_step st, void *s)
{
	hipStream_t sm = (hipStream_t)s;
	if (t->n_tiles == 0)
		return 0;
	if (cnt_bytes == 4) {
#define EXU(RC) launch_ex3<uint32_t, RC>(t, rc, cnt, uniq, A, B, C, T, st, sm)
		HEM_RC_SWITCH(chunk, EXU)
#undef EXU
	}
#define EXD(RC) launch_ex3<double, RC>(t, rc, cnt, uniq, A, B, C, T, st, sm)
	HEM_RC_SWITCH(chunk, EXD)
#undef EXD
}

int hemk_batch_alpha(const hemk_tiles *t, hemk_rctl rc, const double *part,
                     uint32_t T, hemk_step st, void *s)
{
	k_batch_alpha<<<rc.R, TBLOCK, 0, (hipStream_t)s>>>(*t, rc, part, T, st);
	return (int)hipGetLastError();
}

int hemk_batch_seq(hemk_rctl rc, uint32_t T, const double *A,
                   const double *B, const double *C, int phase, hemk_step st,
                   void *s)
{
	k_batch_seq<<<cdiv32(rc.R, BSEQ_NR), TBLOCK, 0, (hipStream_t)s>>>(
	    rc, T, A, B, C, phase, st);
	return (int)hipGetLastError();
}

int hemk_batch_chain(hemk_rctl rc, uint32_t T, const double *xr,
                     const double *xv, int phase, void *s)
{
	dim3 grid(rc.R, 2);
	k_batch_chain<<<grid, TBLOCK, 0, (hipStream_t)s>>>(rc, T, xr, xv, phase);
	return (int)hipGetLastError();
}

int hemk_batch_iter_end(hemk_rctl rc, void *s)
{
	k_batch_iter_end<<<1, TBLOCK, 0, (hipStream_t)s>>>(rc);
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

int hemk_batch_finalize(hemk_rctl rc, uint32_t T, const double *A,
                        const uint32_t *len, double *est, double *tpm,
                        void *s)
{
	dim3 grid(cdiv32(T, 32), cdiv32(rc.R, 32));
	k_batch_finalize<<<grid, TBLOCK, 0, (hipStream_t)s>>>(rc, T, A, len, est,
	                                                      tpm);
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
