// se_kernels.hip -- standard errors by batched conjugate gradients (gfx950).
//
// For K transcripts j at once, solves F x_j = b_j with F = -H, H the Hessian
// of the log-likelihood in theta (grad_kernels.hip), and b_j = d pi_j /
// d theta = pi_j (e_j - pi), so Var(pi_j) ~ b_j . x_j (delta method with the
// observed information).  b_j is orthogonal to the null vector of F (the
// ones vector: softmax is shift invariant), so CG stays in F's range.
//
// Layout: every vector is [q][K], q the tile position of a transcript,
// direction fastest.  One CG step is
//   tile pass   Q = M(pi . P)    per tile, LDS only (class pass, then
//                                 transcript pass), K directions in chunks
//   FP          F P = -Q - N (pi . P) pi + delta P,   partial sums of P.FP
//   XR          x += alpha P, r -= alpha F P,         partial sums of r.m r, r.r
//   P           p = m r + beta p,                     partial sums of pi.p
// with m the Jacobi preconditioner (1 / diag F).  Partial sums are per
// (row block, direction) and folded in row-block order, so every run gives
// the same bits.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "hem_dev.h"

#define SB 256        // threads per block
#define SCOLS 64      // directions per elementwise block
#define SROWS 256     // rows per elementwise block
#define SLDS_BUDGET (40 * 1024 - 512)

namespace {

struct STile {
	uint32_t tx_base, nt, cls_base, nc, mb, nz, jb;
};

__device__ __forceinline__ STile s_tile(const hemk_tiles &t, uint32_t i)
{
	const uint32_t *d = t.desc + (size_t)HEMK_TILE_DESC * i;
	STile c;
	c.tx_base = d[0];
	c.nt = d[1];
	c.cls_base = d[2];
	c.nc = d[3];
	c.mb = d[4];
	c.nz = d[5];
	c.jb = d[6];
	return c;
}

#define SPF 2  // register slots per thread for a chunk's P rows

// The damped Fisher product of one tile, K directions in chunks of RC:
//   Dd_c = sum_{s in c} x_s P_s,  Q_t = P_t a_t - x_t sum_{c ni t} kappa_c Dd_c
// is the HVP without its rank-one term; F P = -(H P) + delta P then is
//   Q_t <- -Q_t - N (pi . P) pi_t + delta P_t,
// written in place of Q, and the tile's p . F p per direction goes to
// part[tile * K + k] (its transcripts in order of the four row groups).
template <int RC>
__global__ __launch_bounds__(SB) void k_se_tile(
    hemk_tiles tl, uint32_t K, const double *__restrict__ xq,
    const double *__restrict__ aq, const double *__restrict__ kq,
    const double *__restrict__ piq, const double *__restrict__ pv,
    const double *__restrict__ scal, double delta,
    const double *__restrict__ P, double *__restrict__ Q,
    double *__restrict__ part)
{
	extern __shared__ double smem[];
	__shared__ double red[SB];
	const STile tc = s_tile(tl, blockIdx.x);
	const uint32_t mt = tl.max_tx, mc = tl.max_cls, mz = tl.max_nnz;
	double *Pl = smem;             // mt * RC
	double *Dd = Pl + mt * RC;     // mc * RC
	double *xs = Dd + mc * RC;     // mt
	double *as = xs + mt;          // mt
	double *ps = as + mt;          // mt
	double *ks = ps + mt;          // mc
	uint32_t *kmem = (uint32_t *)(ks + mc);
	uint16_t *coff = (uint16_t *)(kmem + mz);
	uint16_t *toff = coff + (mc + 1);
	uint16_t *tcls = toff + (mt + 1);
	for (uint32_t i = threadIdx.x; i <= tc.nc; i += SB)
		coff[i] = (uint16_t)(tl.coff[tc.cls_base + i] - tc.mb);
	for (uint32_t i = threadIdx.x; i <= tc.nt; i += SB)
		toff[i] = (uint16_t)(tl.toff[tc.tx_base + i] - tc.jb);
	for (uint32_t i = threadIdx.x; i < tc.nz; i += SB) {
		kmem[i] = tl.kmem[tc.mb + i];
		tcls[i] = tl.tcls[tc.jb + i];
	}
	for (uint32_t i = threadIdx.x; i < tc.nt; i += SB) {
		xs[i] = xq[tc.tx_base + i];
		as[i] = aq[tc.tx_base + i];
		ps[i] = piq[tc.tx_base + i];
	}
	const double N = scal[1];
	for (uint32_t i = threadIdx.x; i < tc.nc; i += SB)
		ks[i] = kq[tc.cls_base + i];
	const uint32_t r = threadIdx.x % RC;
	const uint32_t e0 = threadIdx.x / RC;
	constexpr uint32_t ES = SB / RC;
	const uint32_t nchunk = (K + RC - 1) / RC;
	double pf[SPF];
#pragma unroll
	for (int s = 0; s < SPF; ++s) {
		const uint32_t t = (threadIdx.x + SB * s) / RC;
		pf[s] = (t < tc.nt && r < K) ? P[(size_t)(tc.tx_base + t) * K + r]
		                             : 0.0;
	}
	for (uint32_t ch = 0; ch < nchunk; ++ch) {
		const uint32_t k = ch * RC + r;
		const bool kin = k < K;
#pragma unroll
		for (int s = 0; s < SPF; ++s) {
			const uint32_t e = threadIdx.x + SB * s;
			if (e / RC < tc.nt)
				Pl[e] = pf[s];
		}
		for (uint32_t e = threadIdx.x + SB * SPF; e / RC < tc.nt; e += SB)
			Pl[e] = kin ? P[(size_t)(tc.tx_base + e / RC) * K + k] : 0.0;
		__syncthreads();
		if (ch + 1 < nchunk) {
			const uint32_t kn = k + RC;
#pragma unroll
			for (int s = 0; s < SPF; ++s) {
				const uint32_t t = (threadIdx.x + SB * s) / RC;
				pf[s] = (t < tc.nt && kn < K)
				            ? P[(size_t)(tc.tx_base + t) * K + kn]
				            : 0.0;
			}
		}
		// (four members / eight occurrences at a time, their LDS loads
		// issued before the first product or sum)
		for (uint32_t c = e0; c < tc.nc; c += ES) {
			double acc = 0;
			uint32_t j = coff[c];
			const uint32_t j1 = coff[c + 1];
			for (; j + 4 <= j1; j += 4) {
				uint32_t t[4];
				double xv[4], pv4[4];
#pragma unroll
				for (int q = 0; q < 4; ++q)
					t[q] = kmem[j + q] & 0xFFu;
#pragma unroll
				for (int q = 0; q < 4; ++q) {
					xv[q] = xs[t[q]];
					pv4[q] = Pl[t[q] * RC + r];
				}
#pragma unroll
				for (int q = 0; q < 4; ++q)
					acc += xv[q] * pv4[q];
			}
			for (; j < j1; ++j) {
				const uint32_t t = kmem[j] & 0xFFu;
				acc += xs[t] * Pl[t * RC + r];
			}
			Dd[c * RC + r] = ks[c] * acc;
		}
		__syncthreads();
		const double cv = kin ? N * pv[k] : 0.0;
		double pf = 0;
		for (uint32_t t = e0; t < tc.nt; t += ES) {
			double acc = 0;
			uint32_t j = toff[t];
			const uint32_t j1 = toff[t + 1];
			for (; j + 8 <= j1; j += 8) {
				uint32_t c[8];
				double d[8];
#pragma unroll
				for (int q = 0; q < 8; ++q)
					c[q] = tcls[j + q];
#pragma unroll
				for (int q = 0; q < 8; ++q)
					d[q] = Dd[c[q] * RC + r];
#pragma unroll
				for (int q = 0; q < 8; ++q)
					acc += d[q];
			}
			for (; j < j1; ++j)
				acc += Dd[tcls[j] * RC + r];
			if (kin) {
				const double p = Pl[t * RC + r];
				const double hq = p * as[t] - xs[t] * acc;
				const double f = -hq - cv * ps[t] + delta * p;
				Q[(size_t)(tc.tx_base + t) * K + k] = f;
				pf += p * f;
			}
		}
		red[threadIdx.x] = pf;
		__syncthreads();
		if (e0 == 0 && kin) {
			double sum = red[r];
			for (uint32_t i = 1; i < ES; ++i)
				sum += red[i * RC + r];
			part[(size_t)blockIdx.x * K + k] = sum;
		}
		__syncthreads();
	}
}

// The elementwise kernels: block (row block rb, column block cb) covers
// rows [rb*SROWS, +SROWS) and directions [cb*SCOLS, +SCOLS); thread
// (row lane threadIdx.x / SCOLS, column threadIdx.x % SCOLS).  Column
// partial sums over the block's rows land in part[rb*K + k].
struct SIdx {
	uint32_t k, q0, q1;
	bool kin;
};

__device__ __forceinline__ SIdx s_idx(uint32_t T, uint32_t K)
{
	SIdx i;
	i.k = blockIdx.y * SCOLS + threadIdx.x % SCOLS;
	i.kin = i.k < K;
	i.q0 = blockIdx.x * SROWS + threadIdx.x / SCOLS;
	i.q1 = min(T, (blockIdx.x + 1) * SROWS);
	return i;
}

// Sums v over the block's row lanes for each column; writes the partials.
template <int NV>
__device__ __forceinline__ void s_partials(double (&v)[NV], uint32_t K,
                                           double *const (&part)[NV])
{
	__shared__ double sh[SB / SCOLS][SCOLS];
	const uint32_t col = threadIdx.x % SCOLS, row = threadIdx.x / SCOLS;
	const uint32_t k = blockIdx.y * SCOLS + col;
#pragma unroll
	for (int i = 0; i < NV; ++i) {
		sh[row][col] = v[i];
		__syncthreads();
		if (row == 0 && k < K) {
			double s = sh[0][col];
#pragma unroll
			for (int j = 1; j < SB / SCOLS; ++j)
				s += sh[j][col];
			part[i][(size_t)blockIdx.x * K + k] = s;
		}
		__syncthreads();
	}
}

// r = b = pj (e_j - pi), x = 0, p = m r; partials of r.m r, r.r, pi.p
__global__ __launch_bounds__(SB) void k_cg_init(
    const double *__restrict__ piq, const double *__restrict__ mq,
    const uint32_t *__restrict__ jq, const double *__restrict__ pj,
    double delta, uint32_t T, uint32_t K, double *__restrict__ X,
    double *__restrict__ R, double *__restrict__ P, double *__restrict__ prz,
    double *__restrict__ prr, double *__restrict__ ppv)
{
	const SIdx ix = s_idx(T, K);
	double v[3] = {0, 0, 0};
	if (ix.kin) {
		const uint32_t j = jq[ix.k];
		const double b = pj[ix.k];
		for (uint32_t q = ix.q0; q < ix.q1; q += SB / SCOLS) {
			const size_t g = (size_t)q * K + ix.k;
			const double rr = b * ((q == j ? 1.0 : 0.0) - piq[q]);
			const double z = mq[q] * rr;
			X[g] = 0;
			R[g] = rr;
			P[g] = z;
			v[0] += rr * z;
			v[1] += rr * rr;
			v[2] += piq[q] * z;
		}
	}
	(void)delta;
	double *const part[3] = {prz, prr, ppv};
	s_partials<3>(v, K, part);
}

// F p = -Q - N pv pi + delta p (into Q); partials of p.F p
// A direction whose residual is down to CG_TOL |b| is finished: its x and r
// stay as they are (further steps would only stir rounding noise).
#define CG_TOL2 1e-28

// alpha = rz / pFp; x += alpha p; r -= alpha F p; partials of r.m r, r.r
__global__ __launch_bounds__(SB) void k_cg_xr(
    const double *__restrict__ mq, const double *__restrict__ rz,
    const double *__restrict__ pfp, const double *__restrict__ rr,
    const double *__restrict__ bb, uint32_t T, uint32_t K,
    const double *__restrict__ P, const double *__restrict__ FP,
    double *__restrict__ X, double *__restrict__ R, double *__restrict__ prz,
    double *__restrict__ prr)
{
	const SIdx ix = s_idx(T, K);
	double v[2] = {0, 0};
	if (ix.kin) {
		const double d = pfp[ix.k];
		const bool done = rr[ix.k] <= CG_TOL2 * bb[ix.k];
		const double al = (d > 0 && !done) ? rz[ix.k] / d : 0.0;
		for (uint32_t q = ix.q0; q < ix.q1; q += SB / SCOLS) {
			const size_t g = (size_t)q * K + ix.k;
			X[g] += al * P[g];
			const double rr = R[g] - al * FP[g];
			R[g] = rr;
			v[0] += rr * mq[q] * rr;
			v[1] += rr * rr;
		}
	}
	double *const part[2] = {prz, prr};
	s_partials<2>(v, K, part);
}

// beta = rz_new / rz_old; p = m r + beta p; partials of pi.p
__global__ __launch_bounds__(SB) void k_cg_p(
    const double *__restrict__ piq, const double *__restrict__ mq,
    const double *__restrict__ rz_old, const double *__restrict__ rz_new,
    uint32_t T, uint32_t K, const double *__restrict__ R,
    double *__restrict__ P, double *__restrict__ ppv)
{
	const SIdx ix = s_idx(T, K);
	double v[1] = {0};
	if (ix.kin) {
		const double o = rz_old[ix.k];
		const double be = o > 0 ? rz_new[ix.k] / o : 0.0;
		for (uint32_t q = ix.q0; q < ix.q1; q += SB / SCOLS) {
			const size_t g = (size_t)q * K + ix.k;
			const double p = mq[q] * R[g] + be * P[g];
			P[g] = p;
			v[0] += piq[q] * p;
		}
	}
	double *const part[1] = {ppv};
	s_partials<1>(v, K, part);
}

// partials of pi.x
__global__ __launch_bounds__(SB) void k_cg_pix(const double *__restrict__ piq,
                                               uint32_t T, uint32_t K,
                                               const double *__restrict__ X,
                                               double *__restrict__ ppx)
{
	const SIdx ix = s_idx(T, K);
	double v[1] = {0};
	if (ix.kin)
		for (uint32_t q = ix.q0; q < ix.q1; q += SB / SCOLS)
			v[0] += piq[q] * X[(size_t)q * K + ix.k];
	double *const part[1] = {ppx};
	s_partials<1>(v, K, part);
}

// out[k] = sum over row blocks of part[rb*K + k]: FOLD_K directions per
// block; thread (g, kk) adds the row blocks g, g + FOLD_G, ... of direction
// kk in order, then the FOLD_G group sums are added in group order (a
// fixed order, with every block's loads in flight at once rather than one
// thread per direction walking all row blocks).
#define FOLD_K 8
#define FOLD_G (SB / FOLD_K)
__global__ __launch_bounds__(SB) void k_cg_fold(const double *__restrict__ part,
                                                uint32_t nrb, uint32_t K,
                                                double *__restrict__ out)
{
	__shared__ double sh[FOLD_G][FOLD_K];
	const uint32_t kk = threadIdx.x % FOLD_K, g = threadIdx.x / FOLD_K;
	const uint32_t k = blockIdx.x * FOLD_K + kk;
	double s = 0;
	if (k < K)
		for (uint32_t b = g; b < nrb; b += FOLD_G)
			s += part[(size_t)b * K + k];
	sh[g][kk] = s;
	__syncthreads();
	if (g == 0 && k < K) {
		double t = sh[0][kk];
		for (uint32_t i = 1; i < FOLD_G; ++i)
			t += sh[i][kk];
		out[k] = t;
	}
}

// se_k = sqrt(max(0, pj (x[jq] - pi.x))), resid_k = sqrt(rr / bb)
__global__ __launch_bounds__(SB) void k_cg_out(
    const double *__restrict__ X, const uint32_t *__restrict__ jq,
    const double *__restrict__ pj, const double *__restrict__ pix,
    const double *__restrict__ rr, const double *__restrict__ bb, uint32_t K,
    double *__restrict__ se, double *__restrict__ resid)
{
	const uint32_t k = blockIdx.x * SB + threadIdx.x;
	if (k >= K)
		return;
	const double v = pj[k] * (X[(size_t)jq[k] * K + k] - pix[k]);
	se[k] = sqrt(v > 0 ? v : 0.0);
	resid[k] = bb[k] > 0 ? sqrt(rr[k] / bb[k]) : 0.0;
}

// out[i] = in[idx[i]]
__global__ __launch_bounds__(SB) void k_gather_f64(const double *__restrict__ in,
                                                   const uint32_t *__restrict__ idx,
                                                   uint32_t n,
                                                   double *__restrict__ out)
{
	for (uint32_t i = blockIdx.x * SB + threadIdx.x; i < n;
	     i += gridDim.x * SB)
		out[i] = in[idx[i]];
}

__global__ __launch_bounds__(SB) void k_scatter_f64(const double *__restrict__ in,
                                                    const uint32_t *__restrict__ idx,
                                                    uint32_t n,
                                                    double *__restrict__ out)
{
	for (uint32_t i = blockIdx.x * SB + threadIdx.x; i < n;
	     i += gridDim.x * SB)
		out[idx[i]] = in[i];
}

// m_t = 1 / (fd_t + delta), floored so a non-positive diagonal stays usable
__global__ __launch_bounds__(SB) void k_precond(const double *__restrict__ fdq,
                                                const double *__restrict__ piq,
                                                const double *__restrict__ scal,
                                                double delta, uint32_t n,
                                                double *__restrict__ mq)
{
	const double N = scal[1];
	for (uint32_t i = blockIdx.x * SB + threadIdx.x; i < n;
	     i += gridDim.x * SB) {
		const double floor_ = 1e-6 * N * piq[i] + 1e-300;
		const double d = fdq[i] + delta;
		mq[i] = 1.0 / (d > floor_ ? d : floor_);
	}
}

inline uint32_t sdiv(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

size_t se_lds(const hemk_tiles *t, int rc)
{
	return (size_t)rc * ((size_t)t->max_tx * 8 + (size_t)t->max_cls * 8) +
	       (size_t)t->max_tx * 24 + (size_t)t->max_cls * 8 +
	       (size_t)t->max_nnz * 6 + 2 * ((size_t)t->max_cls + 1) +
	       2 * ((size_t)t->max_tx + 1) + 16;
}

template <int RC>
int launch_se_tile(const hemk_tiles *t, uint32_t K, const double *xq,
                   const double *aq, const double *kq, const double *piq,
                   const double *pv, const double *scal, double delta,
                   const double *P, double *Q, double *part, hipStream_t s)
{
	k_se_tile<RC><<<t->n_tiles, SB, se_lds(t, RC), s>>>(
	    *t, K, xq, aq, kq, piq, pv, scal, delta, P, Q, part);
	return (int)hipGetLastError();
}

} // namespace

extern "C" {

int hemk_se_chunk(const hemk_tiles *t, uint32_t K)
{
	int rc = 32;
	while (rc > 1 && ((uint32_t)rc / 2 >= K || se_lds(t, rc) > SLDS_BUDGET))
		rc /= 2;
	return se_lds(t, rc) > SLDS_BUDGET ? 0 : rc;
}

uint32_t hemk_se_rowblocks(uint32_t T) { return sdiv(T, SROWS); }

int hemk_se_tile(const hemk_tiles *t, int chunk, uint32_t K, const double *xq,
                 const double *aq, const double *kq, const double *piq,
                 const double *pv, const double *scal, double delta,
                 const double *P, double *Q, double *part, void *s)
{
	hipStream_t st = (hipStream_t)s;
#define SE_T(RC) launch_se_tile<RC>(t, K, xq, aq, kq, piq, pv, scal, delta, P, Q, part, st)
	switch (chunk) {
	case 1: return SE_T(1);
	case 2: return SE_T(2);
	case 4: return SE_T(4);
	case 8: return SE_T(8);
	case 16: return SE_T(16);
	case 32: return SE_T(32);
	default: return (int)hipErrorInvalidValue;
	}
#undef SE_T
}

int hemk_cg_init(const double *piq, const double *mq, const uint32_t *jq,
                 const double *pj, uint32_t T, uint32_t K, double *X,
                 double *R, double *P, double *prz, double *prr, double *ppv,
                 void *s)
{
	dim3 g(sdiv(T, SROWS), sdiv(K, SCOLS));
	k_cg_init<<<g, SB, 0, (hipStream_t)s>>>(piq, mq, jq, pj, 0.0, T, K, X, R,
	                                        P, prz, prr, ppv);
	return (int)hipGetLastError();
}

int hemk_cg_xr(const double *mq, const double *rz, const double *pfp,
               const double *rr, const double *bb, uint32_t T, uint32_t K,
               const double *P, const double *FP, double *X, double *R,
               double *prz, double *prr, void *s)
{
	dim3 g(sdiv(T, SROWS), sdiv(K, SCOLS));
	k_cg_xr<<<g, SB, 0, (hipStream_t)s>>>(mq, rz, pfp, rr, bb, T, K, P, FP, X,
	                                      R, prz, prr);
	return (int)hipGetLastError();
}

int hemk_cg_p(const double *piq, const double *mq, const double *rz_old,
              const double *rz_new, uint32_t T, uint32_t K, const double *R,
              double *P, double *ppv, void *s)
{
	dim3 g(sdiv(T, SROWS), sdiv(K, SCOLS));
	k_cg_p<<<g, SB, 0, (hipStream_t)s>>>(piq, mq, rz_old, rz_new, T, K, R, P,
	                                     ppv);
	return (int)hipGetLastError();
}

int hemk_cg_pix(const double *piq, uint32_t T, uint32_t K, const double *X,
                double *ppx, void *s)
{
	dim3 g(sdiv(T, SROWS), sdiv(K, SCOLS));
	k_cg_pix<<<g, SB, 0, (hipStream_t)s>>>(piq, T, K, X, ppx);
	return (int)hipGetLastError();
}

int hemk_cg_fold(const double *part, uint32_t nrb, uint32_t K, double *out,
                 void *s)
{
	k_cg_fold<<<sdiv(K, FOLD_K), SB, 0, (hipStream_t)s>>>(part, nrb, K, out);
	return (int)hipGetLastError();
}

int hemk_cg_out(const double *X, const uint32_t *jq, const double *pj,
                const double *pix, const double *rr, const double *bb,
                uint32_t K, double *se, double *resid, void *s)
{
	k_cg_out<<<sdiv(K, SB), SB, 0, (hipStream_t)s>>>(X, jq, pj, pix, rr, bb,
	                                                 K, se, resid);
	return (int)hipGetLastError();
}

int hemk_gather_f64(const double *in, const uint32_t *idx, uint32_t n,
                    double *out, void *s)
{
	if (!n)
		return 0;
	uint32_t g = sdiv(n, SB);
	if (g > 4096)
		g = 4096;
	k_gather_f64<<<g, SB, 0, (hipStream_t)s>>>(in, idx, n, out);
	return (int)hipGetLastError();
}

int hemk_scatter_f64(const double *in, const uint32_t *idx, uint32_t n,
                     double *out, void *s)
{
	if (!n)
		return 0;
	uint32_t g = sdiv(n, SB);
	if (g > 4096)
		g = 4096;
	k_scatter_f64<<<g, SB, 0, (hipStream_t)s>>>(in, idx, n, out);
	return (int)hipGetLastError();
}

int hemk_precond(const double *fdq, const double *piq, const double *scal,
                 double delta, uint32_t n, double *mq, void *s)
{
	if (!n)
		return 0;
	uint32_t g = sdiv(n, SB);
	if (g > 4096)
		g = 4096;
	k_precond<<<g, SB, 0, (hipStream_t)s>>>(fdq, piq, scal, delta, n, mq);
	return (int)hipGetLastError();
}

} // extern "C"
