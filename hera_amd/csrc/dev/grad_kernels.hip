// grad_kernels.hip -- adjoint of the abundance log-likelihood (gfx950).
//
// Objective, with pi = softmax(theta), w_t = exp(-ell_t) (ell = log effective
// length), x_t = pi_t w_t and unique reads as singleton classes:
//   L = sum_c n_c log D_c + sum_t u_t log x_t,   D_c = sum_{t in c} x_t.
// hera's map (src/EM.c:271-302) is the EM map of this objective: at its
// fixed point pi_t = E[cnt_t] / N.  The adjoint is the E-step itself:
//   E_t = x_t sum_{c ni t} n_c / D_c + u_t            (expected counts)
//   dL/dtheta_t = E_t - N pi_t,   dL/dell_t = -E_t,   N = sum_t E_t,
// two passes over the same CSR (classes) and CSC (transcripts) arrays as
// the EM kernels.  A Hessian-vector product in theta is
//   H v = Q + N (pi . v) pi,
//   Q_t = v_t (g_t - u_t) - x_t sum_{c ni t} kappa_c Dd_c,
//   Dd_c = sum_{s in c} x_s v_s,   kappa_c = n_c / D_c^2,
// again one CSR pass and one CSC pass.  FP32 forward: x and D_c in FP32,
// every sum of the adjoint (E_t, N, L) in FP64.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "hem_dev.h"

#define GB 256
#define GWAVE 64
#define NB HEMK_NORM_BLOCKS

namespace {

template <int G, typename T>
__device__ __forceinline__ T gsum(T v)
{
#pragma unroll
	for (int m = G / 2; m >= 1; m >>= 1)
		v += __shfl_xor(v, m, G);
	return v;
}

// Block sum in a fixed order; every thread gets the result.
__device__ __forceinline__ double bsum(double a)
{
	__shared__ double sa[GB / GWAVE];
	a = gsum<GWAVE>(a);
	const int w = threadIdx.x / GWAVE;
	if ((threadIdx.x & (GWAVE - 1)) == 0)
		sa[w] = a;
	__syncthreads();
	a = sa[0];
#pragma unroll
	for (int i = 1; i < GB / GWAVE; ++i)
		a += sa[i];
	__syncthreads();
	return a;
}

__device__ __forceinline__ double bmax(double a)
{
	__shared__ double sa[GB / GWAVE];
#pragma unroll
	for (int m = GWAVE / 2; m >= 1; m >>= 1)
		a = fmax(a, __shfl_xor(a, m, GWAVE));
	const int w = threadIdx.x / GWAVE;
	if ((threadIdx.x & (GWAVE - 1)) == 0)
		sa[w] = a;
	__syncthreads();
	a = sa[0];
#pragma unroll
	for (int i = 1; i < GB / GWAVE; ++i)
		a = fmax(a, sa[i]);
	__syncthreads();
	return a;
}

// Sum of np block (or tile) partials, the same in every block.
__device__ __forceinline__ double foldn(const double *__restrict__ part,
                                        uint32_t np)
{
	double a = 0;
	for (uint32_t i = threadIdx.x; i < np; i += GB)
		a += part[i];
	return bsum(a);
}

__device__ __forceinline__ double fold(const double *__restrict__ part)
{
	return foldn(part, NB);
}

__global__ __launch_bounds__(GB) void k_gmax(const double *__restrict__ th,
                                             uint32_t T,
                                             double *__restrict__ part)
{
	double m = -INFINITY;
	for (uint32_t t = blockIdx.x * GB + threadIdx.x; t < T; t += NB * GB)
		m = fmax(m, th[t]);
	m = bmax(m);
	if (threadIdx.x == 0)
		part[blockIdx.x] = m;
}

// e_t = exp(theta_t - max), w_t = exp(-ell_t); block sums of e.
__global__ __launch_bounds__(GB) void k_gexp(
    const double *__restrict__ th, const double *__restrict__ ell,
    uint32_t T, const double *__restrict__ pmax, double *__restrict__ e,
    double *__restrict__ w, double *__restrict__ part)
{
	double m = -INFINITY;
	for (int i = threadIdx.x; i < NB; i += GB)
		m = fmax(m, pmax[i]);
	m = bmax(m);
	double s = 0;
	for (uint32_t t = blockIdx.x * GB + threadIdx.x; t < T; t += NB * GB) {
		const double v = exp(th[t] - m);
		e[t] = v;
		w[t] = exp(-ell[t]);
		s += v;
	}
	s = bsum(s);
	if (threadIdx.x == 0)
		part[blockIdx.x] = s;
}

// pi = e / sum e (in place), x = pi w (FP64 and FP32 copies).
__global__ __launch_bounds__(GB) void k_gnorm(
    double *__restrict__ pi, const double *__restrict__ w, uint32_t T,
    const double *__restrict__ psum, double *__restrict__ x,
    float *__restrict__ x32)
{
	const double S = fold(psum);
	for (uint32_t t = blockIdx.x * GB + threadIdx.x; t < T; t += NB * GB) {
		const double p = pi[t] / S;
		pi[t] = p;
		const double v = p * w[t];
		x[t] = v;
		x32[t] = (float)v;
	}
}

// Class pass: D_c = sum x over the members (G lanes per class), rho_c =
// n_c / D_c, kappa_c = n_c / D_c^2 (zero when n_c = 0 or D_c = 0), and the
// block partials of sum_c n_c log D_c.
template <int G, typename TX>
__global__ __launch_bounds__(GB) void k_gclass(
    const uint32_t *__restrict__ off, const uint32_t *__restrict__ ref,
    const double *__restrict__ cnt, const TX *__restrict__ x, uint32_t C,
    double *__restrict__ rho, double *__restrict__ Dout,
    double *__restrict__ kappa, double *__restrict__ part)
{
	const uint32_t lane = threadIdx.x % G;
	double L = 0;
	for (uint32_t c0 = blockIdx.x * (GB / G); c0 < C; c0 += NB * (GB / G)) {
		const uint32_t c = c0 + threadIdx.x / G;
		TX acc = 0;
		if (c < C)
			for (uint32_t k = off[c] + lane; k < off[c + 1]; k += G)
				acc += x[ref[k]];
		acc = gsum<G>(acc);
		if (c < C && lane == 0) {
			const double D = (double)acc, n = cnt[c];
			double r = 0, q = 0;
			if (n != 0 && D != 0) {
				r = n / D;
				q = r / D;
				L += n * log(D);
			}
			rho[c] = r;
			Dout[c] = D;
			kappa[c] = q;
		}
	}
	L = bsum(L);
	if (threadIdx.x == 0)
		part[blockIdx.x] = L;
}

// Transcript pass: E_t = x_t sum_{c ni t} rho_c + u_t; the Fisher diagonal
// fd_t = -sum_{c ni t} n_c q (1 - q), q = x_t / D_c (the class part of
// -H_tt; the N pi (1 - pi) part is added with N); block partials of N and
// of sum_t u_t log x_t.
template <int G, typename TX>
__global__ __launch_bounds__(GB) void k_gtx(
    const uint32_t *__restrict__ off, const uint32_t *__restrict__ ref,
    const double *__restrict__ rho, const double *__restrict__ D,
    const double *__restrict__ cnt, const TX *__restrict__ x,
    const uint32_t *__restrict__ unique, uint32_t T, double *__restrict__ E,
    double *__restrict__ fd, double *__restrict__ partN,
    double *__restrict__ partL)
{
	const uint32_t lane = threadIdx.x % G;
	double Nn = 0, L = 0;
	for (uint32_t t0 = blockIdx.x * (GB / G); t0 < T; t0 += NB * (GB / G)) {
		const uint32_t t = t0 + threadIdx.x / G;
		double s = 0, h = 0, xt = 0;
		if (t < T) {
			xt = (double)x[t];
			for (uint32_t k = off[t] + lane; k < off[t + 1]; k += G) {
				const uint32_t c = ref[k];
				const double r = rho[c];
				s += r;
				if (r != 0) {
					const double q = xt / D[c];
					h += cnt[c] * q * (1 - q);
				}
			}
		}
		s = gsum<G>(s);
		h = gsum<G>(h);
		if (t < T && lane == 0) {
			const uint32_t u = unique[t];
			const double e = xt * s + (double)u;
			E[t] = e;
			fd[t] = -h;
			Nn += e;
			if (u)
				L += (double)u * log(xt);
		}
	}
	Nn = bsum(Nn);
	L = bsum(L);
	if (threadIdx.x == 0) {
		partN[blockIdx.x] = Nn;
		partL[blockIdx.x] = L;
	}
}

// g_theta = E - N pi, g_ell = -E, a = g_theta - u (for the HVP), the Fisher
// diagonal completed, and the scalars {L, N} in out[0..1].
__global__ __launch_bounds__(GB) void k_gfin(
    const double *__restrict__ E, const double *__restrict__ pi,
    const uint32_t *__restrict__ unique, uint32_t T,
    const double *__restrict__ partN, const double *__restrict__ partL1,
    const double *__restrict__ partL2, uint32_t np, double *__restrict__ gth,
    double *__restrict__ gell, double *__restrict__ a,
    double *__restrict__ fd, double *__restrict__ out)
{
	const double N = foldn(partN, np);
	const double L1 = foldn(partL1, np);
	const double L2 = foldn(partL2, np);
	if (blockIdx.x == 0 && threadIdx.x == 0) {
		out[0] = L1 + L2;
		out[1] = N;
	}
	for (uint32_t t = blockIdx.x * GB + threadIdx.x; t < T; t += NB * GB) {
		const double e = E[t], p = pi[t];
		const double g = e - N * p;
		gth[t] = g;
		gell[t] = -e;
		a[t] = g - (double)unique[t];
		fd[t] += N * p * (1 - p);
	}
}

// out[b] = the sum of part[b np .. (b+1) np), one block each.
__global__ __launch_bounds__(GB) void k_gfold(const double *__restrict__ part,
                                              uint32_t np,
                                              double *__restrict__ out)
{
	const double v = foldn(part + (size_t)blockIdx.x * np, np);
	if (threadIdx.x == 0)
		out[blockIdx.x] = v;
}

// ---- several samples side by side: a softmax and N per sample ----------
// info[4q..4q+3] = {first transcript, transcripts, first tile, tiles}.

// pi = softmax(theta) within sample q (block q), w = exp(-ell), x = pi w.
__global__ __launch_bounds__(GB) void k_gsoftmax_seg(
    const double *__restrict__ th, const double *__restrict__ ell,
    const uint32_t *__restrict__ info, double *__restrict__ pi,
    double *__restrict__ w, double *__restrict__ x, float *__restrict__ x32)
{
	const uint32_t q = blockIdx.x;
	const uint32_t a = info[4 * q], b = a + info[4 * q + 1];
	double m = -INFINITY;
	for (uint32_t t = a + threadIdx.x; t < b; t += GB)
		m = fmax(m, th[t]);
	m = bmax(m);
	double sum = 0;
	for (uint32_t t = a + threadIdx.x; t < b; t += GB) {
		const double v = exp(th[t] - m);
		pi[t] = v;
		w[t] = exp(-ell[t]);
		sum += v;
	}
	const double S = bsum(sum);
	for (uint32_t t = a + threadIdx.x; t < b; t += GB) {
		const double pq = pi[t] / S;
		pi[t] = pq;
		const double v = pq * w[t];
		x[t] = v;
		x32[t] = (float)v;
	}
}

// tot[3q + k] = the sum of sample q's tile partials of kind k (sum n_c log
// D_c, N, sum u_t log x_t), block q.
__global__ __launch_bounds__(GB) void k_gfold_seg(
    const double *__restrict__ part, uint32_t np,
    const uint32_t *__restrict__ info, double *__restrict__ tot)
{
	const uint32_t q = blockIdx.x;
	const uint32_t a = info[4 * q + 2], n = info[4 * q + 3];
	for (int k = 0; k < 3; ++k) {
		const double v = foldn(part + (size_t)k * np + a, n);
		if (threadIdx.x == 0)
			tot[3 * q + k] = v;
	}
}

// k_gfin with the N of each transcript's sample.
__global__ __launch_bounds__(GB) void k_gfin_seg(
    const double *__restrict__ E, const double *__restrict__ pi,
    const uint32_t *__restrict__ unique, uint32_t T,
    const uint32_t *__restrict__ txsamp, const double *__restrict__ tot,
    double *__restrict__ gth, double *__restrict__ gell,
    double *__restrict__ a, double *__restrict__ fd)
{
	for (uint32_t t = blockIdx.x * GB + threadIdx.x; t < T; t += NB * GB) {
		const double N = tot[3 * (txsamp ? txsamp[t] : 0u) + 1];
		const double e = E[t], p = pi[t];
		const double g = e - N * p;
		gth[t] = g;
		gell[t] = -e;
		a[t] = g - (double)unique[t];
		fd[t] += N * p * (1 - p);
	}
}

// Block partials of sum_t a_t b_t.
__global__ __launch_bounds__(GB) void k_gdot(const double *__restrict__ a,
                                             const double *__restrict__ b,
                                             uint32_t T,
                                             double *__restrict__ part)
{
	double s = 0;
	for (uint32_t t = blockIdx.x * GB + threadIdx.x; t < T; t += NB * GB)
		s += a[t] * b[t];
	s = bsum(s);
	if (threadIdx.x == 0)
		part[blockIdx.x] = s;
}

// HVP class pass: kd_c = kappa_c sum_{s in c} x_s v_s.
template <int G>
__global__ __launch_bounds__(GB) void k_hclass(
    const uint32_t *__restrict__ off, const uint32_t *__restrict__ ref,
    const double *__restrict__ x, const double *__restrict__ v,
    const double *__restrict__ kappa, uint32_t C, double *__restrict__ kd)
{
	const uint32_t lane = threadIdx.x % G;
	for (uint32_t c0 = blockIdx.x * (GB / G); c0 < C; c0 += gridDim.x * (GB / G)) {
		const uint32_t c = c0 + threadIdx.x / G;
		double acc = 0;
		if (c < C)
			for (uint32_t k = off[c] + lane; k < off[c + 1]; k += G) {
				const uint32_t s = ref[k];
				acc += x[s] * v[s];
			}
		acc = gsum<G>(acc);
		if (c < C && lane == 0)
			kd[c] = kappa[c] * acc;
	}
}

// HVP transcript pass: hv_t = v_t a_t - x_t sum_{c ni t} kd_c + N pv pi_t.
template <int G>
__global__ __launch_bounds__(GB) void k_htx(
    const uint32_t *__restrict__ off, const uint32_t *__restrict__ ref,
    const double *__restrict__ kd, const double *__restrict__ x,
    const double *__restrict__ a, const double *__restrict__ pi,
    const double *__restrict__ v, const double *__restrict__ pv_part,
    const double *__restrict__ scal, uint32_t T, double *__restrict__ hv)
{
	const double pv = fold(pv_part);
	const double N = scal[1];
	const uint32_t lane = threadIdx.x % G;
	for (uint32_t t0 = blockIdx.x * (GB / G); t0 < T; t0 += gridDim.x * (GB / G)) {
		const uint32_t t = t0 + threadIdx.x / G;
		double s = 0;
		if (t < T)
			for (uint32_t k = off[t] + lane; k < off[t + 1]; k += G)
				s += kd[ref[k]];
		s = gsum<G>(s);
		if (t < T && lane == 0)
			hv[t] = v[t] * a[t] - x[t] * s + N * pv * pi[t];
	}
}

} // namespace

extern "C" {

int hemk_grad_softmax(const double *theta, const double *ell, uint32_t T,
                      double *pi, double *w, double *x, float *x32,
                      double *part, void *s)
{
	hipStream_t st = (hipStream_t)s;
	k_gmax<<<NB, GB, 0, st>>>(theta, T, part);
	k_gexp<<<NB, GB, 0, st>>>(theta, ell, T, part, pi, w, part + NB);
	k_gnorm<<<NB, GB, 0, st>>>(pi, w, T, part + NB, x, x32);
	return (int)hipGetLastError();
}

int hemk_grad_passes(const hemk_prob *p, int fp32, const double *x,
                     const float *x32, const uint32_t *unique,
                     const double *pi, double *rho, double *D, double *kappa,
                     double *E, double *fd, double *gth, double *gell,
                     double *a, double *part, double *scal, void *s)
{
	hipStream_t st = (hipStream_t)s;
	double *pL1 = part, *pN = part + NB, *pL2 = part + 2 * NB;
	if (fp32) {
		k_gclass<4, float><<<NB, GB, 0, st>>>(p->cls_off, p->cls_ref,
		                                      p->cnt64, x32, p->n_cls, rho,
		                                      D, kappa, pL1);
		k_gtx<8, float><<<NB, GB, 0, st>>>(p->tx_off, p->tx_ref, rho, D,
		                                   p->cnt64, x32, unique, p->n_tx,
		                                   E, fd, pN, pL2);
	} else {
		k_gclass<4, double><<<NB, GB, 0, st>>>(p->cls_off, p->cls_ref,
		                                       p->cnt64, x, p->n_cls, rho,
		                                       D, kappa, pL1);
		k_gtx<8, double><<<NB, GB, 0, st>>>(p->tx_off, p->tx_ref, rho, D,
		                                    p->cnt64, x, unique, p->n_tx,
		                                    E, fd, pN, pL2);
	}
	k_gfin<<<NB, GB, 0, st>>>(E, pi, unique, p->n_tx, pN, pL1, pL2, NB,
	                          gth, gell, a, fd, scal);
	return (int)hipGetLastError();
}

int hemk_grad_fin(const double *E, const double *pi, const uint32_t *unique,
                  uint32_t T, double *part, uint32_t np, double *gth,
                  double *gell, double *a, double *fd, double *scal, void *s)
{
	hipStream_t st = (hipStream_t)s;
	double *tot = part + 3 * (size_t)np;
	// the three runs of partials to one value each first, so the blocks
	// of k_gfin do not each fold all of them
	k_gfold<<<3, GB, 0, st>>>(part, np, tot);
	k_gfin<<<NB, GB, 0, st>>>(E, pi, unique, T, tot + 1, tot, tot + 2, 1,
	                          gth, gell, a, fd, scal);
	return (int)hipGetLastError();
}

int hemk_hvp(const hemk_prob *p, const double *x, const double *a,
             const double *pi, const double *kappa, const double *scal,
             const double *v, double *kd, double *part, double *hv, void *s)
{
	hipStream_t st = (hipStream_t)s;
	const uint32_t gc = (p->n_cls + GB / 4 - 1) / (GB / 4);
	const uint32_t gt = (p->n_tx + GB / 8 - 1) / (GB / 8);
	k_gdot<<<NB, GB, 0, st>>>(pi, v, p->n_tx, part);
	k_hclass<4><<<gc ? gc : 1, GB, 0, st>>>(p->cls_off, p->cls_ref, x, v,
	                                        kappa, p->n_cls, kd);
	k_htx<8><<<gt ? gt : 1, GB, 0, st>>>(p->tx_off, p->tx_ref, kd, x, a, pi,
	                                     v, part, scal, p->n_tx, hv);
	return (int)hipGetLastError();
}

int hemk_grad_softmax_seg(const double *theta, const double *ell,
                          const uint32_t *info, uint32_t ns, double *pi,
                          double *w, double *x, float *x32, void *s)
{
	k_gsoftmax_seg<<<ns, GB, 0, (hipStream_t)s>>>(theta, ell, info, pi, w,
	                                              x, x32);
	return (int)hipGetLastError();
}

int hemk_grad_fin_seg(const double *E, const double *pi,
                      const uint32_t *unique, uint32_t T,
                      const uint32_t *txsamp, const uint32_t *info,
                      uint32_t ns, const double *part, uint32_t np,
                      double *tot, double *gth, double *gell, double *a,
                      double *fd, void *s)
{
	hipStream_t st = (hipStream_t)s;
	k_gfold_seg<<<ns, GB, 0, st>>>(part, np, info, tot);
	k_gfin_seg<<<NB, GB, 0, st>>>(E, pi, unique, T, txsamp, tot, gth, gell,
	                              a, fd);
	return (int)hipGetLastError();
}

int hemk_dot_parts(const double *a, const double *b, uint32_t T,
                   double *part, void *s)
{
	k_gdot<<<NB, GB, 0, (hipStream_t)s>>>(a, b, T, part);
	return (int)hipGetLastError();
}

} // extern "C"
