// em_step.h -- the SQUAREM step arithmetic shared by the plain kernels
// (src/EM.c:333-352), included by HIP sources only.
#ifndef EM_STEP_H_
#define EM_STEP_H_

#include <math.h>
#include <stdint.h>

#include "hem_dev.h"

// k after an iteration end, and the host's copy of it.
__device__ __forceinline__ void em_set_k(hemk_ctl *ctl, uint32_t k)
{
	ctl->k = k;
	if (ctl->hk)
		__hip_atomic_store(ctl->hk, k, __ATOMIC_RELAXED,
		                   __HIP_MEMORY_SCOPE_SYSTEM);
}

// The loop's end: done, converged, and the host's copy of done.
__device__ __forceinline__ void em_set_done(hemk_ctl *ctl, uint32_t stop)
{
	ctl->done = 1;
	ctl->converged = stop;
	if (ctl->hdone)
		__hip_atomic_store(ctl->hdone, 1u, __ATOMIC_RELAXED,
		                   __HIP_MEMORY_SCOPE_SYSTEM);
}

// The end of one SQUAREM iteration (src/EM.c:328 and 357): k++, the loop
// condition from the stop test of its last map (ctl->stopbad, cleared for
// the next iteration), done set when the loop ends.  One thread.
__device__ __forceinline__ void em_iter_end(hemk_ctl *ctl)
{
	const uint32_t stop = ctl->stopbad ? 0u : 1u;
	const uint32_t k = ctl->k + 1;
	ctl->stopbad = 0;
	em_set_k(ctl, k);
	ctl->maps += 3;
	if (!(k < ctl->min_iter || (k < ctl->max_iter && !stop)))
		em_set_done(ctl, stop);
}

// Step length from the two reductions (src/EM.c:338-346): sample sd (A, B
// = centred sums of squares, src/hash_align.c:158-163) or L2 (A, B = sums
// of squares, hera-EM/utility.c:16-23).
__device__ __forceinline__ double em_step_length(double A, double B,
                                                 uint32_t n, hemk_step st)
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

// The extrapolated point of one transcript: u - 2*alpha*r + alpha*alpha*v
// with r = u1 - u, v = (u2 - u1) - r, rounded exactly as src/EM.c:334-350
// (no contraction), then the clamp of src/EM.c:351 (ZERO) or
// hera-EM/utility.c:50-53 (DIV10).
__device__ __forceinline__ double em_extrapolate(double u, double u1,
                                                 double u2, double alpha,
                                                 hemk_step st)
{
#pragma clang fp contract(off)
	const double r = u1 - u;
	const double v = (u2 - u1) - r;
	double x = u - 2 * alpha * r + alpha * alpha * v;
	if (st.clamp == 0) {
		if (x < 0)
			x = 0;
	} else if (x <= 1e-300) {
		x = u2 / 10;
	}
	return x;
}

// a / b, bit for bit, when b lies in [2^-300, 2^300] and a is 0 or lies in
// [2^-300, 2^300]: the compiler's IEEE FP64 division is v_div_scale, v_rcp,
// two Newton steps on the reciprocal, a multiply, a residual fma, v_div_fmas
// and v_div_fixup; for such operands the scalings are identities (the
// exponents are far from the limits), v_div_fmas is a plain fma and the
// fixup only treats zero, infinite and NaN divisors, so the remaining
// steps give the same bits (checked on 2^32 hashed pairs in each of six
// exponent ranges up to those bounds, and a = 0).  Callers test the range
// once per workgroup and keep the division otherwise.
#define EM_QUOT_LO 0x1p-300
#define EM_QUOT_HI 0x1p300
__device__ __forceinline__ double em_quot(double a, double b)
{
#pragma clang fp contract(off)
	double r = __builtin_amdgcn_rcp(b);
	double e = __builtin_fma(-b, r, 1.0);
	r = __builtin_fma(r, e, r);
	e = __builtin_fma(-b, r, 1.0);
	r = __builtin_fma(r, e, r);
	const double q = a * r;
	const double rem = __builtin_fma(-b, q, a);
	return __builtin_fma(rem, r, q);
}

// Count, mean and centred sum of squares of a run of values, merged by
// the pairwise rule of Chan, Golub and LeVeque (1979); an empty side is
// the identity.
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

// Marks the stop test as failed: one atomic OR, skipped once the flag
// reads as set (an L2 read, so the check sees other blocks' ORs), which
// keeps thousands of failing blocks from serialising on one word.
__device__ __forceinline__ void em_mark_bad(uint32_t *flag)
{
	if (!__hip_atomic_load(flag, __ATOMIC_RELAXED,
	                       __HIP_MEMORY_SCOPE_AGENT))
		atomicOr(flag, 1u);
}

#endif
