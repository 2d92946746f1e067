// spec_common.h -- device helpers shared by the exact-sum kernels
// (spec_sum.hip: speculation over candidate windows; walk_sum.hip: chunk
// programs and an in-order walk).  Included inside an anonymous namespace.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hem_dev.h"

#define SBLOCK 256
#define WAVE 64

namespace {

// Chain slot q of a launch -> the row it sums (rows of x are ld apart).
// Pairs: slots 2i, 2i+1 are rows i and R + i of replicate list entry i.
__device__ __forceinline__ bool chain_row(const hemk_spec &sp, uint32_t q,
                                          uint32_t &row)
{
	uint32_t rep = sp.pairs ? q >> 1 : q;
	if (sp.alist) {
		if (rep >= sp.alist[0])
			return false;
		rep = sp.alist[1 + rep];
	} else if (q >= sp.nchain) {
		return false;
	}
	row = sp.pairs ? (q & 1) * sp.R + rep : rep;
	return true;
}

// The value a chain adds at one position of its row: x, (x - m)^2 with m
// the row's mean (src/hash_align.c:156-159), or x * x.
template <int XF>
__device__ __forceinline__ double xval(double x, double m)
{
#pragma clang fp contract(off)
	if (XF == HEMK_SPEC_CENTRED)
		return (x - m) * (x - m);
	if (XF == HEMK_SPEC_SQUARE)
		return x * x;
	return x;
}

// m = sum / n, as `sum /= n` of src/hash_align.c:156.
template <int XF>
__device__ __forceinline__ double chain_mean(const hemk_spec &sp,
                                             uint32_t row)
{
	return XF == HEMK_SPEC_CENTRED ? sp.mres[row] / (double)sp.n : 0.0;
}

__device__ __forceinline__ bool skip(const hemk_ctl *ctl)
{
	return ctl && ctl->done;
}

// acc = acc + y in one FP64 add (round to nearest, as the compiler's own
// v_add_f64), kept in place: with the running sum pinned to one register
// the loop below needs no register copies between its buffers.
__device__ __forceinline__ void add_in_place(double &acc, double y)
{
	asm volatile("v_add_f64 %0, %0, %1" : "+v"(acc) : "v"(y));
}

// acc + x[0] + ... + x[cnt-1] left to right (per lane, x in LDS).  Two
// register buffers of sixteen values alternate: one is read from LDS while
// the other is added, so the reads' latency hides behind the adds.
__device__ __forceinline__ double run_chain(double acc, const double *x,
                                            uint32_t cnt)
{
#pragma clang fp contract(off)
	uint32_t j = 0;
	if (cnt >= 16) {
		double a[16], b[16];
#pragma unroll
		for (int q = 0; q < 16; ++q)
			a[q] = x[q];
		// a holds x[j-16 .. j-1], not yet added
		for (j = 16; j + 32 <= cnt; j += 32) {
#pragma unroll
			for (int q = 0; q < 16; ++q)
				b[q] = x[j + q];
#pragma unroll
			for (int q = 0; q < 16; ++q)
				add_in_place(acc, a[q]);
#pragma unroll
			for (int q = 0; q < 16; ++q)
				a[q] = x[j + 16 + q];
#pragma unroll
			for (int q = 0; q < 16; ++q)
				add_in_place(acc, b[q]);
		}
#pragma unroll
		for (int q = 0; q < 16; ++q)
			add_in_place(acc, a[q]);
	}
	for (; j < cnt; ++j)
		add_in_place(acc, x[j]);
	return acc;
}

__device__ __forceinline__ double readfirst_d(double v)
{
	const long long b = __double_as_longlong(v);
	const int lo = __builtin_amdgcn_readfirstlane((int)b);
	const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
	return __longlong_as_double(
	    (long long)(((unsigned long long)(unsigned)hi << 32) |
	                (unsigned)lo));
}

__device__ __forceinline__ long long readfirst_ll(long long b)
{
	const int lo = __builtin_amdgcn_readfirstlane((int)b);
	const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
	return (long long)(((unsigned long long)(unsigned)hi << 32) |
	                   (unsigned)lo);
}

// Stage cnt chain values of a row into ys with nthr threads, sixteen loads
// in flight per thread.
template <int XF>
__device__ __forceinline__ void stage(const double *__restrict__ x,
                                      uint32_t cnt, double m, double *ys,
                                      int nthr)
{
	const int t = threadIdx.x % nthr;
	uint32_t i = t;
	for (; i + 15 * nthr < cnt; i += 16 * nthr) {
		double v[16];
#pragma unroll
		for (int q = 0; q < 16; ++q)
			v[q] = x[i + q * nthr];
#pragma unroll
		for (int q = 0; q < 16; ++q)
			ys[i + q * nthr] = xval<XF>(v[q], m);
	}
	double v[16];
#pragma unroll
	for (int q = 0; q < 16; ++q)
		if (i + q * nthr < cnt)
			v[q] = x[i + q * nthr];
#pragma unroll
	for (int q = 0; q < 16; ++q)
		if (i + q * nthr < cnt)
			ys[i + q * nthr] = xval<XF>(v[q], m);
}

// One wave's chunk (cnt <= HEMK_SPEC_LMAX values) through registers:
// chunk_load only issues the loads (lane l takes values l, l + 64, ...), so
// whatever the caller loads next is in flight with them; chunk_store
// applies the transform and writes the values to LDS.
#define CH_REG (HEMK_SPEC_LMAX / WAVE)
__device__ __forceinline__ void chunk_load(const double *__restrict__ x,
                                           uint32_t cnt, double *v)
{
	const uint32_t lane = threadIdx.x % WAVE;
#pragma unroll
	for (int q = 0; q < CH_REG; ++q)
		if (lane + q * WAVE < cnt)
			v[q] = x[lane + q * WAVE];
}

template <int XF>
__device__ __forceinline__ void chunk_store(const double *v, uint32_t cnt,
                                            double m, double *ys)
{
	const uint32_t lane = threadIdx.x % WAVE;
#pragma unroll
	for (int q = 0; q < CH_REG; ++q)
		if (lane + q * WAVE < cnt)
			ys[lane + q * WAVE] = xval<XF>(v[q], m);
}

// Block sum of one double per thread (SBLOCK threads), fixed order; every
// thread gets the result.
__device__ __forceinline__ double block_sum(double a)
{
	__shared__ double red[SBLOCK / WAVE];
#pragma unroll
	for (int k = WAVE / 2; k >= 1; k >>= 1)
		a += __shfl_xor(a, k, WAVE);
	__syncthreads();
	if (threadIdx.x % WAVE == 0)
		red[threadIdx.x / WAVE] = a;
	__syncthreads();
	double t = red[0];
#pragma unroll
	for (int w = 1; w < SBLOCK / WAVE; ++w)
		t += red[w];
	return t;
}

} // namespace
