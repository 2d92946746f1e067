/*
 * philox.h -- Philox-4x32-10 counter-based generator.
 *
 * Algorithm and constants from J. Salmon, M. Moraes, R. Dror, D. Shaw,
 * "Parallel random numbers: as easy as 1, 2, 3", SC'11: ten rounds of two
 * 32x32->64 multiplies (0xD2511F53, 0xCD9E8D57) with a Weyl key schedule
 * (0x9E3779B9, 0xBB67AE85).  A (key, counter) pair names a draw, so every
 * replicate, class or transcript has its own stream with no shared state.
 * Shared by the host C code and the HIP kernels.
 */
#ifndef HEM_PHILOX_H_
#define HEM_PHILOX_H_

#include <stdint.h>

#ifdef __HIPCC__
#define HEM_HD __host__ __device__ __forceinline__
#else
#define HEM_HD static inline
#endif

typedef struct {
	uint32_t v[4];
} hem_u32x4;

HEM_HD hem_u32x4 hem_philox(uint64_t key, uint32_t c0, uint32_t c1,
                            uint32_t c2, uint32_t c3)
{
	uint32_t k0 = (uint32_t)key, k1 = (uint32_t)(key >> 32);
	uint32_t x0 = c0, x1 = c1, x2 = c2, x3 = c3;
	for (int r = 0; r < 10; ++r) {
		const uint64_t p0 = (uint64_t)0xD2511F53u * x0;
		const uint64_t p1 = (uint64_t)0xCD9E8D57u * x2;
		const uint32_t y0 = (uint32_t)(p1 >> 32) ^ x1 ^ k0;
		const uint32_t y1 = (uint32_t)p1;
		const uint32_t y2 = (uint32_t)(p0 >> 32) ^ x3 ^ k1;
		const uint32_t y3 = (uint32_t)p0;
		x0 = y0;
		x1 = y1;
		x2 = y2;
		x3 = y3;
		k0 += 0x9E3779B9u;
		k1 += 0xBB67AE85u;
	}
	hem_u32x4 o = {{x0, x1, x2, x3}};
	return o;
}

/* Uniform double in (0, 1) from two 32-bit words (53 random bits). */
HEM_HD double hem_u01(uint32_t hi, uint32_t lo)
{
	const uint64_t x = (((uint64_t)hi << 32) | lo) >> 11;
	return ((double)x + 0.5) * (1.0 / 9007199254740992.0);
}

/* Stream key for (seed, stage). */
HEM_HD uint64_t hem_key(uint64_t seed, uint32_t stage)
{
	hem_u32x4 h = hem_philox(seed, stage, 0x68657261u, 0, 0);
	return ((uint64_t)h.v[0] << 32) | h.v[1];
}

#endif
