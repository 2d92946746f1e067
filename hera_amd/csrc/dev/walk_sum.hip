// walk_sum.hip -- left-to-right FP64 sums, bitwise, by chunk programs and an
// in-order walk.
//
// The reference adds its norm sums one value after another (Norm,
// src/hash_align.c:153-154 and 158-159).  Fix a unit u = 2^G.  While a
// running sum s = S u is a multiple of u below 2^53 u in magnitude, adding
// a value y gives
//   * (S + y / u) u exactly when y is a multiple of u and the result stays
//     below 2^53 u, whatever binade the sum is in, and
//   * (S + q) u, q = round(y / u) to the nearest integer, when the sums
//     before and after stay in the top binade [2^52 u, 2^53 u), whose ulp
//     is u, except at a tie (y / u = k + 1/2), which goes to the even one
//     of S + k and S + k + 1.
// So a run of such values adds a fixed integer D to S once the parity of S
// before its first tie is known: the run is cut at that tie, which is
// added by itself (the hardware's addition rounds it as the reference
// does), and the rest of the run starts from an even S (a tie always ends
// on an even integer).  Each value's unit is either its own binade's ulp
// (the run then stays in that binade) or, for a row of 64 values, the ulp
// of the binade just above every partial sum of the row (the run can then
// cross the binades below without a record, as sums near zero do).  A
// chunk of L values becomes a short program of records, each one
//   * a segment: such a run, summarised as D u, the interval of starts
//     [lo, hi] for which every partial sum stays where its rule needs it
//     with a margin of one unit, and, where it matters, the unit (or twice
//     the unit, after a cut) whose multiple the start must be, or
//   * a single value (one that crosses a binade off the grid, lies near a
//     binade boundary or zero, or is a run's first tie), added as it is.
// Which values form segments is decided from an approximate trajectory (the
// chunk's start guessed from the earlier chunks' plain sums, which the
// programs publish to each other); that choice only affects speed.
// Exactness comes from the walk: one wave adds the records of every chunk
// in order from +0, and a segment whose actual start lies outside [lo, hi]
// or off its grid is added value by value instead.  Adding D u to a valid
// start is exact and equals the run's additions, and a single value is the
// reference's own addition, so the walk's sum is the reference's, bit for
// bit, whatever the guesses were.
//
// On SQUAREM's r and v rows of 200,000 values a chunk of 1,024 values needs
// a handful of records: a few hundred per row early in a solve, up to about
// 2,000 (r) and 3,000 (v) late, when the running sums hover near zero,
// against 200,000 chained adds.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <climits>

#include "em_step.h"
#include "hem_dev.h"
#include "spec_common.h"

#define WS_T 256                         // threads per chunk program
#define WS_NW (WS_T / WAVE)
#define WS_E (HEMK_SPEC_LMAX / WS_T)     // values per thread, at most 4
#define WS_RB 2048                       // records per walk batch
#define WS_TAU 0x1p-30   // segment margin, relative to the largest |sum|
#define WS_SPIN (1u << 14)  // polls of a published chunk sum

namespace {

// biased exponent field of v
__device__ __forceinline__ uint32_t bexp(double v)
{
	return (uint32_t)((unsigned long long)__double_as_longlong(v) >> 52) &
	       0x7FFu;
}

// max(e, 1) + trailing zeros of the significand (hidden bit included) of
// a nonzero v with exponent field e: its lowest set bit is 2^(that - 1075)
__device__ __forceinline__ int low_code(unsigned long long b)
{
	const int e = (int)((b >> 52) & 0x7FF);
	unsigned long long m = b & 0xFFFFFFFFFFFFFull;
	if (e)
		m |= 1ull << 52;
	return max(e, 1) + (int)__builtin_ctzll(m | (1ull << 63));
}

// y is +0, or a finite nonzero multiple of 2^(code - 1075)
__device__ __forceinline__ bool on_grid(double y, int code)
{
	const unsigned long long b = (unsigned long long)__double_as_longlong(y);
	return (b == 0) | (((b << 1) != 0) & (((b >> 52) & 0x7FF) != 0x7FF) &
	                   (low_code(b) >= code));
}

// a start s that is a signed zero, a NaN or a multiple of 2^(code - 1075)
// (s scaled by 2^(1075 - code) is an integer; a nonzero multiple scales to
// at least 1, so a scaled value that underflows to 0 is not one)
__device__ __forceinline__ bool start_on_grid(double s, int code)
{
	const double t = ldexp(s, 1075 - code);
	return ((t == __builtin_trunc(t)) & ((t != 0.0) | (s == 0.0))) |
	       (s != s);
}

// ---- scans across lanes (DPP moves) and across the workgroup --------------

// A trivially copyable value moved between lanes 32 bits at a time.
template <typename T>
union words {
	T t;
	uint32_t w[sizeof(T) / 4];
};

// DPP move of v (old where the source lane is outside the row or the row
// is masked off).
template <int CTRL, int ROWM, typename T>
__device__ __forceinline__ T dpp_mov(const T &old, const T &v)
{
	words<T> a, b, r;
	a.t = old;
	b.t = v;
#pragma unroll
	for (int i = 0; i < (int)(sizeof(T) / 4); ++i)
		r.w[i] = (uint32_t)__builtin_amdgcn_update_dpp(
		    (int)a.w[i], (int)b.w[i], CTRL, ROWM, 0xF, false);
	return r.t;
}

#define DPP_ROW_SHR(n) (0x110 + (n))
#define DPP_ROW_ROR(n) (0x120 + (n))
#define DPP_WAVE_SHR1 0x138
#define DPP_ROW_BCAST15 0x142
#define DPP_ROW_BCAST31 0x143

// Inclusive scan over the wave's lanes of an associative op (op(earlier,
// later)) with identity id: shifts of 1, 2, 4, 8 lanes inside each row of
// 16, then row totals broadcast to the rows after them.
template <typename T, typename Op>
__device__ __forceinline__ T wave_incl(T v, const T &id, Op op)
{
	v = op(dpp_mov<DPP_ROW_SHR(1), 0xF>(id, v), v);
	v = op(dpp_mov<DPP_ROW_SHR(2), 0xF>(id, v), v);
	v = op(dpp_mov<DPP_ROW_SHR(4), 0xF>(id, v), v);
	v = op(dpp_mov<DPP_ROW_SHR(8), 0xF>(id, v), v);
	v = op(dpp_mov<DPP_ROW_BCAST15, 0xA>(id, v), v);
	v = op(dpp_mov<DPP_ROW_BCAST31, 0xC>(id, v), v);
	return v;
}

// Exclusive scan over the workgroup (WS_NW waves) and the total; wt is a
// WS_NW-entry LDS buffer of this scan's own.
template <typename T, typename Op>
__device__ __forceinline__ T block_excl(T v, const T &id, Op op, T *wt,
                                        T &total)
{
	const int w = threadIdx.x / WAVE;
	const T inc = wave_incl(v, id, op);
	const T ex = dpp_mov<DPP_WAVE_SHR1, 0xF>(id, inc);
	if (threadIdx.x % WAVE == WAVE - 1)
		wt[w] = inc;
	__syncthreads();
	T pre = id, all = id;
#pragma unroll
	for (int i = 0; i < WS_NW; ++i) {
		const T x = wt[i];
		if (i < w)
			pre = op(pre, x);
		all = op(all, x);
	}
	total = all;
	return op(pre, ex);
}

// Segmented carries.  Each thread summarises its values as (h: a segment
// starts among them, then the part since its last start; otherwise the
// part over all of them); the exclusive scan gives every thread the state
// carried in from the threads before it.
struct seg_sum {
	uint32_t h;
	int hp;          // position of the last segment start (-1: none),
	                 // + WS_EVEN when that segment needs an even start
	uint32_t x, pad; // x: an exact row-grid value since the start
	long long acc;   // increments since the last segment start
	long long lo, hi; // bounds on the segment's start since its start
};
#define WS_EVEN 0x40000000

// record count, and the parity transfer (parity after = d ? in ^ v : v)
struct cnt_par {
	uint32_t n, d, v;
};

// a segmented flag: a stretch starts among the values (h), then whether the
// flag (a tie, or an anchor) has been seen since the last start (or over
// all the values)
struct tie_seen {
	uint32_t h, t;
};

struct sum_abs {
	double s, a, q;   // sum, sum of magnitudes, sum of squares
};

// Program of chunk j of each chain: one workgroup of WS_T threads per
// chunk, thread t holding values E t .. E t + E - 1 (E = L / WS_T).  Grid
// (P, chains).
template <int XF>
__global__ __launch_bounds__(WS_T) void k_ws_prog(hemk_spec sp,
                                                  const hemk_ctl *ctl)
{
#pragma clang fp contract(off)
	__shared__ double w_g[WS_NW], w_m[WS_NW];
	__shared__ sum_abs w_t[WS_NW];
	__shared__ cnt_par w_c[WS_NW];
	__shared__ tie_seen w_f[WS_NW];
	__shared__ seg_sum w_s[WS_NW];
	__shared__ tie_seen w_a[WS_NW];
	__shared__ uint32_t sg[WS_T], sl[WS_T], sf[WS_T], sl2[WS_T];
	uint32_t row;
	if (!chain_row(sp, blockIdx.y, row))
		return;
	const int t = threadIdx.x;
	const uint32_t j = blockIdx.x;
	const size_t cb = (size_t)blockIdx.y * sp.P;
	const uint32_t cnt = min(sp.L, sp.n - j * sp.L);
	const uint32_t E = sp.L / WS_T;
	const uint32_t e0 = t * E;
	const double *x = sp.x + (size_t)row * sp.ld + (size_t)j * sp.L;
	const bool seeded = XF == HEMK_SPEC_CENTRED && sp.gsq;
	// programs that publish their chunk sums (cpub) read the earlier ones
	// once published; otherwise the first 256 earlier chunk sums are
	// requested with the chunk's values, before the control word is read,
	// so the loads overlap
	const bool pub = sp.cpub && !seeded;
	double y[WS_E];
	uint32_t vmask = 0;
#pragma unroll
	for (int k = 0; k < WS_E; ++k) {
		y[k] = 0.0;
		if ((uint32_t)k < E && e0 + k < cnt) {
			y[k] = x[e0 + k];
			vmask |= 1u << k;
		}
	}
	double c1 = 0.0, c2 = 0.0;
	uint32_t ep0 = 0;
	if (pub) {
		// the epoch of the last launch: only thread 0 reads it, as only
		// thread 0 writes this word, after the read
		if (t == 0)
			ep0 = (uint32_t)sp.cpub[cb + j] & 0xFFFFu;
	} else if ((uint32_t)t < j) {
		if (seeded) {
			c1 = sp.gsq[cb + t];
			c2 = sp.gsum[cb + t];
		} else {
			c1 = sp.csum[cb + t];
		}
	}
	if (skip(ctl))
		return;
	const double m = chain_mean<XF>(sp, row);
#pragma unroll
	for (int k = 0; k < WS_E; ++k)
		if ((vmask >> k) & 1u)
			y[k] = xval<XF>(y[k], m);
	const auto fadd = [](double a, double b) {
#pragma clang fp contract(off)
		return a + b;
	};
	const auto fmx = [](double a, double b) { return fmax(a, b); };

	// the chunk's own sums (any order), each thread's partial sums before
	// its values, and with publishing the chunk sum and the epoch
	__shared__ uint32_t sh_ep;
	double tl = 0.0;
	double a[WS_E];
	sum_abs tt, ex;
	{
		double ta = 0.0, tq = 0.0;
#pragma unroll
		for (int k = 0; k < WS_E; ++k) {
			tl += y[k];
			a[k] = tl;
			ta += fabs(y[k]);
			tq += y[k] * y[k];
		}
		const auto pad = [](sum_abs p, sum_abs q) {
#pragma clang fp contract(off)
			sum_abs r;
			r.s = p.s + q.s;
			r.a = p.a + q.a;
			r.q = p.q + q.q;
			return r;
		};
		if (pub && t == 0)
			sh_ep = (ep0 + 1u) & 0xFFFFu;
		ex = block_excl(sum_abs{tl, ta, tq}, sum_abs{0.0, 0.0, 0.0}, pad,
		                w_t, tt);
		// the chunk's sum of squares, for the guesses of a centred pass
		if (XF == HEMK_SPEC_IDENT && sp.csq && t == 0)
			sp.csq[cb + j] = tt.q;
	}
	const uint32_t epoch = pub ? sh_ep : 0u;
	// one word carries the sum and the epoch, so a relaxed store and
	// relaxed reads suffice (no fence, which would write back the L2)
	if (pub && t == 0)
		__hip_atomic_store(sp.cpub + cb + j,
		                   ((unsigned long long)__double_as_longlong(tt.s) &
		                    ~0xFFFFull) | epoch,
		                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);

	// guessed start g of the chunk (the earlier chunks' sums, any order)
	// and the largest magnitude M of the earlier chunk starts.  A
	// published sum is taken once its word holds this launch's epoch; the
	// wait is bounded, and a sum not seen in time counts as 0 (the guess
	// only steers which values form runs)
	double g = 0.0, M = 0.0;
	{
		const double lm2 = (double)sp.L * m * m;
		for (uint32_t i0 = 0; i0 < j; i0 += WS_T) {
			const uint32_t i = i0 + t;
			if (pub) {
				c1 = 0.0;
				if (i < j) {
					uint32_t spin = 0;
					unsigned long long w;
					for (;;) {
						w = __hip_atomic_load(sp.cpub + cb + i,
						                      __ATOMIC_RELAXED,
						                      __HIP_MEMORY_SCOPE_AGENT);
						if ((w & 0xFFFFull) == epoch || ++spin >= WS_SPIN)
							break;
						__builtin_amdgcn_s_sleep(1);
					}
					if ((w & 0xFFFFull) == epoch)
						c1 = __longlong_as_double((long long)w);
				}
			} else if (i0 && i < j) {
				if (seeded) {
					c1 = sp.gsq[cb + i];
					c2 = sp.gsum[cb + i];
				} else {
					c1 = sp.csum[cb + i];
				}
			}
			double c = 0.0;
			if (i < j)
				c = seeded ? (c1 - 2.0 * m * c2) + lm2 : c1;
			double tot;
			const double exg = block_excl(c, 0.0, fadd, w_g, tot);
			if (i < j)
				M = fmax(M, fabs(g + exg));
			g += tot;
			__syncthreads();
		}
		double mt;
		(void)block_excl(M, 0.0, fmx, w_m, mt);
		M = fmax(mt, fabs(g));
	}

	// approximate partial sums a_k after each value, and a bound on every
	// partial sum's magnitude in this chunk
	double prev;
	double tau;
	{
		prev = g + ex.s;
#pragma unroll
		for (int k = 0; k < WS_E; ++k)
			a[k] = g + (ex.s + a[k]);
		tau = (M + fabs(g) + tt.a) * WS_TAU;
	}

	// Each value is summarised on one of two grids.
	//  * Its binade's: while the partial sums before and after it stay in
	//    one binade (sign and exponent) with the margin tau, the value adds
	//    q = round(y / u) ulps u of that binade.
	//  * Its row's: gc is the largest binade code of this DPP row's partial
	//    sums (16 threads, 16 E values), margin included, so every partial
	//    sum is below 2^53 units u = 2^(gc - 1075).  A value that is a
	//    multiple of u (or +0) adds exactly y / u units wherever the sum is
	//    below that bound, and a value whose partial sums stay in the top
	//    binade (gc) adds round(y / u) as above, so a run of such values
	//    can cross the binades below the top one without a record.  Its
	//    start must be a multiple of u: a row-grid run begins at an anchor
	//    (a value whose partial sum before lies in the top binade) and goes
	//    on over the values after it with the same grid and the same sign
	//    of the partial sums on both sides, which keeps its total below
	//    2^53 units.
	// A round-half tie (only a top-binade value can be one) is kept at
	// floor for now: it rounds to the even neighbour, so it depends on the
	// running sum's last bit.
	// (only for plain sums: squares and centred squares only grow, so
	// their runs never cross binades on a row's grid)
	constexpr bool GRID = XF == HEMK_SPEC_IDENT;
	uint32_t gc = 0;
	if (GRID) {
#pragma unroll
		for (int k = 0; k < WS_E; ++k) {
			const double pk = k ? a[k - 1] : prev;
			gc = max(gc, bexp(fabs(pk) + tau));
			gc = max(gc, bexp(fabs(a[k]) + tau));
		}
		gc = max(gc, dpp_mov<DPP_ROW_ROR(8), 0xF>(gc, gc));
		gc = max(gc, dpp_mov<DPP_ROW_ROR(4), 0xF>(gc, gc));
		gc = max(gc, dpp_mov<DPP_ROW_ROR(2), 0xF>(gc, gc));
		gc = max(gc, dpp_mov<DPP_ROW_ROR(1), 0xF>(gc, gc));
	}
	const bool gok = gc >= 2 && gc <= 2044;
	const double gtop = __longlong_as_double((long long)gc << 52);
	uint32_t bmask = 0, btie = 0;   // binade grid: usable values, ties
	uint32_t emask_x = 0;           // multiples of the row's unit (or +0)
	uint32_t amask = 0, nmask = 0;  // anchors; partial sum after < 0
	uint32_t gkey[WS_E], bkey[WS_E];
	uint32_t glast = 0;   // gkey of the thread's last value
	long long q[WS_E], qx[WS_E];
#pragma unroll
	for (int k = 0; k < WS_E; ++k) {
		const double ak = a[k], pk = k ? a[k - 1] : prev;
		const uint32_t ka =
		    (uint32_t)((unsigned long long)__double_as_longlong(ak) >> 52);
		const uint32_t kp =
		    (uint32_t)((unsigned long long)__double_as_longlong(pk) >> 52);
		const uint32_t be = ka & 0x7FFu;
		const bool valid = (vmask >> k) & 1u;
		bool s = valid && ka == kp && be >= 2 && be <= 2044;
		const double blo = __longlong_as_double((long long)be << 52);
		const double bhi = __longlong_as_double((long long)(be + 1) << 52);
		const double A = fabs(ak), B = fabs(pk);
		s = s && A - blo > tau && bhi - A > tau && B - blo > tau &&
		    bhi - B > tau;
		const double xs = ldexp(y[k], 1075 - (int)be);
		const double kf = floor(xs);
		const double f = xs - kf;
		s = s && fabs(xs) < 0x1p52;
		bkey[k] = ka;
		q[k] = s ? (long long)kf + (f > 0.5 ? 1 : 0) : 0;
		bmask |= (uint32_t)s << k;
		btie |= (uint32_t)(s && f == 0.5) << k;
		const bool same = (ak > 0.0 && pk > 0.0) || (ak < 0.0 && pk < 0.0);
		const bool xk = gok && valid && same && on_grid(y[k], (int)gc);
		const bool rk = s && be == gc;
		const bool ge = xk || rk;
		qx[k] = xk ? (long long)ldexp(y[k], 1075 - (int)gc) : 0;
		emask_x |= (uint32_t)xk << k;
		nmask |= (uint32_t)(ak < 0.0) << k;
		gkey[k] = ge ? ((gc << 1) | (uint32_t)(ak < 0.0)) + 1u : 0u;
		if ((uint32_t)k + 1 == E)
			glast = gkey[k];
		amask |= (uint32_t)(ge && B - tau >= gtop) << k;
	}
	// row-grid values: those of a stretch (consecutive values with the
	// same nonzero gkey) from its first anchor on
	uint32_t gmask = 0;
	if (GRID) {
		sg[t] = glast;
		__syncthreads();
		const auto tsop = [](tie_seen l, tie_seen r) {
			tie_seen o;
			o.h = l.h | r.h;
			o.t = r.h ? r.t : (l.t | r.t);
			return o;
		};
		uint32_t pkey = t ? sg[t - 1] : 0u, hm = 0;
		tie_seen la = {0u, 0u};
#pragma unroll
		for (int k = 0; k < WS_E; ++k) {
			if (gkey[k] && gkey[k] != pkey) {
				la.h = 1;
				la.t = 0;
				hm |= 1u << k;
			}
			la.t |= (amask >> k) & 1u;
			pkey = gkey[k];
		}
		tie_seen atot;
		const tie_seen ca = block_excl(la, tie_seen{0u, 0u}, tsop, w_a, atot);
		uint32_t seen = ca.t;
#pragma unroll
		for (int k = 0; k < WS_E; ++k) {
			if ((hm >> k) & 1u)
				seen = 0;
			seen |= (amask >> k) & 1u;
			if (gkey[k] && seen)
				gmask |= 1u << k;
		}
	}
	// labels: the unit's binade code and the sign of the partial sums
	// (values of both kinds on one unit form one run), 0 for a value added
	// by itself; units and increments; X marks exact row-grid values,
	// whose partial sums only need to stay below 2^53 units
	uint32_t lab[WS_E], code[WS_E];
	uint32_t llast = 0;   // label of the thread's last value
	uint32_t cmask = 0, tmask = 0, xmask = 0;
#pragma unroll
	for (int k = 0; k < WS_E; ++k) {
		const uint32_t bit = 1u << k;
		lab[k] = 0;
		code[k] = 0;
		if (gmask & bit) {
			lab[k] = gkey[k];
			code[k] = gc;
			if (emask_x & bit) {
				q[k] = qx[k];
				xmask |= bit;
			} else {
				tmask |= btie & bit;
			}
		} else if (bmask & bit) {
			code[k] = bkey[k] & 0x7FFu;
			lab[k] = ((code[k] << 1) | (bkey[k] >> 11)) + 1u;
			tmask |= btie & bit;
		}
		cmask |= (uint32_t)(lab[k] != 0) << k;
		if ((uint32_t)k + 1 == E)
			llast = lab[k];
	}
	// the first tie of each candidate run (consecutive values with the
	// same label) is added by itself (its rounding needs the true start's
	// last bit); after it the running sum is even, so the rest of the run
	// is a segment whose start must be even, its later ties settled by the
	// parity below
	sl[t] = llast;
	__syncthreads();
	uint32_t split;   // first ties of candidate runs
	{
		uint32_t plab = t ? sl[t - 1] : 0u, ch = 0;
#pragma unroll
		for (int k = 0; k < WS_E; ++k) {
			if (lab[k] && lab[k] != plab)
				ch |= 1u << k;
			plab = lab[k];
		}
		tie_seen lt = {0u, 0u};
#pragma unroll
		for (int k = 0; k < WS_E; ++k) {
			if ((ch >> k) & 1u) {
				lt.h = 1;
				lt.t = 0;
			}
			lt.t |= (tmask >> k) & 1u;
		}
		const auto tsop = [](tie_seen l, tie_seen r) {
			tie_seen o;
			o.h = l.h | r.h;
			o.t = r.h ? r.t : (l.t | r.t);
			return o;
		};
		tie_seen ttot;
		const tie_seen ct =
		    block_excl(lt, tie_seen{0u, 0u}, tsop, w_f, ttot);
		uint32_t seen = ct.t;
		split = 0;
#pragma unroll
		for (int k = 0; k < WS_E; ++k) {
			if ((ch >> k) & 1u)
				seen = 0;
			if ((tmask >> k) & 1u) {
				if (!seen)
					split |= 1u << k;
				seen = 1;
			}
		}
	}
	const uint32_t smask = cmask & ~split;
	tmask &= ~split;
	// segment starts (in smask after a value that is not, or after one
	// with another label; an even start when that value is a split tie),
	// ends (before such a value) and records (values not in smask, and
	// starts)
	sf[t] = (lab[0] << 2) | (smask & 1u);
	sl2[t] = (llast << 2) | ((smask >> (E - 1)) & 1u) |
	         (((split >> (E - 1)) & 1u) << 1);
	__syncthreads();
	uint32_t hmask = 0, evmask = 0, emask = 0;
	{
		const uint32_t pw = t ? sl2[t - 1] : 0u;
		const uint32_t nw = t + 1 < WS_T ? sf[t + 1] : 0u;
		uint32_t pin = pw & 1u, psp = (pw >> 1) & 1u, plab = pw >> 2;
#pragma unroll
		for (int k = 0; k < WS_E; ++k) {
			const uint32_t bit = 1u << k;
			const bool in = smask & bit;
			if (in && (!pin || plab != lab[k])) {
				hmask |= bit;
				if (psp)
					evmask |= bit;
			}
			const bool inside = k + 1 < WS_E && (uint32_t)k + 1 < E;
			const uint32_t nin = inside ? (smask >> (k + 1)) & 1u : nw & 1u;
			const uint32_t nlab =
			    inside ? lab[k + 1 < WS_E ? k + 1 : k] : nw >> 2;
			if (in && (!nin || nlab != lab[k]))
				emask |= bit;
			pin = in;
			psp = (split >> k) & 1u;
			plab = lab[k];
		}
	}
	const uint32_t rmask = vmask & (~smask | hmask);

	// parity before each segment value for a segment started at an even
	// integer: reset at a segment start and after a tie (a tie always
	// ends on an even integer), flipped by each odd increment; before
	// this thread's first reset it is the carried-in parity c xor pval.
	// Other values pass it on unchanged (the next segment value after one
	// starts a segment, which resets it), and so do the empty slots past
	// E when a thread holds fewer than WS_E values.
	uint32_t pdep = 0, pval = 0;
	uint32_t dep = 1, pv = 0;
#pragma unroll
	for (int k = 0; k < WS_E; ++k) {
		if (!((smask >> k) & 1u))
			continue;
		if ((hmask >> k) & 1u) {
			dep = 0;
			pv = 0;
		}
		pdep |= dep << k;
		pval |= pv << k;
		if ((tmask >> k) & 1u) {
			dep = 0;
			pv = 0;
		} else {
			pv ^= (uint32_t)q[k] & 1u;
		}
	}
	const auto cpop = [](cnt_par l, cnt_par r) {
		cnt_par o;
		o.n = l.n + r.n;
		o.d = l.d & r.d;
		o.v = r.d ? (l.v ^ r.v) : r.v;
		return o;
	};
	cnt_par ctot;
	const cnt_par cex =
	    block_excl(cnt_par{(uint32_t)__popc(rmask), dep, pv},
	               cnt_par{0u, 1u, 0u}, cpop, w_c, ctot);
	const uint32_t rbase = cex.n, nrec = ctot.n;
	const uint32_t c = cex.v;   // parity carried into this thread
	const uint32_t cap = sp.L / 2;
	if (nrec > cap) {
		if (t == 0)
			sp.wcnt[cb + j] = HEMK_WS_RAW;
		return;
	}
	if (t == 0)
		sp.wcnt[cb + j] = nrec;
	// ties inside segments: q = k + ((parity + k) & 1)
#pragma unroll
	for (int k = 0; k < WS_E; ++k) {
		if ((tmask >> k) & 1u) {
			const uint32_t pb = ((pdep >> k) & 1u)
			                        ? (c ^ ((pval >> k) & 1u))
			                        : ((pval >> k) & 1u);
			q[k] += (pb + (uint32_t)q[k]) & 1u;
		}
	}

	// increments since each segment start (P) and the start's position,
	// carried across threads
	// Increments since each segment start (P), the start's position, and
	// the starts S (in units) each segment value allows: with P before and
	// P' after it, S + P and S + P' in [2^52 + 1, 2^53 - 1] (negated for
	// a negative binade) for a value rounded in its binade, and
	// |S + P'| <= 2^53 - 1 for an exact row-grid value (with P' = P + q:
	// S >= L - min(0, q) - P and S <= U - max(0, q) - P); lo and hi are
	// the tightest bounds since the segment's start.  Each thread's part
	// is relative to P at its first value (or at its last segment start);
	// the scan shifts a later part's bounds by the P of the part before.
	const long long B52 = 1LL << 52, B53 = 1LL << 53;
	const long long NOB = 1LL << 62;   // no bound yet (|bounds| < 2^56)
	long long wl[WS_E], wu[WS_E];
#pragma unroll
	for (int k = 0; k < WS_E; ++k) {
		const bool xk = (xmask >> k) & 1u, nk = (nmask >> k) & 1u;
		const long long L = xk || nk ? -(B53 - 1) : B52 + 1;
		const long long U = xk || !nk ? B53 - 1 : -(B52 + 1);
		wl[k] = L - min(0LL, q[k]);
		wu[k] = U - max(0LL, q[k]);
	}
	const auto bounds = [&](uint32_t k, long long p0, long long &lo,
	                        long long &hi) {
		lo = max(lo, wl[k] - p0);
		hi = min(hi, wu[k] - p0);
	};
	seg_sum ls = {0u, -1, 0u, 0u, 0, -NOB, NOB};
#pragma unroll
	for (int k = 0; k < WS_E; ++k) {
		if (!((smask >> k) & 1u))
			continue;
		if ((hmask >> k) & 1u) {
			ls.h = 1;
			ls.hp = (int)(e0 + k) | (((evmask >> k) & 1u) ? WS_EVEN : 0);
			ls.x = 0;
			ls.acc = 0;
			ls.lo = -NOB;
			ls.hi = NOB;
		}
		bounds(k, ls.acc, ls.lo, ls.hi);
		ls.acc += q[k];
		ls.x |= (xmask >> k) & 1u;
	}
	const auto ssop = [](seg_sum l, seg_sum r) {
		seg_sum o;
		o.h = l.h | r.h;
		o.hp = r.h ? r.hp : l.hp;
		o.x = r.h ? r.x : (l.x | r.x);
		o.pad = 0;
		o.acc = r.h ? r.acc : l.acc + r.acc;
		o.lo = r.h ? r.lo : max(l.lo, r.lo - l.acc);
		o.hi = r.h ? r.hi : min(l.hi, r.hi - l.acc);
		return o;
	};
	seg_sum stot;
	const seg_sum cs = block_excl(
	    ls, seg_sum{0u, -1, 0u, 0u, 0, -NOB, NOB}, ssop, w_s, stot);

	// records: single values at their own index, a segment at its
	// start's index (no record lies between a segment's start and end).
	// A segment adds D = P(end) units for starts S in [lo, hi] that are
	// multiples of its unit (of twice the unit after a split tie); the
	// record keeps its first value's index and, when the start's grid
	// needs a check (an exact row-grid value or a split tie: any start in
	// [lo, hi] of a run rounded in its binade is on the unit's grid), the
	// binade code of the unit it must be a multiple of, plus one
	seg_sum r = cs;
	long long slo = cs.lo, shi = cs.hi;
	uint32_t sx = cs.x;
	uint32_t nr = 0;   // records at this thread's values so far
	const size_t rb0 = (cb + j) * cap + rbase;
	double2 *wlh = (double2 *)sp.wlh;
#pragma unroll
	for (int k = 0; k < WS_E; ++k) {
		const uint32_t bit = 1u << k;
		if (smask & bit) {
			if (hmask & bit) {
				r.acc = 0;
				r.hp = (int)(e0 + k) | ((evmask & bit) ? WS_EVEN : 0);
				slo = LLONG_MIN;
				shi = LLONG_MAX;
				sx = 0;
			}
			bounds(k, r.acc, slo, shi);
			r.acc += q[k];
			sx |= (xmask >> k) & 1u;
		}
		if (rmask & bit)
			++nr;
		if ((rmask & bit) && !(smask & bit)) {
			const size_t ri = rb0 + nr - 1;
			sp.wa[ri] = y[k];
			wlh[ri] = make_double2(-__builtin_inf(), __builtin_inf());
			sp.wrg[ri] = e0 + k;
		}
		if (emask & bit) {
			const size_t ri = rb0 + nr - 1;
			const int sh = (int)code[k] - 1075;
			double lo = __builtin_inf(), hi = -__builtin_inf();
			double da = 0.0;
			if (slo <= shi && r.acc > -B53 && r.acc < B53) {
				lo = ldexp((double)slo, sh);
				hi = ldexp((double)shi, sh);
				da = ldexp((double)r.acc, sh);
			}
			const uint32_t ev = (r.hp & WS_EVEN) ? 1u : 0u;
			const uint32_t chk = sx | ev ? code[k] + ev + 1u : 0u;
			sp.wa[ri] = da;
			wlh[ri] = make_double2(lo, hi);
			sp.wrg[ri] = (uint32_t)(r.hp & 0x7FF) | (chk << 11);
		}
	}
}

// Add records r0 .. nr-1 to s (every lane of the calling wave the same),
// keeping the start of each group of WS_WB records (from r0) in sb[g]; sb
// is lane 0's array, the other lanes write the same values to their own
// slots of a sink, so the stores need no lane mask.  Two register buffers
// of WS_WB addends alternate, the scheduling barriers keeping the next
// buffer's LDS reads ahead of the current buffer's adds.
#define WS_WB 16
__device__ __forceinline__ double ws_walk(double s, const double *a,
                                          double *sb, double *sink,
                                          uint32_t r0, uint32_t nr)
{
#pragma clang fp contract(off)
	const int lane = threadIdx.x % WAVE;
	double *out = lane ? sink + lane : sb;
	uint32_t r = r0, b = 0;
	if (r + WS_WB <= nr) {
		double p[WS_WB], q[WS_WB];
#pragma unroll
		for (int i = 0; i < WS_WB; ++i)
			p[i] = a[r + i];
		// p holds records r .. r + WS_WB - 1, not yet added
		while (r + 3 * WS_WB <= nr) {
#pragma unroll
			for (int i = 0; i < WS_WB; ++i)
				q[i] = a[r + WS_WB + i];
			__builtin_amdgcn_sched_barrier(0);
			out[b++] = s;
#pragma unroll
			for (int i = 0; i < WS_WB; ++i)
				s = s + p[i];
			r += WS_WB;
			__builtin_amdgcn_sched_barrier(0);
#pragma unroll
			for (int i = 0; i < WS_WB; ++i)
				p[i] = a[r + WS_WB + i];
			__builtin_amdgcn_sched_barrier(0);
			out[b++] = s;
#pragma unroll
			for (int i = 0; i < WS_WB; ++i)
				s = s + q[i];
			r += WS_WB;
			__builtin_amdgcn_sched_barrier(0);
		}
		if (r + 2 * WS_WB <= nr) {
#pragma unroll
			for (int i = 0; i < WS_WB; ++i)
				q[i] = a[r + WS_WB + i];
			out[b++] = s;
#pragma unroll
			for (int i = 0; i < WS_WB; ++i)
				s = s + p[i];
			r += WS_WB;
			out[b++] = s;
#pragma unroll
			for (int i = 0; i < WS_WB; ++i)
				s = s + q[i];
			r += WS_WB;
		} else {
			out[b++] = s;
#pragma unroll
			for (int i = 0; i < WS_WB; ++i)
				s = s + p[i];
			r += WS_WB;
		}
	}
	if (r < nr) {
		out[b] = s;
		for (; r < nr; ++r)
			s = s + a[r];
	}
	return s;
}

// One workgroup per chain: batches of chunks whose records fit WS_RB are
// staged in LDS, wave 0 walks them, then every thread checks the segments
// of one group of records from the group's start (the same additions as
// the walk); the first record whose start lies outside its interval is
// added value by value and the walk resumes after it.  A chunk without a
// program (too many records) is added value by value.
template <int XF>
__global__ __launch_bounds__(SBLOCK) void k_ws_walk(hemk_spec sp,
                                                    const hemk_ctl *ctl)
{
#pragma clang fp contract(off)
	constexpr int RPT = WS_RB / SBLOCK;   // records staged per thread
	constexpr int NG = WS_RB / WS_WB + 1; // walk groups of a batch
	__shared__ double la[WS_RB];
	__shared__ double2 llh[WS_RB];
	__shared__ uint32_t lrg[WS_RB];
	__shared__ uint8_t lch[WS_RB];
	__shared__ double lsb[NG];
	__shared__ double sink[NG + WAVE];
	__shared__ double ys[HEMK_SPEC_LMAX];
	__shared__ uint32_t loff[SBLOCK + 1];
	__shared__ uint32_t wsum[SBLOCK / WAVE];
	__shared__ uint32_t sh_fail;
	__shared__ double sh_fs;
	const int tid = threadIdx.x;
	const size_t cb = (size_t)blockIdx.x * sp.P;
	// the first batch's counts are requested before the control word is
	// read, so the two loads overlap
	uint32_t c = (uint32_t)tid < sp.P ? sp.wcnt[cb + tid] : HEMK_WS_RAW;
	uint32_t row;
	if (skip(ctl) || !chain_row(sp, blockIdx.x, row))
		return;
	const bool walker = tid < WAVE;
	const uint32_t cap = sp.L / 2;
	const double *xr = sp.x + (size_t)row * sp.ld;
	const auto uadd = [](uint32_t a, uint32_t b) { return a + b; };
	double s = 0.0;
	uint32_t reruns = 0;
	for (uint32_t j = 0; j < sp.P;) {
		// the batch: chunks j .. j+nb-1 whose records fit WS_RB; a chunk
		// without a program closes it (and is a batch of its own)
		if (j) {
			const uint32_t jc = j + tid;
			c = jc < sp.P ? sp.wcnt[cb + jc] : HEMK_WS_RAW;
		}
		const uint32_t cc = c > cap ? WS_RB + 1 : c;
		const uint32_t w = tid / WAVE;
		uint32_t inc = wave_incl(cc, 0u, uadd);
		if (tid % WAVE == WAVE - 1)
			wsum[w] = inc;
		__syncthreads();
		for (uint32_t w2 = 0; w2 < w; ++w2)
			inc += wsum[w2];
		const uint32_t nb = __syncthreads_count(inc <= WS_RB);
		if (tid < (int)nb)
			loff[tid + 1] = inc;
		if (tid == 0)
			loff[0] = 0;
		__syncthreads();
		if (nb == 0) {
			if (walker) {
				const double m = chain_mean<XF>(sp, row);
				const uint32_t n1 = min(sp.L, sp.n - j * sp.L);
				double v[CH_REG];
				chunk_load(xr + (size_t)j * sp.L, n1, v);
				chunk_store<XF>(v, n1, m, ys);
				__builtin_amdgcn_wave_barrier();
				s = run_chain(s, ys, n1);
			}
			++reruns;
			++j;
			__syncthreads();
			continue;
		}
		const uint32_t nr = loff[nb];
		// which chunk of the batch each record belongs to
		if (tid < (int)nb)
			for (uint32_t r = inc - cc; r < inc; ++r)
				lch[r] = (uint8_t)tid;
		__syncthreads();
		// each thread's records requested RPT / 2 at a time, then stored;
		// a slot past the batch's last record reads that record again, so
		// the loads are unconditional and stay in flight together
#pragma unroll
		for (int h = 0; h < 2; ++h) {
			constexpr int HR = RPT / 2;
			if ((uint32_t)(h * HR * SBLOCK) >= nr)
				break;
			double av[HR], lo[HR], hi[HR];
			uint32_t rg[HR];
#pragma unroll
			for (int i = 0; i < HR; ++i) {
				const uint32_t r = min(tid + (h * HR + i) * SBLOCK, nr - 1);
				const uint32_t t = lch[r];
				const size_t src = (cb + j + t) * cap + (r - loff[t]);
				av[i] = sp.wa[src];
				lo[i] = sp.wlh[2 * src];
				hi[i] = sp.wlh[2 * src + 1];
				rg[i] = sp.wrg[src];
			}
#pragma unroll
			for (int i = 0; i < HR; ++i) {
				const uint32_t r = tid + (h * HR + i) * SBLOCK;
				if (r < nr) {
					la[r] = av[i];
					llh[r] = make_double2(lo[i], hi[i]);
					lrg[r] = rg[i];
				}
			}
		}
		__syncthreads();
		if (walker)
			s = ws_walk(s, la, lsb, sink, 0, nr);
		__syncthreads();
		uint32_t from = 0;
		for (;;) {
			if (tid == 0)
				sh_fail = 0xFFFFFFFFu;
			__syncthreads();
			// group g: records from + g WS_WB .. (at most WS_WB), its
			// start lsb[g]
			const uint32_t ng = (nr - from + WS_WB - 1) / WS_WB;
			uint32_t myf = 0xFFFFFFFFu;
			double myv = 0.0;
			for (uint32_t g = tid; g < ng; g += SBLOCK) {
				// all WS_WB records without branches (a slot past the
				// last reads it again, masked); the first that fails and
				// its start are kept, the sums after it are not used
				double v = lsb[g];
				const uint32_t rb = from + g * WS_WB;
				const uint32_t re = min(rb + WS_WB, nr);
				uint32_t gf = 0xFFFFFFFFu;
				double gv = 0.0;
#pragma unroll
				for (int i = 0; i < WS_WB; ++i) {
					const uint32_t r = rb + i;
					const uint32_t rr = min(r, nr - 1);
					const double2 lh = llh[rr];
					const uint32_t rg = lrg[rr];
					const double add = la[rr];
					const int gcode = (int)((rg >> 11) & 0x7FFu);
					const bool bad =
					    (r < re) &
					    ((v < lh.x) | (v > lh.y) |
					     ((gcode != 0) & !start_on_grid(v, gcode - 1)));
					const bool first = bad & (gf == 0xFFFFFFFFu);
					gf = first ? r : gf;
					gv = first ? v : gv;
					v = v + add;
				}
				if (gf < myf) {
					myf = gf;
					myv = gv;
				}
			}
			if (myf != 0xFFFFFFFFu)
				atomicMin(&sh_fail, myf);
			__syncthreads();
			const uint32_t f = sh_fail;
			if (f >= nr)
				break;
			if (myf == f)
				sh_fs = myv;
			__syncthreads();
			if (walker) {
				// the segment's values one by one from its start
				const double m = chain_mean<XF>(sp, row);
				// its values: up to the next record of the chunk, or
				// the chunk's end
				const uint32_t t = lch[f];
				const uint32_t n1 = min(sp.L, sp.n - (j + t) * sp.L);
				const uint32_t beg = min(lrg[f] & 0x7FFu, n1);
				const uint32_t end =
				    f + 1 < nr && lch[f + 1] == t
				        ? min(max(lrg[f + 1] & 0x7FFu, beg), n1)
				        : n1;
				double v[CH_REG];
				chunk_load(xr + (size_t)(j + t) * sp.L + beg, end - beg, v);
				chunk_store<XF>(v, end - beg, m, ys);
				__builtin_amdgcn_wave_barrier();
				s = run_chain(sh_fs, ys, end - beg);
				s = ws_walk(s, la, lsb, sink, f + 1, nr);
			}
			++reruns;
			from = f + 1;
			__syncthreads();
		}
		j += nb;
	}
	if (tid == 0) {
		sp.res[row] = s;
		if (reruns && sp.miss)
			atomicAdd(sp.miss, reruns);
		// the previous iteration's end, off the longer walk's path: the
		// kernels after this one read its done flag
		if (sp.iend && blockIdx.x == 0)
			em_iter_end(sp.iend);
	}
}

template <int XF>
int ws_launch(const hemk_spec &sp, const hemk_ctl *ctl, hipStream_t st)
{
	k_ws_prog<XF><<<dim3(sp.P, sp.nchain), WS_T, 0, st>>>(sp, ctl);
	k_ws_walk<XF><<<sp.nchain, SBLOCK, 0, st>>>(sp, ctl);
	return (int)hipGetLastError();
}

} // namespace

extern "C" {

size_t hemk_ws_records(uint32_t L, uint32_t chains, uint32_t P)
{
	return (size_t)chains * P * (L / 2);
}

int hemk_ws_finish(const hemk_spec *sp, const hemk_ctl *ctl, void *s)
{
	hipStream_t st = (hipStream_t)s;
	if (!sp->wa || !sp->wlh || !sp->wrg || !sp->wcnt ||
	    sp->L > HEMK_SPEC_LMAX || sp->L % WS_T != 0)
		return (int)hipErrorInvalidValue;
	switch (sp->xf) {
	case HEMK_SPEC_IDENT:
		return ws_launch<HEMK_SPEC_IDENT>(*sp, ctl, st);
	case HEMK_SPEC_CENTRED:
		return ws_launch<HEMK_SPEC_CENTRED>(*sp, ctl, st);
	case HEMK_SPEC_SQUARE:
		return ws_launch<HEMK_SPEC_SQUARE>(*sp, ctl, st);
	}
	return (int)hipErrorInvalidValue;
}

} // extern "C"
