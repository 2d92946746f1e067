// walk_sum.hip -- left-to-right FP64 sums, bitwise, by chunk programs and an
// in-order walk.
//
// The reference adds its norm sums one value after another (Norm,
// src/hash_align.c:153-154 and 158-159).  Inside one binade [2^E, 2^(E+1))
// every double is an integer multiple S of the binade's ulp u = 2^(E-52),
// and while a running sum s = S u and its result stay in that binade,
//     fl(s + y) = (S + q) u,   q = round(y / u) to the nearest integer,
// except at a tie (y / u = k + 1/2), which goes to the even one of S + k and
// S + k + 1.  So a run of values whose partial sums stay in one binade adds
// a fixed integer D to S once the parity of S before its first tie is
// known: the run is cut at that tie, which is added by itself (the
// hardware's addition rounds it as the reference does), and the rest of
// the run starts from an even S (a tie always ends on an even integer).  A
// chunk of L values becomes a short program of records, each one
//   * a segment: such a run, summarised as D u and the interval of starts
//     [lo, hi] for which every partial sum stays in the binade with a
//     margin of one ulp (and, after a cut, the need for an even start), or
//   * a single value (one that crosses a binade, lies near a binade
//     boundary or zero, or is a run's first tie), added as it is.
// Which values form segments is decided from an approximate trajectory (the
// chunk's start guessed from the earlier chunks' plain sums); that choice
// only affects speed.  Exactness comes from the walk: one wave adds the
// records of every chunk in order from +0, and a segment whose actual start
// lies outside [lo, hi] (or is odd where it must be even) is added value by
// value instead.  Adding D u to a
// start in [lo, hi] is exact (the result is on the binade's grid) and equals
// the run's rounded additions, and a single value is the reference's own
// addition, so the walk's sum is the reference's, bit for bit, whatever the
// guesses were.
//
// On SQUAREM's r and v rows of 200,000 values a chunk of 1,024 values needs
// a handful of records where the running sum stays away from zero and
// powers of two, and its single-value records are where it crosses them:
// about 760 (r) and 1,370 (v) records per row against 200,000 chained adds.
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

namespace {

__device__ __forceinline__ uint32_t lo32(double v)
{
	return (uint32_t)__double_as_longlong(v);
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
// starts among them, then the value since its last start; otherwise the
// value over all of them); the exclusive scan gives every thread the value
// carried in from the threads before it.
struct seg_sum {
	uint32_t h;
	int hp;          // position of the last segment start (-1: none),
	                 // + WS_EVEN when that segment needs an even start
	long long acc;   // increments since the last segment start
};
#define WS_EVEN 0x40000000

// record count, and the parity transfer (parity after = d ? in ^ v : v)
struct cnt_par {
	uint32_t n, d, v;
};

// a segment-candidate run starts among the values (h), then whether a tie
// has been seen since the last start (or over all the values)
struct tie_seen {
	uint32_t h, t;
};

struct seg_rng {
	uint32_t h, pad;
	long long mn, mx;   // smallest / largest partial increment since start
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
	__shared__ seg_rng w_r[WS_NW];
	__shared__ uint32_t sm[WS_T], sm2[WS_T];
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
	// the chunk's values and the first 256 earlier chunk sums are
	// requested before the control word is read, so the loads overlap
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
	if ((uint32_t)t < j) {
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

	// guessed start g of the chunk (the earlier chunks' sums, any order)
	// and the largest magnitude M of the earlier chunk starts
	double g = 0.0, M = 0.0;
	{
		const double lm2 = (double)sp.L * m * m;
		for (uint32_t i0 = 0; i0 < j; i0 += WS_T) {
			const uint32_t i = i0 + t;
			if (i0 && i < j) {
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
			const double ex = block_excl(c, 0.0, fadd, w_g, tot);
			if (i < j)
				M = fmax(M, fabs(g + ex));
			g += tot;
			__syncthreads();
		}
		double mt;
		(void)block_excl(M, 0.0, fmx, w_m, mt);
		M = fmax(mt, fabs(g));
	}

	// approximate partial sums a_k after each value, and a bound on every
	// partial sum's magnitude in this chunk
	double a[WS_E];
	double prev;
	double tau;
	{
		double tl = 0.0, ta = 0.0, tq = 0.0;
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
		sum_abs tt;
		const sum_abs ex = block_excl(sum_abs{tl, ta, tq},
		                              sum_abs{0.0, 0.0, 0.0}, pad, w_t, tt);
		// the chunk's sum of squares, for the guesses of a centred pass
		if (XF == HEMK_SPEC_IDENT && sp.csq && t == 0)
			sp.csq[cb + j] = tt.q;
		prev = g + ex.s;
#pragma unroll
		for (int k = 0; k < WS_E; ++k)
			a[k] = g + (ex.s + a[k]);
		tau = (M + fabs(g) + tt.a) * WS_TAU;
	}

	// segment candidates: the partial sums before and after stay in one
	// binade (sign and exponent) with a margin; q = the value in units of
	// the binade's ulp, rounded to the nearest integer, a round-half tie
	// (tmask) kept at floor for now: its rounding goes to the even
	// neighbour and so depends on the running sum's last bit
	uint32_t cmask = 0, tmask = 0;
	uint32_t key[WS_E];
	long long q[WS_E];
#pragma unroll
	for (int k = 0; k < WS_E; ++k) {
		const double ak = a[k], pk = k ? a[k - 1] : prev;
		const uint32_t ka =
		    (uint32_t)((unsigned long long)__double_as_longlong(ak) >> 52);
		const uint32_t kp =
		    (uint32_t)((unsigned long long)__double_as_longlong(pk) >> 52);
		const uint32_t be = ka & 0x7FFu;
		bool s = ((vmask >> k) & 1u) && ka == kp && be >= 2 && be <= 2044;
		const double blo = __longlong_as_double((long long)be << 52);
		const double bhi = __longlong_as_double((long long)(be + 1) << 52);
		const double A = fabs(ak), B = fabs(pk);
		s = s && A - blo > tau && bhi - A > tau && B - blo > tau &&
		    bhi - B > tau;
		const double xs = ldexp(y[k], 1075 - (int)be);
		const double kf = floor(xs);
		const double f = xs - kf;
		s = s && fabs(xs) < 0x1p52;
		key[k] = ka;
		q[k] = s ? (long long)kf + (f > 0.5 ? 1 : 0) : 0;
		cmask |= (uint32_t)s << k;
		tmask |= (uint32_t)(s && f == 0.5) << k;
	}
	const uint32_t fullm = (1u << E) - 1u;
	// the first tie of each candidate run is added by itself (its
	// rounding needs the true start's last bit); after it the running sum
	// is even, so the rest of the run is a segment whose start must be
	// even, its later ties settled by the parity below
	sm[t] = cmask;
	__syncthreads();
	uint32_t split;   // first ties of candidate runs
	{
		const uint32_t pc = t ? (sm[t - 1] >> (E - 1)) & 1u : 0u;
		const uint32_t ch = cmask & ~(((cmask << 1) | pc) & fullm);
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
	// segment starts (safe after a value that is not; an even start when
	// that value is a split tie), ends (safe before one that is not) and
	// records (values that are not safe, and starts)
	sm2[t] = smask | (split << 8);
	__syncthreads();
	const uint32_t pw = t ? sm2[t - 1] : 0u;
	const uint32_t psafe = (pw >> (E - 1)) & 1u;
	const uint32_t psplit = (pw >> (8 + E - 1)) & 1u;
	const uint32_t nsafe = t + 1 < WS_T ? sm2[t + 1] & 1u : 0u;
	const uint32_t hmask = smask & ~(((smask << 1) | psafe) & fullm);
	const uint32_t evmask = hmask & (((split << 1) | psplit) & fullm);
	const uint32_t emask = smask & ~((smask >> 1) | (nsafe << (E - 1)));
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
	seg_sum ls = {0u, -1, 0};
#pragma unroll
	for (int k = 0; k < WS_E; ++k) {
		if (!((smask >> k) & 1u))
			continue;
		if ((hmask >> k) & 1u) {
			ls.h = 1;
			ls.acc = 0;
			ls.hp = (int)(e0 + k) | (((evmask >> k) & 1u) ? WS_EVEN : 0);
		}
		ls.acc += q[k];
	}
	const auto ssop = [](seg_sum l, seg_sum r) {
		seg_sum o;
		o.h = l.h | r.h;
		o.acc = r.h ? r.acc : l.acc + r.acc;
		o.hp = r.h ? r.hp : l.hp;
		return o;
	};
	seg_sum stot;
	const seg_sum cs = block_excl(ls, seg_sum{0u, -1, 0}, ssop, w_s, stot);
	// smallest and largest P since each segment start
	seg_rng lr = {0u, 0u, LLONG_MAX, LLONG_MIN};
	{
		long long acc = cs.acc;
#pragma unroll
		for (int k = 0; k < WS_E; ++k) {
			if (!((smask >> k) & 1u))
				continue;
			if ((hmask >> k) & 1u) {
				lr.h = 1;
				lr.mn = LLONG_MAX;
				lr.mx = LLONG_MIN;
				acc = 0;
			}
			acc += q[k];
			lr.mn = min(lr.mn, acc);
			lr.mx = max(lr.mx, acc);
		}
	}
	const auto srop = [](seg_rng l, seg_rng r) {
		seg_rng o;
		o.h = l.h | r.h;
		o.pad = 0;
		o.mn = r.h ? r.mn : min(l.mn, r.mn);
		o.mx = r.h ? r.mx : max(l.mx, r.mx);
		return o;
	};
	seg_rng rtot;
	const seg_rng cr = block_excl(
	    lr, seg_rng{0u, 0u, LLONG_MAX, LLONG_MIN}, srop, w_r, rtot);

	// records: single values at their own index, a segment at its
	// start's index (no record lies between a segment's start and end).
	// A segment adds D = P(end) ulps; every partial sum S + P_i stays in
	// [2^52 + 1, 2^53 - 1] ulps (in magnitude) for starts S in [lo, hi]
	// (and, after a split tie, S even)
	seg_sum r = cs;
	long long mn = cr.mn, mx = cr.mx;
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
				mn = LLONG_MAX;
				mx = LLONG_MIN;
			}
			r.acc += q[k];
			mn = min(mn, r.acc);
			mx = max(mx, r.acc);
		}
		if (rmask & bit)
			++nr;
		if ((rmask & bit) && !(smask & bit)) {
			const size_t ri = rb0 + nr - 1;
			sp.wa[ri] = y[k];
			wlh[ri] = make_double2(-__builtin_inf(), __builtin_inf());
			sp.wrg[ri] = (e0 + k) | ((e0 + k + 1) << 11);
		}
		if (emask & bit) {
			const size_t ri = rb0 + nr - 1;
			const int be = (int)(key[k] & 0x7FFu);
			const bool neg = key[k] >> 11;
			const long long B52 = 1LL << 52, B53 = 1LL << 53;
			long long loS, hiS;
			if (!neg) {
				loS = max(B52, B52 + 1 - mn);
				hiS = min(B53 - 1, B53 - 1 - mx);
			} else {
				loS = max(-(B53 - 1), -(B53 - 1) - mn);
				hiS = min(-B52, -(B52 + 1) - mx);
			}
			double lo = __builtin_inf(), hi = -__builtin_inf();
			double da = 0.0;
			if (loS <= hiS) {
				const int sh = be - 1075;
				lo = ldexp((double)loS, sh);
				hi = ldexp((double)hiS, sh);
				da = ldexp((double)r.acc, sh);
			}
			sp.wa[ri] = da;
			wlh[ri] = make_double2(lo, hi);
			sp.wrg[ri] = (uint32_t)(r.hp & 0x7FF) |
			             ((e0 + k + 1) << 11) |
			             ((r.hp & WS_EVEN) ? HEMK_WS_EVEN : 0u);
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
		// each thread's records requested RPT / 2 at a time, then stored
#pragma unroll
		for (int h = 0; h < 2; ++h) {
			constexpr int HR = RPT / 2;
			double av[HR], lo[HR], hi[HR];
			uint32_t rg[HR];
#pragma unroll
			for (int i = 0; i < HR; ++i) {
				const uint32_t r = tid + (h * HR + i) * SBLOCK;
				av[i] = lo[i] = hi[i] = 0.0;
				rg[i] = 0;
				if (r < nr) {
					const uint32_t t = lch[r];
					const size_t src =
					    (cb + j + t) * cap + (r - loff[t]);
					av[i] = sp.wa[src];
					lo[i] = sp.wlh[2 * src];
					hi[i] = sp.wlh[2 * src + 1];
					rg[i] = sp.wrg[src] | (t << 22);   // chunk of the batch
				}
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
				double v = lsb[g];
				const uint32_t rb = from + g * WS_WB;
				const uint32_t re = min(rb + WS_WB, nr);
				for (uint32_t r = rb; r < re; ++r) {
					const double2 lh = llh[r];
					const bool odd = (lrg[r] & HEMK_WS_EVEN) &&
					                 (lo32(v) & 1u);
					if (v < lh.x || v > lh.y || odd) {
						if (r < myf) {
							myf = r;
							myv = v;
						}
						break;
					}
					v = v + la[r];
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
				const uint32_t rg = lrg[f];
				const uint32_t beg = rg & 0x7FFu;
				const uint32_t end = (rg >> 11) & 0x7FFu;
				const uint32_t t = (rg >> 22) & 0xFFu;
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
