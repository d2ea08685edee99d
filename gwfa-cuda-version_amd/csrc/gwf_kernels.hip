// gwf_kernels.hip — CDNA4 (gfx950) kernels for GWFA graph edit distance.
//
// A workgroup of NW 64-lane wavefronts aligns one read.  Per score step s
// (gwf_ed_extend, gwf-ed.cu:1064-1290) it:
//   P1 extends every wavefront state (gwf_extend1, gwf-ed.cu:806-839): each
//      lane compares 16 bytes with two 8-byte XORs; states still matching are
//      then extended cooperatively by their wave, 512 bytes per wave step
//      (8 bytes per lane; a 64-bit ballot and __ffsll pick the first
//      mismatching lane);
//   P2 emits the in-vertex successors (the WFA recurrence of
//      gwf_ed_extend_batch, gwf-ed.cu:891-1018) directly in sorted order: the
//      state on diagonal d owns output diagonals d-1 / d / d+1 unless a
//      neighbour owns them and takes the max over its neighbours' candidates;
//   P3 runs the vertex-crossing closure (the A queue, gwf-ed.cu:1106-1255)
//      level by level over a frontier, reading each vertex's CSR arcs;
//   P4 sorts the closure's successors (bitonic) and P6 merges them in place
//      into the in-vertex successors (merge path, right to left);
//   P5 folds this step's exits into the forbidden-interval history
//      (gwf_dedup, gwf-ed.cu:780-804);
//   P7 keeps the furthest state of every diagonal (gwf_diag_dedup,
//      gwf-ed.cu:653-687) outside the forbidden intervals (gwf_mixed_dedup,
//      gwf-ed.cu:691-747).
// The same step code runs on arrays in LDS (persistent kernel) or in HBM
// (plain kernel, and a persistent read whose wavefront outgrew its LDS).  A
// step that would overflow LDS leaves its input untouched and is re-run in
// HBM; an HBM overflow reports GWF_ST_OVERFLOW and the host re-runs the read
// with a larger workspace.  All arithmetic is exact integer.
#include "gwf_dev.h"

// This file is compiled once per key layout (GWF_KEY_LAYOUT = 0 .. GWF_N_LAYOUTS - 1,
// gwf_key_layout in gwf_dev.h); the layout's kernels live in namespace
// gwf_layout<n>.  Layout 0's build also holds the code that does not depend
// on the layout (binning, fills, sizes, occupancy) and the dispatchers that
// launch the layout a DGraph names.
#ifndef GWF_KEY_LAYOUT
#define GWF_KEY_LAYOUT 0
#endif
#define GWF_KNS_CAT2(a, b) a##b
#define GWF_KNS_CAT(a, b) GWF_KNS_CAT2(a, b)
#define GWF_KNS GWF_KNS_CAT(gwf_layout, GWF_KEY_LAYOUT)

#define WAVE 64
#define VIS_SLOTS 256      // per-step visited set (w, i+1), LDS
#define SORT_CHUNK 1024    // plain kernel: run length sorted in LDS
#define EMPTY_KEY (~0ull)

// Element types of LDS arrays.  Pointers to them are 32-bit LDS addresses,
// so the kernels' many array pointers do not take 64-bit scalar register
// pairs, and every access through them is a ds_ instruction.
#define LDS_AS __attribute__((address_space(3)))
typedef LDS_AS u64 lu64;
typedef LDS_AS i32 li32;

namespace GWF_KNS {

// ------------------------------------------------------------ key helpers
// A state (v, d, k) packed as v | d + D_BIAS | KINV_MAX - k in fields of
// KL.bv, KL.bd and KL.bk bits (see gwf_dev.h).
constexpr GwfKeyLayout KL = gwf_key_layout(GWF_KEY_LAYOUT);
constexpr i32 D_BIAS = 1 << (KL.bd - 1);
constexpr i32 KINV_MAX = (1 << KL.bk) - 2;
constexpr u32 D_MASK = (1u << KL.bd) - 1, K_MASK = (1u << KL.bk) - 1;
__device__ __forceinline__ u64 mk(u32 v, i32 d, i32 k) {
	return (u64)v << (KL.bd + KL.bk) | (u64)(u32)(d + D_BIAS) << KL.bk | (u64)(u32)(KINV_MAX - k);
}
__device__ __forceinline__ u32 key_v(u64 x) { return (u32)(x >> (KL.bd + KL.bk)); }
__device__ __forceinline__ i32 key_d(u64 x) { return (i32)((x >> KL.bk) & D_MASK) - D_BIAS; }
__device__ __forceinline__ i32 key_k(u64 x) { return KINV_MAX - (i32)(x & K_MASK); }
__device__ __forceinline__ u64 key_diag(u64 x) { return x >> KL.bk; }
__device__ __forceinline__ u64 diag_id(u32 v, i32 d) { return (u64)v << KL.bd | (u64)(u32)(d + D_BIAS); }
// The same state with offset k (d and v unchanged).
__device__ __forceinline__ u64 with_k(u64 x, i32 k) { return (x & ~(u64)K_MASK) | (u64)(u32)(KINV_MAX - k); }
// One diagonal in key units.  A state's d + D_BIAS is at least 1 (d >= -(vertex
// length) > -D_BIAS) and a successor on d + 1 stays inside the read (d + 1 <
// read length < D_BIAS), so key +- D_ONE never carries into or borrows from the
// vertex field.
constexpr u64 D_ONE = (u64)1 << KL.bk;

// ------------------------------------------------------------ wave helpers
__device__ __forceinline__ u32 lane_id() { return __lane_id(); }

__device__ __forceinline__ u32 rank_below(u64 mask) {
	return __builtin_amdgcn_mbcnt_hi((u32)(mask >> 32), __builtin_amdgcn_mbcnt_lo((u32)mask, 0u));
}

// Wave prefix operations on DPP lane moves (no LDS traffic): shifts by 1,
// 2, 4, 8 within each row of 16 lanes, then row 0's last lane into row 1,
// row 2's into row 3, and lane 31 into rows 2 and 3.  Lanes a move has no
// source for receive `old` (the operation's identity).
#define DPP_ROW_SHR(n) (0x110 + (n))
#define DPP_ROW_BCAST15 0x142
#define DPP_ROW_BCAST31 0x143
#define DPP_WAVE_SHR1 0x138
#define DPP_WAVE_SHL1 0x130

// (with full row masks a lane whose source is out of range reads 0, the old
// value every caller passes, and the destination needs no copy of it)
template <int CTRL, int ROWS>
__device__ __forceinline__ i32 dpp_i32(i32 old, i32 v) {
	return __builtin_amdgcn_update_dpp(old, v, CTRL, ROWS, 0xf, ROWS == 0xf);
}
template <int CTRL, int ROWS>
__device__ __forceinline__ u64 dpp_u64(u64 v) {
	const u32 lo = (u32)dpp_i32<CTRL, ROWS>(0, (i32)(u32)v);
	const u32 hi = (u32)dpp_i32<CTRL, ROWS>(0, (i32)(u32)(v >> 32));
	return (u64)hi << 32 | lo;
}

// exclusive prefix sum over the wave; *total receives the wave sum
__device__ __forceinline__ i32 wave_scan(i32 x, i32 *total) {
	i32 v = x;
	v += dpp_i32<DPP_ROW_SHR(1), 0xf>(0, v);
	v += dpp_i32<DPP_ROW_SHR(2), 0xf>(0, v);
	v += dpp_i32<DPP_ROW_SHR(4), 0xf>(0, v);
	v += dpp_i32<DPP_ROW_SHR(8), 0xf>(0, v);
	v += dpp_i32<DPP_ROW_BCAST15, 0xa>(0, v);
	v += dpp_i32<DPP_ROW_BCAST31, 0xc>(0, v);
	*total = __builtin_amdgcn_readlane(v, WAVE - 1);
	return v - x;
}

// the same in unsigned arithmetic (packed fields that may reach bit 31)
__device__ __forceinline__ u32 wave_scan_u32(u32 x, u32 *total) {
	u32 v = x;
	v += (u32)dpp_i32<DPP_ROW_SHR(1), 0xf>(0, (i32)v);
	v += (u32)dpp_i32<DPP_ROW_SHR(2), 0xf>(0, (i32)v);
	v += (u32)dpp_i32<DPP_ROW_SHR(4), 0xf>(0, (i32)v);
	v += (u32)dpp_i32<DPP_ROW_SHR(8), 0xf>(0, (i32)v);
	v += (u32)dpp_i32<DPP_ROW_BCAST15, 0xa>(0, (i32)v);
	v += (u32)dpp_i32<DPP_ROW_BCAST31, 0xc>(0, (i32)v);
	*total = (u32)__builtin_amdgcn_readlane((i32)v, WAVE - 1);
	return v - x;
}

__device__ __forceinline__ u64 umax64(u64 a, u64 b) { return a > b ? a : b; }

// inclusive prefix max over the wave (u64)
__device__ __forceinline__ u64 wave_scan_max(u64 x) {
	x = umax64(x, dpp_u64<DPP_ROW_SHR(1), 0xf>(x));
	x = umax64(x, dpp_u64<DPP_ROW_SHR(2), 0xf>(x));
	x = umax64(x, dpp_u64<DPP_ROW_SHR(4), 0xf>(x));
	x = umax64(x, dpp_u64<DPP_ROW_SHR(8), 0xf>(x));
	x = umax64(x, dpp_u64<DPP_ROW_BCAST15, 0xa>(x));
	x = umax64(x, dpp_u64<DPP_ROW_BCAST31, 0xc>(x));
	return x;
}

// lane l's value, l uniform over the wave
__device__ __forceinline__ i32 lane_bcast(i32 v, i32 l) { return __builtin_amdgcn_readlane(v, l); }
__device__ __forceinline__ u64 lane_bcast64(u64 v, i32 l) {
	return (u64)(u32)__builtin_amdgcn_readlane((i32)(u32)v, l) |
	       (u64)(u32)__builtin_amdgcn_readlane((i32)(u32)(v >> 32), l) << 32;
}

// ------------------------------------------------------------ block helpers
// Collectives over the NW waves of a workgroup.  Every thread must call them
// the same number of times (trip counts of loops around them are uniform).
// The per-wave partials alternate between two LDS slots, so one barrier per
// collective suffices: a slot is written again only two collectives later,
// after a barrier that every reader of its previous contents has passed.
template <int NW>
struct Blk {
	static constexpr int NT = WAVE * NW;
	li32 *red;   // LDS scratch, 2 NW ints
	lu64 *redu;  // LDS scratch, 2 NW u64
	li32 *xch;   // LDS, 4 ints: results handed from one wave to the others
	mutable i32 par = 0;  // slot of the next collective
	__device__ __forceinline__ u32 tid() const { return threadIdx.x; }
	__device__ __forceinline__ u32 wid() const { return NW == 1 ? 0u : threadIdx.x / WAVE; }
	__device__ __forceinline__ void sync() const { __syncthreads(); }
	// prefix over waves of a per-wave value (uniform within each wave)
	__device__ __forceinline__ i32 wave_prefix(i32 c, i32 *tot) const {
		if constexpr (NW == 1) {
			*tot = c;
			return 0;
		} else {
			li32 *r = red + par * NW;
			par ^= 1;
			if (lane_id() == 0) r[wid()] = c;
			sync();
			i32 pre = 0, t = 0;
#pragma unroll
			for (int w = 0; w < NW; ++w) {
				const i32 v = r[w];
				pre += w < (int)wid() ? v : 0;
				t += v;
			}
			*tot = t;
			return pre;
		}
	}
	__device__ __forceinline__ i32 scan(i32 x, i32 *tot) const {
		i32 wt, t;
		const i32 p = wave_scan(x, &wt);
		const i32 pre = wave_prefix(wt, &t);
		*tot = t;
		return p + pre;
	}
	// two exclusive scans in one (per-thread values below 256, block totals
	// below 2^16): *pc receives the second, *ta / *tc the totals
	__device__ __forceinline__ i32 scan2(i32 a, i32 c, i32 *pc, i32 *ta, i32 *tc) const {
		i32 t;
		const i32 p = scan(a | c << 16, &t);
		*pc = p >> 16;
		*ta = t & 0xffff;
		*tc = t >> 16;
		return p & 0xffff;
	}
	// three exclusive scans in one, in unsigned arithmetic (per-thread values
	// a <= 2, bb and c <= 6; block totals below 2^10, 2^11 and 2^11 for at
	// most 256 threads): *pb / *pc receive the second and third, *ta, *tb,
	// *tc the totals
	__device__ __forceinline__ i32 scan3(i32 a, i32 bb, i32 c, i32 *pb, i32 *pc, i32 *ta, i32 *tb, i32 *tc) const {
		static_assert(NW <= 4, "packed fields sized for at most 256 threads");
		u32 wt;
		u32 p = wave_scan_u32((u32)a | (u32)bb << 10 | (u32)c << 21, &wt);
		u32 t = wt;
		if constexpr (NW > 1) {
			li32 *r = red + par * NW;
			par ^= 1;
			if (lane_id() == 0) r[wid()] = (i32)wt;
			sync();
			u32 pre = 0;
			t = 0;
#pragma unroll
			for (int w = 0; w < NW; ++w) {
				const u32 v = (u32)r[w];
				pre += w < (int)wid() ? v : 0u;
				t += v;
			}
			p += pre;
		}
		*pb = (i32)((p >> 10) & 0x7ff);
		*pc = (i32)(p >> 21);
		*ta = (i32)(t & 0x3ff);
		*tb = (i32)((t >> 10) & 0x7ff);
		*tc = (i32)(t >> 21);
		return (i32)(p & 0x3ff);
	}
	// compaction slot of a predicate; *n advances by the block count
	__device__ __forceinline__ i32 slot(bool pred, i32 *n) const {
		const u64 m = __ballot(pred);
		i32 t;
		const i32 pre = wave_prefix(__popcll(m), &t);
		const i32 s = *n + pre + (i32)rank_below(m);
		*n += t;
		return s;
	}
	// slots of two predicates with one collective (*s2 receives the second)
	__device__ __forceinline__ i32 slot2(bool p1, bool p2, i32 *n1, i32 *n2, i32 *s2) const {
		const u64 m1 = __ballot(p1), m2 = __ballot(p2);
		i32 t;
		const i32 pre = wave_prefix(__popcll(m1) | __popcll(m2) << 16, &t);
		const i32 s1 = *n1 + (pre & 0xffff) + (i32)rank_below(m1);
		*s2 = *n2 + (pre >> 16) + (i32)rank_below(m2);
		*n1 += t & 0xffff;
		*n2 += t >> 16;
		return s1;
	}
	__device__ __forceinline__ bool any(bool p) const {
		if constexpr (NW == 1) return __ballot(p) != 0;
		else return __syncthreads_or(p) != 0;
	}
	// block max of non-negative values
	__device__ __forceinline__ i32 max(i32 x) const {
		x = ::max(x, dpp_i32<DPP_ROW_SHR(1), 0xf>(0, x));
		x = ::max(x, dpp_i32<DPP_ROW_SHR(2), 0xf>(0, x));
		x = ::max(x, dpp_i32<DPP_ROW_SHR(4), 0xf>(0, x));
		x = ::max(x, dpp_i32<DPP_ROW_SHR(8), 0xf>(0, x));
		x = ::max(x, dpp_i32<DPP_ROW_BCAST15, 0xa>(0, x));
		x = ::max(x, dpp_i32<DPP_ROW_BCAST31, 0xc>(0, x));
		x = lane_bcast(x, WAVE - 1);
		if constexpr (NW == 1) {
			return x;
		} else {
			li32 *r = red + par * NW;
			par ^= 1;
			if (lane_id() == 0) r[wid()] = x;
			sync();
			i32 m = r[0];
#pragma unroll
			for (int w = 1; w < NW; ++w) m = m > r[w] ? m : r[w];
			return m;
		}
	}
	// inclusive prefix max (u64) over the block; *excl = exclusive prefix max
	// (0 for the first thread), *tot = block max
	__device__ __forceinline__ u64 scan_max2(u64 x, u64 *excl, u64 *tot) const {
		const u64 wi = wave_scan_max(x);
		const u64 we = dpp_u64<DPP_WAVE_SHR1, 0xf>(wi);  // lane 0: 0
		if constexpr (NW == 1) {
			*excl = we;
			*tot = lane_bcast64(wi, WAVE - 1);
			return wi;
		} else {
			lu64 *r = redu + par * NW;
			par ^= 1;
			if (lane_id() == WAVE - 1) r[wid()] = wi;
			sync();
			u64 pre = 0, t = 0;
#pragma unroll
			for (int w = 0; w < NW; ++w) {
				const u64 v = r[w];
				if (w < (int)wid()) pre = pre > v ? pre : v;
				t = t > v ? t : v;
			}
			*excl = we > pre ? we : pre;
			*tot = t;
			return wi > pre ? wi : pre;
		}
	}
};

// Collectives over one wave, with the Blk interface (lane ids as thread ids,
// no LDS partials, no barriers): for work that one wave of a workgroup does
// while the others do something else.
struct WaveB {
	static constexpr int NT = WAVE;
	__device__ __forceinline__ u32 tid() const { return lane_id(); }
	__device__ __forceinline__ u32 wid() const { return 0u; }
	// A wave's memory accesses are performed in order (LDS, and global
	// memory through the CU's vector cache, as the workgroup barrier already
	// relies on for the waves of one CU); this only keeps the compiler from
	// moving memory accesses across the point
	__device__ __forceinline__ void sync() const {
		__builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
		__builtin_amdgcn_wave_barrier();
	}
	__device__ __forceinline__ i32 scan(i32 x, i32 *tot) const { return wave_scan(x, tot); }
	__device__ __forceinline__ i32 scan2(i32 a, i32 c, i32 *pc, i32 *ta, i32 *tc) const {
		i32 t;
		const i32 p = wave_scan(a | c << 16, &t);
		*pc = p >> 16;
		*ta = t & 0xffff;
		*tc = t >> 16;
		return p & 0xffff;
	}
	__device__ __forceinline__ i32 slot(bool pred, i32 *n) const {
		const u64 m = __ballot(pred);
		const i32 s = *n + (i32)rank_below(m);
		*n += __popcll(m);
		return s;
	}
	__device__ __forceinline__ i32 slot2(bool p1, bool p2, i32 *n1, i32 *n2, i32 *s2) const {
		const u64 m1 = __ballot(p1), m2 = __ballot(p2);
		const i32 s1 = *n1 + (i32)rank_below(m1);
		*s2 = *n2 + (i32)rank_below(m2);
		*n1 += __popcll(m1);
		*n2 += __popcll(m2);
		return s1;
	}
	__device__ __forceinline__ bool any(bool p) const { return __ballot(p) != 0; }
	__device__ __forceinline__ i32 max(i32 x) const {  // of non-negative values
		x = ::max(x, dpp_i32<DPP_ROW_SHR(1), 0xf>(0, x));
		x = ::max(x, dpp_i32<DPP_ROW_SHR(2), 0xf>(0, x));
		x = ::max(x, dpp_i32<DPP_ROW_SHR(4), 0xf>(0, x));
		x = ::max(x, dpp_i32<DPP_ROW_SHR(8), 0xf>(0, x));
		x = ::max(x, dpp_i32<DPP_ROW_BCAST15, 0xa>(0, x));
		x = ::max(x, dpp_i32<DPP_ROW_BCAST31, 0xc>(0, x));
		return lane_bcast(x, WAVE - 1);
	}
	__device__ __forceinline__ u64 scan_max2(u64 x, u64 *excl, u64 *tot) const {
		const u64 wi = wave_scan_max(x);
		*excl = dpp_u64<DPP_WAVE_SHR1, 0xf>(wi);  // lane 0: 0
		*tot = lane_bcast64(wi, WAVE - 1);
		return wi;
	}
};

// Unaligned 8-byte load of base[off..off+8) from a buffer padded past its
// end (base 8-byte aligned and uniform): one 16-byte load from the enclosing
// 8-byte boundary, addressed as base + 32-bit offset (no 64-bit address
// arithmetic per lane), then two byte-aligns of the dwords the window covers.
typedef u32 u32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ u64 load8u(const uint8_t *base, u32 off) {
	const u32x4 w = *(const u32x4 *)(base + (size_t)(off & ~7u));
	const bool hi = (off & 4u) != 0;
	const u32 b = off & 3u;
	const u32 w0 = hi ? w.y : w.x, w1 = hi ? w.z : w.y, w2 = hi ? w.w : w.z;
	return (u64)__builtin_amdgcn_alignbyte(w2, w1, b) << 32 | __builtin_amdgcn_alignbyte(w1, w0, b);
}

// first index in a[0..n) with a[i] >= x (lower) or a[i] > x (upper)
template <class PA>
__device__ __forceinline__ i32 lower_bound(PA a, i32 n, u64 x) {
	i32 lo = 0, hi = n;
	while (lo < hi) {
		const i32 m = (lo + hi) >> 1;
		if (a[m] < x) lo = m + 1;
		else hi = m;
	}
	return lo;
}
template <class PA>
__device__ __forceinline__ i32 upper_bound(PA a, i32 n, u64 x) {
	i32 lo = 0, hi = n;
	while (lo < hi) {
		const i32 m = (lo + hi) >> 1;
		if (a[m] <= x) lo = m + 1;
		else hi = m;
	}
	return lo;
}

// ------------------------------------------------------------ extension
// LCE of graph bytes ts[k+1..] and query bytes q[d+k+1..] up to max_k: the
// signed-max_k form of gwf_extend1 (gwf-ed.cu:806-839).  All lanes of the
// wave call it together; lanes with act == false pass through.
__device__ __forceinline__ i32 extend_wave(bool act, const uint8_t *seq, u32 seq_off, const uint8_t *q, i32 d, i32 k, i32 max_k,
                           int64_t *chars) {
	bool more = false;
	if (act) {
#pragma unroll
		for (int it = 0; it < 2; ++it) {
			const i32 avail = max_k - k;
			if (avail <= 0) break;
			const u64 x = load8u(seq, seq_off + (u32)(k + 1)) ^ load8u(q, (u32)(d + k + 1));
			const i32 m = x ? (i32)(__builtin_ctzll(x) >> 3) : 8;
			if (m >= avail) { *chars += avail; k = max_k; break; }
			if (m < 8) { *chars += m + 1; k += m; break; }
			*chars += 8;
			k += 8;
			if (it == 1) more = k < max_k;
		}
	}
	u64 lm = __ballot(more);
	const u32 lane = lane_id();
	while (lm) {  // cooperative extension of the long runs, one state at a time
		const int src = __ffsll((unsigned long long)lm) - 1;
		lm &= lm - 1;
		const i32 kk = lane_bcast(k, src), mk_ = lane_bcast(max_k, src), dd = lane_bcast(d, src);
		const u32 so = (u32)lane_bcast((i32)seq_off, src);
		i32 base = kk, res;
		for (;;) {
			const i32 avail = mk_ - base;
			const i32 my = (i32)lane * 8;
			i32 m = 8;
			if (my < avail) {
				const u64 x = load8u(seq, so + (u32)(base + 1 + my)) ^ load8u(q, (u32)(dd + base + 1 + my));
				m = x ? (i32)(__builtin_ctzll(x) >> 3) : 8;
			}
			const bool stop = (my < avail) && (m < 8 || my + 8 >= avail);
			const u64 sm = __ballot(stop);
			if (sm) {
				const int f = __ffsll((unsigned long long)sm) - 1;
				const i32 got = lane_bcast(my + m, f);
				res = base + (got < avail ? got : avail);
				break;
			}
			if (avail <= 0) { res = base; break; }
			base += WAVE * 8;
		}
		if ((int)lane == src) {
			*chars += res - kk + (res < mk_ ? 1 : 0);
			k = res;
		}
	}
	return k;
}

// ------------------------------------------------------------ sort / merge
// The value of lane (this lane ^ S): quad permutations for S = 1, 2,
// swizzles within 32 lanes for S = 4, 8, 16, a permute for S = 32.
template <int S>
__device__ __forceinline__ u32 xor_lane32(u32 v) {
	if constexpr (S == 1) return (u32)dpp_i32<0xB1, 0xf>(0, (i32)v);  // quad_perm [1,0,3,2]
	else if constexpr (S == 2) return (u32)dpp_i32<0x4E, 0xf>(0, (i32)v);  // quad_perm [2,3,0,1]
	else if constexpr (S < 32) return (u32)__builtin_amdgcn_ds_swizzle((i32)v, 0x1f | (S << 10));
	else return (u32)__shfl_xor((i32)v, 32);
}
template <int S>
__device__ __forceinline__ u64 xor_lane(u64 x) {
	return (u64)xor_lane32<S>((u32)(x >> 32)) << 32 | xor_lane32<S>((u32)x);
}

template <int SIZE, int STRIDE>
__device__ __forceinline__ u64 bitonic_stage(u64 x, u32 lane) {
	const u64 y = xor_lane<STRIDE>(x);
	// the lower lane of an ascending pair (or the upper of a descending one)
	// keeps the smaller key; the comparison and this choice combine as masks
	const bool take_min = ((lane & SIZE) == 0) == ((lane & STRIDE) == 0);
	return ((x < y) == take_min) ? x : y;
}
template <int SIZE, int STRIDE>
__device__ __forceinline__ u64 bitonic_merge_stages(u64 x, u32 lane) {
	x = bitonic_stage<SIZE, STRIDE>(x, lane);
	if constexpr (STRIDE > 1) x = bitonic_merge_stages<SIZE, STRIDE / 2>(x, lane);
	return x;
}

// Bitonic sort of 64 keys held one per lane (ascending by lane).
__device__ __forceinline__ u64 wave_sort64(u64 x) {
	const u32 lane = lane_id();
	x = bitonic_merge_stages<2, 1>(x, lane);
	x = bitonic_merge_stages<4, 2>(x, lane);
	x = bitonic_merge_stages<8, 4>(x, lane);
	x = bitonic_merge_stages<16, 8>(x, lane);
	x = bitonic_merge_stages<32, 16>(x, lane);
	x = bitonic_merge_stages<64, 32>(x, lane);
	return x;
}

// One round of pairwise merges of the sorted runs of width wdt in src[0..n)
// into dst (left run first on ties).  Thread t writes outputs [t E, t E + E)
// with E odd (above 2): a power-of-two stride between the lanes' reads and
// writes would put a wave's accesses on a few LDS banks.  A range that crosses
// the boundary of a pair is merged piece by piece: per piece one merge-path
// search, then a sequential merge from the two run heads.
template <int NW, class PS, class PD>
__device__ __forceinline__ void merge_round(PS src, PD dst, i32 n, i32 wdt, const Blk<NW> &b) {
	constexpr i32 NT = Blk<NW>::NT;
	i32 E = (n + NT - 1) / NT;
	if (E > 2) E |= 1;
	const i32 o0 = (i32)b.tid() * E;
	const i32 o1 = o0 + E < n ? o0 + E : n;
	for (i32 ob = o0; ob < o1;) {
		const i32 l = ob / (2 * wdt) * (2 * wdt);
		const i32 mid = l + wdt < n ? l + wdt : n, r = l + 2 * wdt < n ? l + 2 * wdt : n;
		const PS A = src + l, C = src + mid;
		const i32 na = mid - l, nc = r - mid, dg = ob - l;
		i32 lo = dg - nc > 0 ? dg - nc : 0, hi = dg < na ? dg : na;
		while (lo < hi) {
			const i32 m = (lo + hi) >> 1;
			if (A[m] <= C[dg - 1 - m]) lo = m + 1;
			else hi = m;
		}
		i32 i = lo, j = dg - lo;
		u64 ah = i < na ? A[i] : EMPTY_KEY, ch = j < nc ? C[j] : EMPTY_KEY;
		const i32 oe = o1 < r ? o1 : r;
		for (i32 o = ob; o < oe; ++o) {
			const bool ta = ah <= ch;  // an exhausted run's head is EMPTY_KEY, above every key
			dst[o] = ta ? ah : ch;
			if (ta) {
				++i;
				ah = i < na ? A[i] : EMPTY_KEY;
			} else {
				++j;
				ch = j < nc ? C[j] : EMPTY_KEY;
			}
		}
		ob = oe;
	}
}

// Sort src[0..n) into dst: every wave sorts runs of 64 in registers (no
// barriers), then each key's final rank is found by binary search in the
// other runs (one pass, when there are at most mw runs), or runs are merged
// pairwise through dst/src (log2(runs) passes).  mw is a tuning parameter.
template <int NW, class PT>
__device__ __forceinline__ void sort_runs(PT src, i32 n, PT dst, i32 mw, const Blk<NW> &b) {
	const i32 k = (n + WAVE - 1) / WAVE;
	for (i32 c = (i32)b.wid(); c < k; c += NW) {
		const i32 j = c * WAVE + (i32)lane_id();
		u64 x = j < n ? src[j] : EMPTY_KEY;
		x = wave_sort64(x);
		if (j < n) src[j] = x;
	}
	b.sync();
	if (k <= 1) {
		for (i32 j = (i32)b.tid(); j < n; j += Blk<NW>::NT) dst[j] = src[j];
		b.sync();
		return;
	}
	if (k <= mw) {
		for (i32 j = (i32)b.tid(); j < n; j += Blk<NW>::NT) {
			const u64 x = src[j];
			const i32 r = j / WAVE;
			i32 rank = j - r * WAVE;
			for (i32 o = 0; o < k; ++o) {
				if (o == r) continue;
				const i32 len = (o + 1) * WAVE <= n ? WAVE : n - o * WAVE;
				rank += o < r ? upper_bound(src + o * WAVE, len, x) : lower_bound(src + o * WAVE, len, x);
			}
			dst[rank] = x;
		}
		b.sync();
		return;
	}
	PT a = src, c = dst;
	for (i32 wdt = WAVE; wdt < n; wdt <<= 1) {
		merge_round(a, c, n, wdt, b);
		b.sync();
		PT t = a;
		a = c;
		c = t;
	}
	if (a != dst) {
		for (i32 j = (i32)b.tid(); j < n; j += Blk<NW>::NT) dst[j] = a[j];
		b.sync();
	}
}

// Sort p[0..n) (HBM): chunks of lc / 2 keys sorted in the LDS scratch (lc,
// a power of two >= 128, elements of it), then merge-path rounds between p
// and tmp (n elements).
template <int NW>
__device__ __forceinline__ void sort_hbm(u64 *p, i32 n, lu64 *lds, i32 lc, u64 *tmp, const Blk<NW> &b) {
	if (n <= 1) return;
	const i32 cw = lc / 2;
	for (i32 c0 = 0; c0 < n; c0 += cw) {
		const i32 m = n - c0 < cw ? n - c0 : cw;
		stage(lds, p + c0, m, b);
		b.sync();
		sort_runs(lds, m, lds + cw, 0, b);
		for (i32 t = (i32)b.tid(); t < m; t += Blk<NW>::NT) p[c0 + t] = lds[cw + t];
		b.sync();
	}
	u64 *src = p, *dst = tmp;
	for (i32 wdt = cw; wdt < n; wdt <<= 1) {
		merge_round(src, dst, n, wdt, b);
		b.sync();
		u64 *t = src;
		src = dst;
		dst = t;
	}
	if (src != p) {
		for (i32 t = (i32)b.tid(); t < n; t += Blk<NW>::NT) p[t] = src[t];
		b.sync();
	}
}

// Merge two interval lists sorted by lo (a before c on ties) into olo/ohi.
template <class B, class PA, class PC, class PO>
__device__ __forceinline__ void merge_intervals2(PA alo, PA ahi, i32 na, PC clo, PC chi, i32 nc, PO olo, PO ohi,
                                                 const B &b) {
	for (i32 j = (i32)b.tid(); j < na; j += B::NT) {
		const u64 x = alo[j];
		const i32 p = j + lower_bound(clo, nc, x);
		olo[p] = x, ohi[p] = ahi[j];
	}
	for (i32 t = (i32)b.tid(); t < nc; t += B::NT) {
		const u64 x = clo[t];
		const i32 p = t + upper_bound(alo, na, x);
		olo[p] = x, ohi[p] = chi[t];
	}
	b.sync();
}

// The same merge, partitioned into contiguous ranges of the output: one
// merge-path search per thread, then a sequential merge (for lists in HBM,
// where a binary search per element is a chain of dependent loads).
template <class B, class PA, class PC, class PO>
__device__ __forceinline__ void merge_intervals_ranges(PA alo, PA ahi, i32 na, PC clo, PC chi, i32 nc, PO olo, PO ohi,
                                                       const B &b) {
	constexpr i32 NT = B::NT;
	const i32 n = na + nc;
	const i32 E = (n + NT - 1) / NT;
	const i32 o0 = (i32)b.tid() * E;
	const i32 o1 = o0 + E < n ? o0 + E : n;
	if (o0 < n) {
		i32 lo = o0 - nc > 0 ? o0 - nc : 0, up = o0 < na ? o0 : na;
		while (lo < up) {
			const i32 mid = (lo + up) >> 1;
			if (alo[mid] <= clo[o0 - 1 - mid]) lo = mid + 1;
			else up = mid;
		}
		i32 i = lo, j = o0 - lo;
		u64 ah = i < na ? alo[i] : EMPTY_KEY, ch = j < nc ? clo[j] : EMPTY_KEY;
		for (i32 o = o0; o < o1; ++o) {
			if (i < na && (j >= nc || ah <= ch)) {
				olo[o] = ah, ohi[o] = ahi[i];
				++i;
				ah = i < na ? alo[i] : EMPTY_KEY;
			} else {
				olo[o] = ch, ohi[o] = chi[j];
				++j;
				ch = j < nc ? clo[j] : EMPTY_KEY;
			}
		}
	}
	b.sync();
}

// Runs of consecutive ids in the sorted, duplicate-free T[0..n) as half-open
// intervals [lo, hi); returns their count.  Each thread takes K consecutive
// ids per pass (their loads and its neighbours' in flight together), so a
// pass needs one block scan for K * NT ids.
template <i32 K, class B, class PT, class PO>
__device__ __forceinline__ i32 ti_runs_k(PT T, i32 n, PO lo, PO hi, const B &b) {
	i32 n_ti = 0;
	for (i32 base = 0; base < n; base += K * B::NT) {
		const i32 j0 = base + (i32)b.tid() * K;
		u64 x[K + 2];  // T[j0 - 1 .. j0 + K]
#pragma unroll
		for (i32 i = 0; i < K + 2; ++i) {
			const i32 j = j0 - 1 + i;
			x[i] = j >= 0 && j < n ? T[j] : 0;
		}
		u32 sm = 0, em = 0;  // run starts / ends among the thread's K ids
#pragma unroll
		for (i32 i = 0; i < K; ++i) {
			const i32 j = j0 + i;
			if (j < n) {
				if (j == 0 || x[i] + 1 != x[i + 1]) sm |= 1u << i;
				if (j + 1 == n || x[i + 2] != x[i + 1] + 1) em |= 1u << i;
			}
		}
		i32 cnt;
		// index of the run holding the thread's current id (before its first
		// start: the run begun by an earlier thread)
		i32 r = n_ti + b.scan(__popc(sm), &cnt) - 1;
#pragma unroll
		for (i32 i = 0; i < K; ++i) {
			if ((sm >> i) & 1) lo[++r] = x[i + 1];
			if ((em >> i) & 1) hi[r] = x[i + 1] + 1;
		}
		n_ti += cnt;
	}
	return n_ti;
}
// (one id per thread where one pass holds them all)
template <class B, class PT, class PO>
__device__ __forceinline__ i32 ti_runs(PT T, i32 n, PO lo, PO hi, const B &b) {
	return n <= B::NT ? ti_runs_k<1>(T, n, lo, hi, b) : ti_runs_k<4>(T, n, lo, hi, b);
}

// Coalesce overlapping or adjacent intervals of a list sorted by lo
// (gwf_intv_merge_adj, gwf-ed.cu:266-296).  Returns the output count (the
// caller guarantees room for n).
template <i32 K, class B, class PI, class PO>
__device__ __forceinline__ i32 coalesce_k(PI lo, PI hi, i32 n, PO olo, PO ohi, const B &b) {
	// (K consecutive intervals per thread per pass)
	u64 carry = 0;  // max hi over previous passes
	i32 ng = 0;     // groups started so far
	for (i32 base = 0; base < n; base += K * B::NT) {
		const i32 j0 = base + (i32)b.tid() * K;
		u64 l[K + 1], h[K];
#pragma unroll
		for (i32 i = 0; i < K; ++i) {
			const i32 j = j0 + i;
			l[i] = j < n ? lo[j] : EMPTY_KEY;
			h[i] = j < n ? hi[j] : 0;
		}
		l[K] = j0 + K < n ? lo[j0 + K] : EMPTY_KEY;
		u64 sm = h[0];
#pragma unroll
		for (i32 i = 1; i < K; ++i) sm = sm > h[i] ? sm : h[i];
		u64 ex, tmax;
		b.scan_max2(sm, &ex, &tmax);
		u64 run = ex > carry ? ex : carry;  // max hi before the thread's first interval
		u32 stm = 0, enm = 0;               // group starts / ends among its K
#pragma unroll
		for (i32 i = 0; i < K; ++i) {
			const i32 j = j0 + i;
			if (j < n && (j == 0 || l[i] > run)) stm |= 1u << i;
			run = run > h[i] ? run : h[i];
			h[i] = run;  // the group's hi so far
			if (j < n && (j + 1 == n || l[i + 1] > run)) enm |= 1u << i;
		}
		i32 cnt;
		i32 r = ng + b.scan(__popc(stm), &cnt) - 1;
#pragma unroll
		for (i32 i = 0; i < K; ++i) {
			if ((stm >> i) & 1) olo[++r] = l[i];
			if ((enm >> i) & 1) ohi[r] = h[i];
		}
		ng += cnt;
		carry = carry > tmax ? carry : tmax;
	}
	b.sync();
	return ng;
}
template <class B, class PI, class PO>
__device__ __forceinline__ i32 coalesce(PI lo, PI hi, i32 n, PO olo, PO ohi, const B &b) {
	return n <= B::NT ? coalesce_k<1>(lo, hi, n, olo, ohi, b) : coalesce_k<4>(lo, hi, n, olo, ohi, b);
}

template <class PH>
__device__ __forceinline__ bool forbidden(PH hlo, PH hhi, i32 nh, u64 dg) {
	const i32 p = upper_bound(hlo, nh, dg) - 1;
	return p >= 0 && dg < hhi[p];
}

// Drop the states of X[0..n) whose diagonal is forbidden (in place, order
// kept); returns the count left.  Such a state is never kept by the filter,
// and neither is any other state of its diagonal, so the result is the same.
// Each chunk is read whole before any of it is written (the block scan's
// barrier), and a kept state moves only down, below the next chunk.
template <int NW, class PX, class PH>
__device__ __forceinline__ i32 drop_forbidden(PX X, i32 n, PH hlo, PH hhi, i32 nh, const Blk<NW> &b) {
	constexpr i32 NT = Blk<NW>::NT;
	i32 nk = 0;
	for (i32 base = 0; base < n; base += NT) {
		const i32 j = base + (i32)b.tid();
		const bool act = j < n;
		const u64 x = act ? X[j] : 0;
		const bool keep = act && !forbidden(hlo, hhi, nh, key_diag(x));
		const i32 sl = b.slot(keep, &nk);
		if (keep) X[sl] = x;
	}
	b.sync();
	return nk;
}

// LDS-qualified view of a pointer into the arena: loads through it compile
// to LDS instructions instead of flat ones.
typedef const __attribute__((address_space(3))) u64 *lds_cptr;

// dst[0..n) = src[0..n), four loads in flight per thread (no barrier).
template <class B, class PD, class PS>
__device__ __forceinline__ void stage(PD dst, PS src, i32 n, const B &b) {
	constexpr i32 NT = B::NT;
	for (i32 j0 = (i32)b.tid(); j0 < n; j0 += 4 * NT) {
		u64 v[4];
#pragma unroll
		for (i32 r = 0; r < 4; ++r) {
			const i32 j = j0 + r * NT;
			v[r] = j < n ? src[j] : 0;
		}
#pragma unroll
		for (i32 r = 0; r < 4; ++r) {
			const i32 j = j0 + r * NT;
			if (j < n) dst[j] = v[r];
		}
	}
}

// One step of the fused merge below: takes the next state of Q / X (Q
// first on ties) into x and returns whether it is kept (first of its
// diagonal, outside the history intervals).  The cursors live in the caller;
// an exhausted list's head is EMPTY_KEY, above every key; h is the first
// history interval ending after the previous kept diagonal, (hl, hh) its
// bounds (EMPTY_KEY past the last one).
template <typename QP, typename XP, typename HP>
__device__ __forceinline__ bool mf_next(QP Q, i32 n_s, XP X, i32 n_x, HP Hl, HP Hh, i32 nh, i32 &i,
                                        i32 &jx, u64 &qh, u64 &xh, u64 &prev_dg, i32 &h, u64 &hl, u64 &hh, u64 &x) {
	const bool takeq = qh <= xh;
	x = takeq ? qh : xh;
	if (takeq) {
		++i;
		qh = i < n_s ? Q[i] : EMPTY_KEY;
	} else {
		++jx;
		xh = jx < n_x ? X[jx] : EMPTY_KEY;
	}
	const u64 dg = key_diag(x);
	bool keep = dg != prev_dg;
	prev_dg = dg;
	if (keep) {
		while (hh <= dg) {
			++h;
			hl = h < nh ? Hl[h] : EMPTY_KEY;
			hh = h < nh ? Hh[h] : EMPTY_KEY;
		}
		keep = hl > dg;  // coalesced intervals: disjoint, sorted
	}
	return keep;
}

// Merge of the sorted successor lists Q (n_s) and X (n_x) fused with the
// furthest-state-per-diagonal rule and the history test (P6+P7 for arrays
// in HBM).  Thread t takes the contiguous range [t E, t E + E) of the merged
// order: one merge-path search gives its start in Q and X; it then merges
// sequentially (Q first on ties), keeps the first (furthest) state of each
// diagonal and tests the history (coalesced intervals Hl/Hh, nh of them)
// with a pointer that only moves forward.  Kept states go to P: straight from
// registers, in chunks of 8 steps (see below).  Returns their count.
// QP, XP and HP are lds_cptr or const u64 * (where each list lives).  The
// kept states go to P[p_base..); prev0 is the diagonal of the state merged
// just before Q[0] and X[0] (~0: none).
template <int NW, typename QP, typename XP, typename HP>
__device__ __forceinline__ i32 merge_filter(QP Q, i32 n_s, XP X, i32 n_x, HP Hl, HP Hh, i32 nh, u64 *M,
                                            u64 *P, const Blk<NW> &b, i32 p_base = 0, u64 prev0 = ~0ull) {
	constexpr i32 NT = Blk<NW>::NT;
	constexpr i32 CH = 8;
	const i32 n_m = n_s + n_x;
	const i32 E = (n_m + NT - 1) / NT;
	const i32 o0 = (i32)b.tid() * E;
	const i32 o1 = o0 + E < n_m ? o0 + E : n_m;
	i32 i = 0, jx = 0, h = 0;
	u64 prev_dg = prev0, qh = EMPTY_KEY, xh = EMPTY_KEY, hl = EMPTY_KEY, hh = EMPTY_KEY;
	if (o0 < n_m) {
		i32 lo = o0 - n_x > 0 ? o0 - n_x : 0, up = o0 < n_s ? o0 : n_s;
		while (lo < up) {
			const i32 mid = (lo + up) >> 1;
			if (Q[mid] <= X[o0 - 1 - mid]) lo = mid + 1;
			else up = mid;
		}
		i = lo, jx = o0 - lo;
		if (o0 > 0) {  // diag of the merged state before o0 (none: o0 == 0)
			const bool from_q = i > 0 && (jx == 0 || Q[i - 1] > X[jx - 1]);
			prev_dg = key_diag(from_q ? Q[i - 1] : X[jx - 1]);
		}
		qh = i < n_s ? Q[i] : EMPTY_KEY, xh = jx < n_x ? X[jx] : EMPTY_KEY;
		// the first history interval ending after this range's first diagonal
		const u64 dg0 = key_diag(qh <= xh ? qh : xh);
		i32 c = nh;
		while (h < c) {
			const i32 mid = (h + c) >> 1;
			if (Hh[mid] <= dg0) h = mid + 1;
			else c = mid;
		}
		hl = h < nh ? Hl[h] : EMPTY_KEY;
		hh = h < nh ? Hh[h] : EMPTY_KEY;
	}
	// Chunks of CH merge steps in registers.  The last chunk's kept states
	// stay in registers until the scan has given their place in P; earlier
	// chunks (E > CH) are stored to M after their steps, packed at the front
	// of the thread's range (on this GPU a wait for a load also waits for the
	// stores issued before it, so a store between every two steps would add
	// a store round trip to each step).
	i32 n_new, in_m = 0;
	u64 kv[CH];
	u32 km = 0;
	for (i32 o = o0; o < o1; o += CH) {
		if (km) {  // the previous chunk was not the last one
#pragma unroll
			for (i32 u = 0; u < CH; ++u)
				if ((km >> u) & 1) M[o0 + in_m + __popc(km & ((1u << u) - 1))] = kv[u];
			in_m += __popc(km);
		}
		km = 0;
#pragma unroll
		for (i32 u = 0; u < CH; ++u) {
			u64 x = 0;
			if (o + u < o1 && mf_next(Q, n_s, X, n_x, Hl, Hh, nh, i, jx, qh, xh, prev_dg, h, hl, hh, x)) km |= 1u << u;
			kv[u] = x;
		}
	}
	const i32 at = p_base + b.scan(in_m + __popc(km), &n_new);
	for (i32 e0 = 0; e0 < in_m; e0 += 4) {
		u64 t[4];
#pragma unroll
		for (i32 r = 0; r < 4; ++r) t[r] = e0 + r < in_m ? M[o0 + e0 + r] : 0;
#pragma unroll
		for (i32 r = 0; r < 4; ++r)
			if (e0 + r < in_m) P[at + e0 + r] = t[r];
	}
#pragma unroll
	for (i32 u = 0; u < CH; ++u)
		if ((km >> u) & 1) P[at + in_m + __popc(km & ((1u << u) - 1))] = kv[u];
	b.sync();
	return n_new;
}

// ------------------------------------------------------------ visited set
__device__ __forceinline__ u32 vis_hash(u64 x) {
	x ^= x >> 31;
	x *= 0x7fb5d329728ea185ULL;
	x ^= x >> 27;
	return (u32)x;
}
// The claims of one arc group (key[u] for the bits u of `live`) in a per-step
// visited set (open addressing, linear probing over 16 positions): the first
// probes are issued together, the rare collisions probed on one key at a time.
// Returns the mask of the keys that were absent (and are now present; a full
// probe window reports "absent", which only costs duplicate work).
__device__ __forceinline__ u32 vis_insert4(lu64 *vis, u32 vmask, const u64 (&key)[4], u32 live) {
	u32 h[4];
	u64 old[4];
#pragma unroll
	for (int u = 0; u < 4; ++u) {
		h[u] = vis_hash(key[u]) & vmask;
		old[u] = EMPTY_KEY;
		if ((live >> u) & 1)
			__hip_atomic_compare_exchange_strong(&vis[h[u]], &old[u], key[u], __ATOMIC_RELAXED, __ATOMIC_RELAXED,
			                                     __HIP_MEMORY_SCOPE_WORKGROUP);
	}
	u32 absent = 0, pend = 0;
#pragma unroll
	for (int u = 0; u < 4; ++u)
		if ((live >> u) & 1) {
			if (old[u] == EMPTY_KEY) absent |= 1u << u;
			else if (old[u] != key[u]) pend |= 1u << u;
		}
	if (__ballot(pend != 0)) {
#pragma unroll
		for (int u = 0; u < 4; ++u)
			if ((pend >> u) & 1) {
				u32 hh = h[u];
				bool res = true;
#pragma unroll 1
				for (int p = 1; p < 16; ++p) {
					hh = (hh + 1) & vmask;
					u64 o = EMPTY_KEY;
					if (__hip_atomic_compare_exchange_strong(&vis[hh], &o, key[u], __ATOMIC_RELAXED, __ATOMIC_RELAXED,
					                                         __HIP_MEMORY_SCOPE_WORKGROUP))
						break;
					if (o == key[u]) {
						res = false;
						break;
					}
				}
				if (res) absent |= 1u << u;
			}
	}
	return absent;
}

// Slot of key K in a read's visited set (w.VK/VS/vmask), inserted when absent
// if `insert`; -1 when the probe window is exhausted (the set counts as full).
template <class BF>
__device__ __forceinline__ i32 vs_slot(const BF &w, u64 K, bool insert) {
	u32 h = vis_hash(K) & w.vmask;
	for (int p = 0; p < 32; ++p) {
		const u64 old = insert ? atomicCAS((unsigned long long *)&w.VK[h], (unsigned long long)EMPTY_KEY,
		                                   (unsigned long long)K)
		                       : __hip_atomic_load(&w.VK[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
		if (old == K) return (i32)h;
		if (old == EMPTY_KEY) {
			if (!insert) return -1;
			__hip_atomic_fetch_add(w.vfill, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
			return (i32)h;
		}
		h = (h + 1) & w.vmask;
	}
	return -1;
}

// ------------------------------------------------------------ working arrays
// One read's arrays: in LDS (T = lu64) or in HBM (T = u64).
template <class E>
struct BufsT {
	E *P, *Q, *X, *F0, *F1, *T;
	E *H0lo, *H0hi, *H1lo, *H1hi;  // forbidden-interval history, double buffered
	lu64 *sort_lds;   // HBM mode: LDS scratch for sorting runs and staging
	i32 sort_cap;     // its capacity for bitonic runs (power of two)
	i32 sort_n;       // its size in elements
	i32 mw;           // LDS sort: runs of 64 merged in one rank pass when at most mw of them
	u64 *sort_tmp;    // HBM mode: merge temp (max(x, t) elements)
	// per-read visited set (graphs with mixed in-overlaps; VK == nullptr:
	// none): keys, and per key the stamp of the candidate that owns it
	u64 *VK, *VS;
	u32 vmask;        // slots - 1
	li32 *vfill;      // LDS: slots taken
	// HBM mode: the first frc entries of the frontier lists F0 and F1 live in
	// the bottom of the LDS scratch (FL0, FL1; 2 frc elements, kept free until
	// the closure ends), so a closure level's small lists make no HBM round trip
	lu64 *FL0, *FL1;
	i32 frc;
	Caps c;
};
typedef BufsT<lu64> LBufs;
typedef BufsT<u64> HBufs;
template <class E> struct IsHbm { static constexpr bool v = false; };
template <> struct IsHbm<u64> { static constexpr bool v = true; };
#define FRC 64  // frontier entries per list kept in the LDS scratch in HBM mode (2 FRC in large scratches)

// Entry j of a frontier list (A: F0 or F1 in the workspace, L: its LDS part).
template <class E>
__device__ __forceinline__ u64 front_get(const BufsT<E> &w, const E *A, const lu64 *L, i32 j) {
	if constexpr (IsHbm<E>::v) return j < w.frc ? L[j] : A[j];
	else return A[j];
}
template <class E>
__device__ __forceinline__ void front_put(const BufsT<E> &w, E *A, lu64 *L, i32 j, u64 x) {
	if constexpr (IsHbm<E>::v) {
		if (j < w.frc) L[j] = x;
		else A[j] = x;
	} else {
		A[j] = x;
	}
}
// Give the frontier lists the bottom of an HBM workspace's LDS scratch: FRC
// entries each when it holds 8 times that, 2 FRC when it holds 32 times FRC.
__device__ __forceinline__ void front_in_scratch(BufsT<u64> &w) {
	const i32 f = w.sort_n >= 64 * FRC ? 2 * FRC : w.sort_n >= 16 * FRC ? FRC : 0;
	if (f) w.FL0 = w.sort_lds, w.FL1 = w.sort_lds + f, w.frc = f;
}

__host__ __device__ inline size_t ws_elems(const Caps &c) {
	return 2 * (size_t)c.m + c.x + 2 * (size_t)c.f + c.t + 4 * (size_t)c.h + (size_t)(c.x > c.t ? c.x : c.t) + 8 +
	       2 * (size_t)c.v;
}
// elements between equal plain-kernel workspaces (gwf_ws_bytes / 8)
__host__ __device__ inline u64 plain_stride(const Caps &c) { return (ws_elems(c) * 8 + 255) / 256 * 32; }
__host__ __device__ inline size_t lds_elems(const Caps &c) {
	return 2 * (size_t)c.m + c.x + 2 * (size_t)c.f + c.t + 4 * (size_t)c.h;
}

template <class E>
__device__ inline void carve_arrays(BufsT<E> &w, E *p, const Caps &c) {
	w.c = c;
	w.P = p; p += c.m;
	w.Q = p; p += c.m;
	w.X = p; p += c.x;
	w.F0 = p; p += c.f;
	w.F1 = p; p += c.f;
	w.T = p; p += c.t;
	w.H0lo = p; p += c.h;
	w.H0hi = p; p += c.h;
	w.H1lo = p; p += c.h;
	w.H1hi = p; p += c.h;
	w.sort_tmp = nullptr;
	w.VK = w.VS = nullptr;
	w.vmask = 0;
	w.vfill = nullptr;
	w.FL0 = w.FL1 = nullptr;
	w.frc = 0;
	w.sort_lds = nullptr;
	w.sort_n = 0;
	w.mw = 8;
	w.sort_cap = 0;
}
// LDS-resident arrays (no merge temp, no visited set of their own)
__device__ inline LBufs carve_lds(lu64 *p, const Caps &c) {
	LBufs w;
	carve_arrays(w, p, c);
	return w;
}
// An HBM workspace: the arrays, the merge temp and (mixed in-overlaps) the
// read's visited set
__device__ inline HBufs carve(u64 *p, const Caps &c) {
	HBufs w;
	carve_arrays(w, p, c);
	p += lds_elems(c);
	w.sort_tmp = p;
	p += (c.x > c.t ? c.x : c.t) + 8;
	w.VK = c.v ? p : nullptr;
	w.VS = p + c.v;
	w.vmask = c.v ? (u32)c.v - 1 : 0;
	return w;
}

// In-vertex successors of the state cur = (v, d, k) (the extend_batch
// recurrence, gwf-ed.cu:913-963): deletion -> (d-1, k+1), mismatch ->
// (d, k+1), insertion -> (d+1, k).  The state owns output diagonal d-1
// unless its left neighbour (same vertex) is on d-1 or d-2, d+1 unless its
// right neighbour is on d+1, and d always; an owner takes the max over the
// candidates of its neighbours too.  s* = kept in the vertex and the query,
// t* = leaves the vertex (an exit for the history).
struct Succ {
	bool s0, s1, s2, t0, t1, t2;
	i32 k0, k1, k2;
};
__device__ __forceinline__ Succ successors(bool act, u64 cur, u64 prv, u64 nxt, i32 vl, i32 ql) {
	// neighbours compared by diagonal id (vertex and diagonal in one number;
	// consecutive diagonals of a vertex have consecutive ids)
	const u64 cdg = key_diag(cur), pdg = key_diag(prv), ndg = key_diag(nxt);
	const i32 d = key_d(cur), k = key_k(cur);
	const bool hp = act && (pdg >> KL.bd) == (cdg >> KL.bd);  // EMPTY_KEY never matches a vertex
	const bool hn = act && (ndg >> KL.bd) == (cdg >> KL.bd);
	const i32 kp = key_k(prv), kn = key_k(nxt);
	Succ o;
	const bool p1 = hp && pdg + 1 == cdg, n1 = hn && ndg == cdg + 1;
	const bool e0 = act && !(hp && pdg + 2 >= cdg);
	o.k0 = k + 1;
	o.k1 = k + 1;
	if (p1 && kp > o.k1) o.k1 = kp;
	if (n1 && kn + 1 > o.k1) o.k1 = kn + 1;
	const bool e1 = act;
	const bool e2 = act && !n1;
	o.k2 = k;
	if (hn && ndg == cdg + 2 && kn + 1 > o.k2) o.k2 = kn + 1;
	o.s0 = e0 && (d - 1) + o.k0 < ql && o.k0 < vl, o.t0 = e0 && o.k0 == vl;
	o.s1 = e1 && d + o.k1 < ql && o.k1 < vl, o.t1 = e1 && o.k1 == vl;
	o.s2 = e2 && (d + 1) + o.k2 < ql && o.k2 < vl, o.t2 = e2 && o.k2 == vl;
	return o;
}

enum { STEP_NEXT = 0, STEP_FOUND = 1, STEP_EMPTY = 2, STEP_SPILL = 3 };

// P5 for a spilled read, in the LDS scratch: exit runs of T[0..n_tt),
// current history (n_h intervals), merged list; the new history goes to the
// top of the scratch (2 (n_h + n_tt) elements reserved there), where it
// stays for the filter, and to Hlo/Hhi in HBM for the next step; the bottom
// `base` elements are left alone.  Needs base + 2 n_tt + 2 n_h + 4 (n_h +
// n_tt) <= sort_n.  Returns its interval count.
template <class B>
__device__ __forceinline__ i32 hist_lds(const HBufs &w, const B &b, const u64 *Hc_lo, const u64 *Hc_hi, i32 n_h, i32 n_tt,
                                        u64 *Hlo, u64 *Hhi, i32 base) {
	const i32 cap_h = n_h + n_tt;
	lu64 *l = w.sort_lds + base;
	lu64 *hc = l + 2 * n_tt, *Clo = hc + 2 * n_h, *Chi = Clo + cap_h;
	lu64 *Olo = w.sort_lds + w.sort_n - 2 * cap_h, *Ohi = Olo + cap_h;
	stage(hc, Hc_lo, n_h, b);
	stage(hc + n_h, Hc_hi, n_h, b);
	const i32 n_ti = ti_runs(w.T, n_tt, l, l + n_tt, b);
	b.sync();
	merge_intervals2(hc, hc + n_h, n_h, l, l + n_tt, n_ti, Clo, Chi, b);
	const i32 nh2 = coalesce(Clo, Chi, n_h + n_ti, Olo, Ohi, b);
	for (i32 t = (i32)b.tid(); t < nh2; t += B::NT) Hlo[t] = Olo[t], Hhi[t] = Ohi[t];
	return nh2;
}

// The crossing part of one closure level on a graph whose vertices have
// in-arcs with different overlaps.  There a visited key (w, i+1) admits a
// different entry depending on which arc claims it first, so the reference's
// order is followed exactly (gwf-ed.cu:1106-1227: a FIFO queue, each popped
// state's arcs in CSR order, one visited set for the whole read): pass 1
// ranks every (frontier state, arc) candidate in that order and lets it claim
// its key in the read's set with the stamp (level, rank), the lowest stamp
// staying; pass 2 keeps the candidates that own their key and writes the
// next level's entries state by state, arc by arc.  Returns true when the
// set is full (the caller reports a spill; the read is re-run larger).
template <class B, class E>
__device__ __forceinline__ bool closure_mixed(const DGraph &g, const uint8_t *q, i32 ql, const BufsT<E> &w, const B &b,
                                              i32 v1, i32 n_f, bool closure_origin, u64 &lvl, i32 &n_na, i32 &n_x,
                                              i32 &n_t) {
	constexpr i32 NT = B::NT;
	const i32 tid = (i32)b.tid();
	const Caps &cap = w.c;
	const u64 lv = ++lvl;
	bool full = false;
	i32 rank_base = 0;
	for (i32 base = 0; base < n_f; base += NT) {
		const i32 j = base + tid;
		const bool act = j < n_f;
		const u64 key = act ? front_get(w, w.F0, w.FL0, j) : 0;
		const i32 i = key_d(key) + key_k(key);
		const uint4 vi = act ? g.vinfo[2 * (size_t)key_v(key)] : make_uint4(0, 0, 0, 0);
		const i32 deg = act && i + 1 != ql ? (i32)vi.w : 0;  // a query end claims no key
		i32 tot;
		const i32 r0 = rank_base + b.scan(deg, &tot);
		rank_base += tot;
		for (i32 e = 0; e < deg && !full; ++e) {
			const i32 sl = vs_slot(w, (u64)g.arc[vi.z + e].x << 32 | (u32)(i + 1), true);
			if (sl < 0) full = true;
			else atomicMin((unsigned long long *)&w.VS[sl], (unsigned long long)(lv << 24 | (u64)(r0 + e)));
		}
	}
	// every claim is in memory before any candidate reads its key's stamp
	__builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "agent");
	b.sync();
	if (b.any(full || rank_base >= (1 << 24) || *w.vfill > (i32)(w.vmask >> 1))) return true;
	rank_base = 0;
	for (i32 base = 0; base < n_f; base += NT) {
		const i32 j = base + tid;
		const bool act = j < n_f;
		const u64 key = act ? front_get(w, w.F0, w.FL0, j) : 0;
		const u32 v = key_v(key);
		const i32 d = key_d(key), k = key_k(key), i = d + k;
		const uint4 vi = act ? g.vinfo[2 * (size_t)v] : make_uint4(0, 0, 0, 0);
		const bool rec = act && closure_origin;
		// query end at a vertex end (only with v1 >= 0): delete into every next
		// vertex, nothing else (gwf-ed.cu:1244-1254)
		const bool qe = act && i + 1 == ql;
		const uint8_t c = act ? q[i + 1] : 0;
		const i32 degall = act ? (i32)vi.w : 0;
		i32 tot;
		const i32 r0 = rank_base + b.scan(qe ? 0 : degall, &tot);
		rank_base += tot;
		auto owns = [&](u32 wv, i32 e) {
			const i32 sl = vs_slot(w, (u64)wv << 32 | (u32)(i + 1), false);
			return sl >= 0 && __hip_atomic_load(&w.VS[sl], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ==
			                      (lv << 24 | (u64)(r0 + e));
		};
		i32 na = 0, nx = 0, n_ext = 0;
		for (i32 e = 0; e < degall; ++e) {
			const uint4 ar = g.arc[vi.z + e];
			if (qe) {
				nx += 1;
				continue;
			}
			const bool m = (uint8_t)(ar.y >> 24) == c;
			n_ext += m;
			const bool own = owns(ar.x, e);
			na += own && m;
			nx += own && !m ? 2 : 0;
		}
		const bool pi = rec && !qe && (degall == 0 || n_ext != degall);
		const bool t2 = rec && !qe && v1 >= 0;
		i32 ta, tx, tt;
		i32 sa = n_na + b.scan(na, &ta);
		i32 sx = n_x + b.scan(nx + (pi ? 1 : 0), &tx);
		const i32 st = n_t + b.scan(t2 ? 1 : 0, &tt);
		n_na += ta;
		n_x += tx;
		n_t += tt;
		for (i32 e = 0; e < degall; ++e) {
			const uint4 ar = g.arc[vi.z + e];
			const u32 wv = ar.x;
			const i32 ol = (i32)(ar.y & GWF_ARC_OL_MASK);
			if (qe) {
				if (sx < cap.x) w.X[sx] = mk(wv, i - ol, ol);
				++sx;
				continue;
			}
			const bool m = (uint8_t)(ar.y >> 24) == c;
			if (!owns(wv, e)) continue;
			if (m) {
				if (sa < cap.f) front_put(w, w.F1, w.FL1, sa, mk(wv, i + 1 - ol, ol));
				++sa;
			} else {
				if (sx + 1 < cap.x) {
					w.X[sx] = mk(wv, i - ol, ol);
					w.X[sx + 1] = mk(wv, i + 1 - ol, ol);
				}
				sx += 2;
			}
		}
		if (pi && sx < cap.x) w.X[sx] = mk(v, d + 1, k);
		if (t2 && st < cap.t) w.T[st] = diag_id(v, d);
	}
	return false;
}

// P3, the vertex-crossing closure of one step, by frontier levels: a
// frontier state at a vertex end follows each out-arc (one 16-byte record:
// target, overlap, first base); a match enters the target and is extended
// to its end (the next level) or to a mismatch, a mismatch gives the
// deletion and mismatch successors, and a state that did not match every
// arc an insertion.  A per-step visited set keeps each (vertex, query
// position) once.  Appends successors to X and (v1 >= 0) exits to T;
// returns STEP_NEXT, STEP_FOUND or STEP_SPILL.  B: the whole workgroup
// (Blk) or one wave (WaveB).
template <bool MIXED, class B, class E>
__device__ __forceinline__ int closure(const DGraph &g, const uint8_t *q, i32 ql, const BufsT<E> &w, const B &b, lu64 *vis,
                                       u32 vmask,
                                       i32 v1, i32 n_f, i32 &n_x, i32 &n_t, int64_t &chars, u64 &lvl) {
	constexpr i32 NT = B::NT;
	const i32 tid = (i32)b.tid();
	const Caps &cap = w.c;
	bool found = false;
	for (i32 t = tid; t <= (i32)vmask; t += NT) vis[t] = EMPTY_KEY;
	b.sync();
	bool closure_origin = false;
	while (n_f > 0) {
		i32 n_na = 0;
		if constexpr (MIXED) {  // mixed in-overlaps: the reference's order and one visited set per read
			if (closure_mixed(g, q, ql, w, b, v1, n_f, closure_origin, lvl, n_na, n_x, n_t)) return STEP_SPILL;
		}
		for (i32 base = 0; !MIXED && base < n_f; base += NT) {
			const i32 j = base + tid;
			const bool act = j < n_f;
			const u64 key = act ? front_get(w, w.F0, w.FL0, j) : 0;
			const u32 v = key_v(key);
			const i32 d = key_d(key), k = key_k(key);
			const i32 i = d + k;
			// the vertex record and, beside it, its first two out-arcs
			const uint4 vi = act ? g.vinfo[2 * (size_t)v] : make_uint4(0, 0, 0, 0);
			const uint4 va = act ? g.vinfo[2 * (size_t)v + 1] : make_uint4(0, 0, 0, 0);
			const bool rec = act && closure_origin;
			// query end at a vertex end (only with v1 >= 0): delete into every
			// next vertex, nothing else (gwf-ed.cu:1244-1254)
			const bool qe = act && i + 1 == ql;
			const uint8_t c = act ? q[i + 1] : 0;
			const i32 deg = act ? (i32)vi.w : 0;
			const i32 maxdeg = b.max(deg);
			i32 n_ext = 0;
			// arcs in groups of AG: the group's arc records are loaded together
			// (from a valid address when the lane has no arc there: loads under
			// a branch are waited for one after the other), its claims made, and
			// its successors placed with one scan.  The first two come with the
			// vertex record; arcs 2 and 3 are loaded only where a vertex has them.
			constexpr i32 AG = 4;
			for (i32 e0 = 0; e0 < maxdeg; e0 += AG) {
				uint2 ar[AG];
				if (e0 == 0) {
					ar[0] = 0 < deg ? make_uint2(va.x, va.y) : make_uint2(0, 0);
					ar[1] = 1 < deg ? make_uint2(va.z, va.w) : make_uint2(0, 0);
					ar[2] = ar[3] = make_uint2(0, 0);
					if (maxdeg > 2) {
#pragma unroll
						for (i32 u = 2; u < AG; ++u) {
							const bool has = u < deg;
							const uint2 t = *(const uint2 *)&g.arc[has ? vi.z + (u32)u : 0u];
							ar[u] = has ? t : make_uint2(0, 0);
						}
					}
				} else {
#pragma unroll
					for (i32 u = 0; u < AG; ++u) {
						const bool has = e0 + u < deg;
						const uint2 t = *(const uint2 *)&g.arc[has ? vi.z + (u32)(e0 + u) : 0u];
						ar[u] = has ? t : make_uint2(0, 0);
					}
				}
				// an arc that is its target's only in-arc claims nothing: the key
				// (target, i + 1) can repeat only where this state does, and the
				// copies it would make are removed with the other duplicates of a
				// diagonal by the merge
				u32 pa_m = 0, px_m = 0, pq_m = 0, live = 0, single = 0;
				i32 ca = 0, cx = 0;
				u64 vk[AG];
#pragma unroll
				for (i32 u = 0; u < AG; ++u) {
					const bool cl = e0 + u < deg && !qe;
					const bool one = (ar[u].y & GWF_ARC_SINGLE_IN) != 0;
					vk[u] = (u64)ar[u].x << 32 | (u32)(i + 1);
					live |= (cl && !one ? 1u : 0u) << u;
					single |= (cl && one ? 1u : 0u) << u;
				}
				static_assert(AG == 4, "one visited-set call per arc group");
				const u32 absent = vis_insert4(vis, vmask, vk, live) | single;
#pragma unroll
				for (i32 u = 0; u < AG; ++u) {
					const bool has = e0 + u < deg;
					const bool m = has && !qe && (uint8_t)(ar[u].y >> 24) == c;
					n_ext += m;
					const bool ab = (absent >> u) & 1;
					pa_m |= (ab && m ? 1u : 0u) << u;
					px_m |= (ab && !m ? 1u : 0u) << u;
					pq_m |= (has && qe ? 1u : 0u) << u;
				}
				ca = __popc(pa_m);
				cx = 2 * __popc(px_m) + __popc(pq_m);
				i32 sx, ta, tx;
				i32 sa = n_na + b.scan2(ca, cx, &sx, &ta, &tx);
				sx += n_x;
				n_na += ta;
				n_x += tx;
#pragma unroll
				for (i32 u = 0; u < AG; ++u) {
					const u32 wv = ar[u].x;
					const i32 ol = (i32)(ar[u].y & GWF_ARC_OL_MASK);
					if ((pa_m >> u) & 1) {
						if (sa < cap.f) front_put(w, w.F1, w.FL1, sa, mk(wv, i + 1 - ol, ol));
						++sa;
					}
					if ((px_m >> u) & 1) {
						if (sx + 1 < cap.x) {
							w.X[sx] = mk(wv, i - ol, ol);
							w.X[sx + 1] = mk(wv, i + 1 - ol, ol);
						}
						sx += 2;
					}
					if ((pq_m >> u) & 1) {
						if (sx < cap.x) w.X[sx] = mk(wv, i - ol, ol);
						++sx;
					}
				}
			}
			const bool pi = rec && !qe && (deg == 0 || n_ext != deg);
			// with v1 >= 0 the exits of states that entered their vertex in
			// this step join the history too (gwf-ed.cu:1177-1183); with
			// v1 < 0 they change no distance and are left out
			const bool t2 = rec && !qe && v1 >= 0;
			i32 si;
			const i32 st2 = b.slot2(t2, pi, &n_t, &n_x, &si);
			if (t2 && st2 < cap.t) w.T[st2] = diag_id(v, d);
			if (pi && si < cap.x) w.X[si] = mk(v, d + 1, k);
		}
		if (n_na > cap.f || n_x > cap.x) return STEP_SPILL;
		b.sync();
		// extend the states that entered a vertex with a match
		i32 n_f2 = 0;
		for (i32 base = 0; base < n_na; base += NT) {
			const i32 j = base + tid;
			const bool act = j < n_na;
			const u64 key = act ? front_get(w, w.F1, w.FL1, j) : 0;
			const u32 v = key_v(key);
			const i32 d = key_d(key);
			i32 k = key_k(key);
			const uint4 vi = act ? g.vinfo[2 * (size_t)v] : make_uint4(0, 0, 0, 0);
			const i32 vl = (i32)vi.y;
			const i32 max_k = (ql - d < vl ? ql - d : vl) - 1;
			k = extend_wave(act, g.seq, vi.x, q, d, k, max_k, &chars);
			const bool qend = act && d + k == ql - 1;
			const bool fnd = qend && (v1 < 0 || (v == (u32)v1 && k == vl - 1));
			found |= fnd;
			const bool vend = act && !fnd && k == vl - 1;
			const bool mid = act && !qend && !vend;
			const bool qmid = act && qend && !fnd && k + 1 < vl;  // query end mid-vertex: deletion only (gwf-ed.cu:1237-1241)
			i32 sx, tf, tx;
			const i32 sf = n_f2 + b.scan2(vend ? 1 : 0, mid ? 3 : qmid ? 1 : 0, &sx, &tf, &tx);
			sx += n_x;
			n_f2 += tf;
			n_x += tx;
			if (vend && sf < cap.f) front_put(w, w.F0, w.FL0, sf, with_k(key, k));
			if (mid && sx + 2 < cap.x) {
				w.X[sx] = mk(v, d - 1, k + 1);
				w.X[sx + 1] = mk(v, d, k + 1);
				w.X[sx + 2] = mk(v, d + 1, k);
			}
			if (qmid && sx < cap.x) w.X[sx] = mk(v, d - 1, k + 1);
		}
		if (b.any(found)) return STEP_FOUND;
		if (n_f2 > cap.f || n_x > cap.x) return STEP_SPILL;
		b.sync();
		n_f = n_f2;
		closure_origin = true;
	}

	return STEP_NEXT;
}

// The per-step visited set of a closure: the static one (vis, VIS_SLOTS
// slots), or, when the LDS scratch from element `base` on is free during the
// closure, the largest power of two of slots up to 16 VIS_SLOTS that fits
// there (a set too small for a large frontier lets duplicate entries
// through: every claim beyond its probe window counts as new).  Returns the
// slot mask.
__device__ __forceinline__ u32 step_vset(const HBufs &w, i32 base, lu64 *vis, lu64 *&vset) {
	u32 slots = VIS_SLOTS;
	while (slots < 16 * VIS_SLOTS && (i32)(2 * slots) <= w.sort_n - base) slots *= 2;
	vset = slots > VIS_SLOTS ? w.sort_lds + base : vis;
	return slots - 1;
}

// One score step.  Reads P[0..n_a) (the wavefront, sorted) and the history
// H[hs] (n_h intervals); on STEP_NEXT leaves the next wavefront in P and the
// updated history in H[hs ^ 1].  On STEP_SPILL nothing the next attempt needs
// was modified (P holds the same states, extended in place, which is
// idempotent).  HBM selects how sorts are staged.
template <bool HBM> struct ElemOf { typedef lu64 type; };
template <> struct ElemOf<true> { typedef u64 type; };

// P6 for arrays that stay where they are: merge the sorted X into Q in place
// (merge path, right to left; Q first on ties).
template <int NW, class PQ, class PX>
__device__ __forceinline__ void merge_in_place(PQ Q, i32 n_s, PX Xs, i32 n_x, const Blk<NW> &b) {
	constexpr i32 NT = Blk<NW>::NT;
	const i32 tid = (i32)b.tid();
	for (i32 hi = n_s + n_x; hi > 0; hi -= NT) {
		const i32 p = hi - NT + tid;
		u64 val = 0;
		if (p >= 0) {
			i32 lo = p - n_x > 0 ? p - n_x : 0, up = p < n_s ? p : n_s;
			while (lo < up) {
				const i32 mid = (lo + up) >> 1;
				if (Q[mid] <= Xs[p - 1 - mid]) lo = mid + 1;
				else up = mid;
			}
			const i32 i = lo, jx = p - lo;
			val = (i < n_s && (jx >= n_x || Q[i] <= Xs[jx])) ? Q[i] : Xs[jx];
		}
		b.sync();
		if (p >= 0) Q[p] = val;
		b.sync();
	}
}

template <int NW, bool HBM, bool MIXED>
__device__ __forceinline__ int gwf_step(const DGraph &g, const uint8_t *q, i32 ql,
                                        const BufsT<typename ElemOf<HBM>::type> &w, const Blk<NW> &b, lu64 *vis,
                                        i32 v1, i32 &n_a, i32 &n_h, i32 &hs, int64_t &chars, u64 &lvl) {
	typedef typename ElemOf<HBM>::type E;
	constexpr i32 NT = Blk<NW>::NT;
	const i32 tid = (i32)b.tid();
	const Caps &cap = w.c;

	// ---- P1+P2: extend every state in place, detect the end, collect vertex
	// ends (the closure's frontier), and emit the in-vertex successors into Q
	// (sorted) and their exits into T.  Each thread takes S consecutive states
	// per round and issues the loads of all of them before using any: one
	// chain of dependent loads (key, vertex record, first 8 bytes) per round.
	// States still matching after 8 bytes continue in extend_wave (16-byte
	// lane steps, then 512-byte cooperative wave steps).  Successors need the
	// extended neighbours on both sides, so they are emitted one round behind,
	// from registers: neighbours come from the adjacent lanes (DPP wave
	// shifts), across waves and rounds from per-round edge slots (three rounds
	// deep) kept in the visited-set buffer, which the closure clears later.
	constexpr i32 S = 2;
	const i32 RS = NT * S;
	const i32 R = (n_a + RS - 1) / RS;
	const i32 wv = (i32)b.wid();
	const u32 lane = lane_id();
	lu64 *edgeF = vis, *edgeL = vis + 3 * NW;  // [3][NW]: first / last state of each wave's range
	bool found = false, ovf = false;
	i32 n_f = 0, n_s = 0, n_t = 0;
	u64 pk[S];   // the previous round's extended states (EMPTY_KEY past the end)
	i32 pvl[S];  // and their vertex lengths
	// A round's vertex ends go to the frontier one round later, with its
	// successors: one scan places all three lists, and a plain barrier
	// publishes the edge slots before the successors
	bool pve[S];  // the previous round's vertex ends
#pragma unroll
	for (i32 s = 0; s < S; ++s) pk[s] = EMPTY_KEY, pvl[s] = 0, pve[s] = false;
	for (i32 r = 0; r <= R; ++r) {
		u64 nk[S];
		i32 nvl[S];
		bool vend[S];  // vertex ends of round r
#pragma unroll
		for (i32 s = 0; s < S; ++s) nk[s] = EMPTY_KEY, nvl[s] = 0, vend[s] = false;
		if (r < R) {  // extend round r
			const i32 j0 = r * RS + tid * S;
			u64 key[S];
			uint2 vi[S];
#pragma unroll
			for (i32 s = 0; s < S; ++s) key[s] = j0 + s < n_a ? w.P[j0 + s] : EMPTY_KEY;
			// loads are unconditional (from a valid address when the lane has no
			// state): loads under a branch are waited for one after the other
#pragma unroll
			for (i32 s = 0; s < S; ++s) {
				const uint2 t = *(const uint2 *)&g.vinfo[key[s] != EMPTY_KEY ? 2 * (size_t)key_v(key[s]) : 0];
				vi[s] = key[s] != EMPTY_KEY ? t : make_uint2(0, 0);
			}
			i32 kk[S], mk8[S];
			u64 x8[S];
#pragma unroll
			for (i32 s = 0; s < S; ++s) {
				const i32 d = key_d(key[s]), vl = (i32)vi[s].y;
				kk[s] = key_k(key[s]);
				mk8[s] = (ql - d < vl ? ql - d : vl) - 1;
				const bool go = key[s] != EMPTY_KEY && mk8[s] > kk[s];
				const u64 gx = load8u(g.seq, go ? vi[s].x + (u32)(kk[s] + 1) : 0u);
				const u64 qx = load8u(q, go ? (u32)(d + kk[s] + 1) : 0u);
				x8[s] = go ? gx ^ qx : 0;
			}
			bool more[S];
#pragma unroll
			for (i32 s = 0; s < S; ++s) {
				more[s] = false;
				const i32 avail = mk8[s] - kk[s];
				if (key[s] != EMPTY_KEY && avail > 0) {
					const i32 m = x8[s] ? (i32)(__builtin_ctzll(x8[s]) >> 3) : 8;
					if (m >= avail) { chars += avail; kk[s] = mk8[s]; }
					else if (m < 8) { chars += m + 1; kk[s] += m; }
					else { chars += 8; kk[s] += 8; more[s] = true; }
				}
			}
#pragma unroll
			for (i32 s = 0; s < S; ++s)  // every thread calls it S times (it is a wave collective)
				kk[s] = extend_wave(more[s], g.seq, vi[s].x, q, key_d(key[s]), kk[s], mk8[s], &chars);
#pragma unroll
			for (i32 s = 0; s < S; ++s) {
				const bool act = key[s] != EMPTY_KEY;
				const u32 v = key_v(key[s]);
				const i32 d = key_d(key[s]), k = kk[s], vl = (i32)vi[s].y;
				if (act) {
					nk[s] = with_k(key[s], k);
					nvl[s] = vl;
					// LDS arrays: keep the extension, so a spilled step restarts from it
					if constexpr (!HBM) w.P[j0 + s] = nk[s];
				}
				const bool qend = act && d + k == ql - 1;
				// the end (gwf-ed.cu:1231-1236): anywhere, or at the last base of v1
				const bool fnd = qend && (v1 < 0 || (v == (u32)v1 && k == vl - 1));
				found |= fnd;
				// vertex ends go to the closure; with v1 >= 0 that includes query
				// ends, which continue by deletion into the next vertices
				vend[s] = act && !fnd && k == vl - 1;
			}
			if (lane == 0) edgeF[(r % 3) * NW + wv] = nk[0];
			if (lane == WAVE - 1) edgeL[(r % 3) * NW + wv] = nk[S - 1];
		}
		if (r > 0) {  // successors and vertex ends of round r - 1
			b.sync();  // publishes the edge slots of round r
			u64 c[S + 2];  // c[0] = left neighbour, c[1..S] = pk, c[S + 1] = right neighbour
			c[0] = dpp_u64<DPP_WAVE_SHR1, 0xf>(pk[S - 1]);
			c[S + 1] = dpp_u64<DPP_WAVE_SHL1, 0xf>(pk[0]);
			if (lane == 0)
				c[0] = wv > 0 ? edgeL[((r - 1) % 3) * NW + wv - 1] : r >= 2 ? edgeL[((r - 2) % 3) * NW + NW - 1] : EMPTY_KEY;
			if (lane == WAVE - 1)
				c[S + 1] = wv < NW - 1 ? edgeF[((r - 1) % 3) * NW + wv + 1] : r < R ? edgeF[(r % 3) * NW] : EMPTY_KEY;
#pragma unroll
			for (i32 s = 0; s < S; ++s) c[s + 1] = pk[s];
			i32 cs = 0, ct = 0;
			Succ sc[S];
#pragma unroll
			for (i32 s = 0; s < S; ++s) {
				sc[s] = successors(pk[s] != EMPTY_KEY, c[s + 1], c[s], c[s + 2], pvl[s], ql);
				cs += (i32)sc[s].s0 + (i32)sc[s].s1 + (i32)sc[s].s2;
				ct += (i32)sc[s].t0 + (i32)sc[s].t1 + (i32)sc[s].t2;
			}
			i32 tot, pt, tot_t, p;
			{
				i32 nvp = 0;
#pragma unroll
				for (i32 s = 0; s < S; ++s) nvp += pve[s];
				i32 pq, tf;
				i32 pos = n_f + b.scan3(nvp, cs, ct, &pq, &pt, &tf, &tot, &tot_t);
#pragma unroll
				for (i32 s = 0; s < S; ++s)
					if (pve[s]) {
						if (pos < cap.f) front_put(w, w.F0, w.FL0, pos, pk[s]);
						++pos;
					}
				n_f += tf;
				p = n_s + pq;
			}
			if (n_s + tot > cap.m) ovf = true;
			else {
#pragma unroll
				for (i32 s = 0; s < S; ++s) {
					const u64 cur = c[s + 1];
					if (sc[s].s0) w.Q[p++] = with_k(cur - D_ONE, sc[s].k0);
					if (sc[s].s1) w.Q[p++] = with_k(cur, sc[s].k1);
					if (sc[s].s2) w.Q[p++] = with_k(cur + D_ONE, sc[s].k2);
				}
			}
			n_s += tot;
			p = n_t + pt;
			if (n_t + tot_t > cap.t) ovf = true;
			else {
#pragma unroll
				for (i32 s = 0; s < S; ++s) {
					const u64 cdg = key_diag(c[s + 1]);
					if (sc[s].t0) w.T[p++] = cdg - 1;
					if (sc[s].t1) w.T[p++] = cdg;
					if (sc[s].t2) w.T[p++] = cdg + 1;
				}
			}
			n_t += tot_t;
		}
#pragma unroll
		for (i32 s = 0; s < S; ++s) pk[s] = nk[s], pvl[s] = nvl[s], pve[s] = vend[s];
	}
	if (b.any(found)) return STEP_FOUND;
	if (n_f > cap.f || ovf) return STEP_SPILL;  // (ovf is uniform: derived from block totals)
	const i32 n_t1 = n_t;  // exits of in-vertex successors (sorted, unique)

	// ---- P3: vertex-crossing closure.  For a spilled read with v1 < 0 (the
	// closure records no exits, so the history update needs only the exits
	// of the in-vertex successors) and two waves, P5 runs beside it: wave 0
	// follows the closure, wave 1 updates the history in the LDS scratch
	// (with four waves both stay block-wide: measured faster).  A spill
	// found by the closure then ends the read's attempt (HBM capacities are
	// re-sized by the host), so P5 having run changes nothing.
	E *Hcur_lo = hs ? w.H1lo : w.H0lo, *Hcur_hi = hs ? w.H1hi : w.H0hi;
	E *Hlo = hs ? w.H0lo : w.H1lo, *Hhi = hs ? w.H0hi : w.H1hi;
	i32 n_x = 0, nh2 = 0;
	i32 hist_res = 0;  // HBM mode: LDS elements reserved at the top of the scratch
	// HBM mode: the new history's copy in the LDS scratch (null: only in HBM)
	const lu64 *Hl_l = nullptr, *Hh_l = nullptr;
	i32 q_res = 0;                       // HBM mode: Q staged in the LDS scratch just below the history
	bool p5_done = false;
	if constexpr (HBM && NW == 2) {
		const i32 cap_h = n_h + n_t;
		// (the bottom 2 frc elements of the scratch hold the closure's frontier lists)
		const i32 fr = 2 * w.frc;
		const i32 hneed = 2 * n_t + 2 * n_h + 4 * cap_h;
		if (v1 < 0 && fr + hneed <= w.sort_n && cap_h <= cap.h) {
			// a large frontier takes a visited set above the frontier lists
			// when the room left beside the history update holds 512 slots
			i32 vs = 0;
			lu64 *vset = vis;
			u32 vmask = VIS_SLOTS - 1;
			if (n_f > 64 && w.sort_n - fr - hneed >= 2 * VIS_SLOTS) {
				vs = 2 * VIS_SLOTS;
				while (vs < 16 * VIS_SLOTS && 2 * vs <= w.sort_n - fr - hneed) vs *= 2;
				vset = w.sort_lds + fr, vmask = (u32)vs - 1;
			}
			const WaveB wb;
			if (b.wid() == 0) {
				i32 nx = 0, nt = n_t;
				const int rc = closure<MIXED>(g, q, ql, w, wb, vset, vmask, v1, n_f, nx, nt, chars, lvl);
				if (lane_id() == 0) b.xch[0] = rc, b.xch[1] = nx;
			} else if (b.wid() == 1) {
				const i32 h2 = hist_lds(w, wb, Hcur_lo, Hcur_hi, n_h, n_t, Hlo, Hhi, fr + vs);
				// and Q, for the merge, right below the new history
				if (fr + vs + n_s <= w.sort_n - 2 * cap_h) stage(w.sort_lds + w.sort_n - 2 * cap_h - n_s, w.Q, n_s, wb);
				if (lane_id() == 0) b.xch[2] = h2;
			}
			b.sync();
			const int rc = b.xch[0];
			n_x = b.xch[1], nh2 = b.xch[2];
			if (rc != STEP_NEXT) return rc;
			hist_res = 2 * cap_h;
			q_res = fr + vs + n_s <= w.sort_n - 2 * cap_h ? n_s : 0;
			Hl_l = w.sort_lds + w.sort_n - 2 * cap_h, Hh_l = Hl_l + cap_h;
			p5_done = true;
		} else if (v1 < 0 && 4 * n_t + 2 * n_h <= cap.m && cap_h <= cap.h) {
			// the same split with the history update in HBM (P, free during the
			// closure, holds its intermediate lists; the LDS scratch above the
			// frontier lists holds the closure's visited set)
			const WaveB wb;
			if (b.wid() == 0) {
				i32 nx = 0, nt = n_t;
				lu64 *vset;
				const u32 vmask = step_vset(w, fr, vis, vset);
				const int rc = closure<MIXED>(g, q, ql, w, wb, vset, vmask, v1, n_f, nx, nt, chars, lvl);
				if (lane_id() == 0) b.xch[0] = rc, b.xch[1] = nx;
			} else if (b.wid() == 1) {
				E *TIlo = w.P, *TIhi = w.P + n_t;
				const i32 n_ti = ti_runs(w.T, n_t, TIlo, TIhi, wb);
				wb.sync();
				E *Clo = w.P + 2 * n_t, *Chi = Clo + (n_h + n_ti);
				merge_intervals_ranges(Hcur_lo, Hcur_hi, n_h, TIlo, TIhi, n_ti, Clo, Chi, wb);
				const i32 h2 = coalesce(Clo, Chi, n_h + n_ti, Hlo, Hhi, wb);
				if (lane_id() == 0) b.xch[2] = h2;
			}
			b.sync();
			const int rc = b.xch[0];
			n_x = b.xch[1], nh2 = b.xch[2];
			if (rc != STEP_NEXT) return rc;
			p5_done = true;
		}
	}
	if (!p5_done) {
		// (the LDS scratch is free during a block-wide closure)
		lu64 *vset = vis;
		u32 vmask = VIS_SLOTS - 1;
		if constexpr (HBM) vmask = step_vset(w, 2 * w.frc, vis, vset);
		const int rc = closure<MIXED>(g, q, ql, w, b, vset, vmask, v1, n_f, n_x, n_t, chars, lvl);
		if (rc != STEP_NEXT) return rc;
	}

	// ---- capacity checks for the destructive part of the step (P is free from here on)
	if (n_s + n_x > cap.m || n_x > cap.m || n_t > cap.t || n_t > cap.m || 4 * n_t + 2 * n_h > cap.m ||
	    n_h + n_t > cap.h)
		return STEP_SPILL;

	// ---- exits recorded by the closure (v1 >= 0 only): sort T, drop repeats
	i32 n_tt = n_t1;
	if (n_t > n_t1) {
		const E *S;
		if constexpr (HBM) {
			sort_hbm(w.T, n_t, w.sort_lds, w.sort_cap, w.sort_tmp, b);
			S = w.T;
		} else {
			sort_runs(w.T, n_t, w.P, w.mw, b);
			S = w.P;
		}
		E *D = HBM ? w.P : w.T;
		i32 nu = 0;
		for (i32 base = 0; base < n_t; base += NT) {
			const i32 j = base + tid;
			const bool act = j < n_t;
			const u64 x = act ? S[j] : 0;
			const bool keep = act && (j == 0 || S[j - 1] != x);
			const i32 sl = b.slot(keep, &nu);
			if (keep) D[sl] = x;
		}
		b.sync();
		if constexpr (HBM) {
			for (i32 j = tid; j < nu; j += NT) w.T[j] = D[j];
			b.sync();
		}
		n_tt = nu;
	}

	// ---- P5: history H[hs ^ 1] = coalesce(H[hs] u runs of this step's exits)
	const i32 cap_h = n_h + n_tt;  // bound on the merged interval count
	if constexpr (HBM) {
		if (!p5_done && 2 * n_tt + 2 * n_h + 4 * cap_h <= w.sort_n) {
			nh2 = hist_lds(w, b, Hcur_lo, Hcur_hi, n_h, n_tt, Hlo, Hhi, 0);
			hist_res = 2 * cap_h;
			Hl_l = w.sort_lds + w.sort_n - 2 * cap_h, Hh_l = Hl_l + cap_h;
			p5_done = true;
		}
	}
	if (!p5_done) {
		E *TIlo = w.P, *TIhi = w.P + n_tt;
		const i32 n_ti = ti_runs(w.T, n_tt, TIlo, TIhi, b);
		b.sync();
		if constexpr (HBM) {
			// the exits form few runs: the merge and coalesce fit the LDS
			// scratch even where the exits did not (current history and runs
			// at the bottom, merged list above them, the new history at the top)
			const i32 cap_r = n_h + n_ti;
			if (2 * n_h + 2 * n_ti + 4 * cap_r <= w.sort_n) {
				lu64 *hc = w.sort_lds, *rr = hc + 2 * n_h, *Clo = rr + 2 * n_ti, *Chi = Clo + cap_r;
				lu64 *Olo = w.sort_lds + w.sort_n - 2 * cap_r, *Ohi = Olo + cap_r;
				stage(hc, Hcur_lo, n_h, b);
				stage(hc + n_h, Hcur_hi, n_h, b);
				stage(rr, TIlo, n_ti, b);
				stage(rr + n_ti, TIhi, n_ti, b);
				b.sync();
				merge_intervals2(hc, hc + n_h, n_h, rr, rr + n_ti, n_ti, Clo, Chi, b);
				nh2 = coalesce(Clo, Chi, cap_r, Olo, Ohi, b);
				for (i32 t = tid; t < nh2; t += NT) Hlo[t] = Olo[t], Hhi[t] = Ohi[t];
				hist_res = 2 * cap_r;
				Hl_l = Olo, Hh_l = Ohi;
				p5_done = true;
			}
		}
		if (!p5_done) {
			E *Clo = w.P + 2 * n_tt, *Chi = Clo + (n_h + n_ti);
			if constexpr (HBM) merge_intervals_ranges(Hcur_lo, Hcur_hi, n_h, TIlo, TIhi, n_ti, Clo, Chi, b);
			else merge_intervals2(Hcur_lo, Hcur_hi, n_h, TIlo, TIhi, n_ti, Clo, Chi, b);
			nh2 = coalesce(Clo, Chi, n_h + n_ti, Hlo, Hhi, b);
		}
	}

	// ---- P4: sort the closure successors: in LDS (Xl), or in HBM (Xg, HBM mode)
	const lu64 *Xl = nullptr;
	const u64 *Xg = nullptr;
	i32 lds_used = 0;    // HBM mode: elements of the LDS scratch holding the sorted X
	bool x_lds = !HBM;   // the sorted X is in LDS
	if constexpr (HBM) {
		// X too large for the free scratch: first drop its forbidden states
		if (2 * n_x > w.sort_n - hist_res - q_res) {
			if (Hl_l) n_x = drop_forbidden(w.X, n_x, (lds_cptr)Hl_l, (lds_cptr)Hh_l, nh2, b);
			else n_x = drop_forbidden(w.X, n_x, (const u64 *)Hlo, (const u64 *)Hhi, nh2, b);
		}
		// X sorted in LDS is worth more than Q or the history staged there:
		// give up their copies (the merge stages them again if room is left)
		// when that lets X fit
		if (2 * n_x > w.sort_n - hist_res - q_res && 2 * n_x <= w.sort_n - hist_res) q_res = 0;
		if (2 * n_x > w.sort_n - hist_res - q_res && 2 * n_x <= w.sort_n) {
			hist_res = 0, q_res = 0;
			Hl_l = Hh_l = nullptr;
		}
		const i32 free_n = w.sort_n - hist_res - q_res;  // LDS scratch below the staged history and Q
		i32 free_cap = 128;
		while (free_cap * 2 <= free_n) free_cap <<= 1;
		if (2 * n_x <= free_n) {  // stage X in LDS, sort it there and merge from there
			// (staged above, sorted into the bottom: the staging half is free
			// again for the history and Q)
			stage(w.sort_lds + n_x, w.X, n_x, b);
			b.sync();
			sort_runs(w.sort_lds + n_x, n_x, w.sort_lds, w.mw, b);
			Xl = w.sort_lds;
			lds_used = n_x;
			x_lds = true;
		} else if ((hist_res || q_res) && free_cap >= 256 && free_cap <= free_n) {
			// X sorted in HBM through the free part of the scratch, so that the
			// history and Q staged above it stay there for the merge (which
			// then takes most of its inputs from LDS)
			sort_hbm(w.X, n_x, w.sort_lds, free_cap, w.sort_tmp, b);
			Xg = w.X;
		} else {
			if (hist_res) {  // the sort takes the whole scratch: the history is staged again later
				b.sync();
				hist_res = 0, q_res = 0;
				Hl_l = Hh_l = nullptr;
			}
			sort_hbm(w.X, n_x, w.sort_lds, w.sort_cap, w.sort_tmp, b);
			Xg = w.X;
		}
	} else {
		sort_runs(w.X, n_x, w.P, w.mw, b);
		Xl = w.P;
	}

	const i32 n_m = n_s + n_x;
	i32 n_new = 0;
	bool fused = false;
	if constexpr (HBM) {
		i32 used = lds_used;
		const i32 avail = w.sort_n - hist_res - q_res;
		if (!Hl_l && n_m <= 2 * cap.f && used + 2 * nh2 <= avail) {
			lu64 *hl = w.sort_lds + used, *hh = hl + nh2;
			stage(hl, Hlo, nh2, b);
			stage(hh, Hhi, nh2, b);
			Hl_l = hl, Hh_l = hh;
			used += 2 * nh2;
		}
		if (n_m <= 2 * cap.f) {
			// ---- P6+P7 fused (arrays in HBM), from LDS copies where they fit:
			// the sorted X and the new history are there; Q follows them.
			fused = true;
			u64 *M = w.F0;  // F0 and F1 are adjacent: 2 cap.f >= n_m
			const bool q_lds = q_res || used + n_s <= avail;
			lu64 *Ql = q_res ? w.sort_lds + w.sort_n - hist_res - q_res : w.sort_lds + used;
			if (q_lds && !q_res) stage(Ql, w.Q, n_s, b);
			b.sync();
			if (Hl_l) {
				const lu64 *Hl = Hl_l, *Hh = Hh_l;
				if (q_lds && x_lds) n_new = merge_filter<NW>((lds_cptr)Ql, n_s, Xl, n_x, Hl, Hh, nh2, M, w.P, b);
				else if (q_lds) n_new = merge_filter<NW>((lds_cptr)Ql, n_s, Xg, n_x, Hl, Hh, nh2, M, w.P, b);
				else if (x_lds && avail - used >= 256) {
					// Q did not fit: staged window by window (Q[qa, qb) and the X
					// entries before Q[qb] in merged order), each merged from LDS
					const i32 QW = avail - used;
					lu64 *qw = w.sort_lds + used;
					i32 qa = 0, xa = 0;
					u64 prev = ~0ull;
					while (qa < n_s || xa < n_x) {
						const i32 qb = qa + QW < n_s ? qa + QW : n_s;
						const i32 xb = qb < n_s ? xa + lower_bound(Xl + xa, n_x - xa, w.Q[qb]) : n_x;
						stage(qw, w.Q + qa, qb - qa, b);
						b.sync();
						n_new += merge_filter<NW>((lds_cptr)qw, qb - qa, Xl + xa, xb - xa, Hl, Hh, nh2, M, w.P, b, n_new,
						                          prev);
						// the last state merged: the larger of the two lists' last ones
						const u64 lq = qb > qa ? w.Q[qb - 1] : 0, lx = xb > xa ? Xl[xb - 1] : 0;
						if (qb > qa || xb > xa) prev = key_diag(lq > lx ? lq : lx);
						qa = qb, xa = xb;
					}
				} else if (x_lds) n_new = merge_filter<NW>((const u64 *)w.Q, n_s, Xl, n_x, Hl, Hh, nh2, M, w.P, b);
				else n_new = merge_filter<NW>((const u64 *)w.Q, n_s, Xg, n_x, Hl, Hh, nh2, M, w.P, b);
			} else {  // the history did not fit in LDS: every list through generic pointers
				const u64 *Qg = q_lds ? (const u64 *)Ql : (const u64 *)w.Q;
				const u64 *Xa = x_lds ? (const u64 *)Xl : Xg;
				n_new = merge_filter<NW>(Qg, n_s, Xa, n_x, (const u64 *)Hlo, (const u64 *)Hhi, nh2, M, w.P, b);
			}
		}
	}
	if (!fused) {
		// ---- P6: merge the sorted X into Q in place
		if (n_x > 0) {
			if (x_lds) merge_in_place(w.Q, n_s, Xl, n_x, b);
			else merge_in_place(w.Q, n_s, Xg, n_x, b);
		}

		// ---- P7: furthest state per diagonal, outside the forbidden intervals
		for (i32 base = 0; base < n_m; base += NT) {
			const i32 j = base + tid;
			const bool act = j < n_m;
			const u64 x = act ? w.Q[j] : 0;
			const u64 dg = key_diag(x);
			const bool first = act && (j == 0 || key_diag(w.Q[j - 1]) != dg);
			const bool keep = first && !forbidden(Hlo, Hhi, nh2, dg);
			const i32 sl = b.slot(keep, &n_new);
			if (keep) w.P[sl] = x;
		}
		b.sync();
	}
	n_a = n_new;
	n_h = nh2;
	hs ^= 1;
	return n_a == 0 ? STEP_EMPTY : STEP_NEXT;
}

// Copy the step input (wavefront + history) from LDS to HBM buffers.
template <int NW>
__device__ __forceinline__ void spill_state(const LBufs &from, const HBufs &to, i32 n_a, i32 n_h, i32 hs, const Blk<NW> &b) {
	for (i32 t = (i32)b.tid(); t < n_a; t += Blk<NW>::NT) to.P[t] = from.P[t];
	const lu64 *flo = hs ? from.H1lo : from.H0lo, *fhi = hs ? from.H1hi : from.H0hi;
	u64 *tlo = hs ? to.H1lo : to.H0lo, *thi = hs ? to.H1hi : to.H0hi;
	for (i32 t = (i32)b.tid(); t < n_h; t += Blk<NW>::NT) {
		tlo[t] = flo[t];
		thi[t] = fhi[t];
	}
	b.sync();
}

struct ReadOut {
	i32 ed, status;
	int64_t steps, states, chars, spill_step;
};

// Heartbeat record (relaxed system-scope stores: the host only reads it).
__device__ __forceinline__ void beat_put(GwfBeat *bt, i32 rid, i32 s, i32 n_a, i32 hbm) {
	__hip_atomic_store(&bt->read, rid, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
	__hip_atomic_store(&bt->step, s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
	__hip_atomic_store(&bt->n_a, n_a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
	__hip_atomic_store(&bt->in_hbm, hbm, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
	__hip_atomic_store(&bt->beat, (u64)wall_clock64(), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Align one read: LDS-resident while it fits (USE_LDS), then in HBM.  beat
// (may be null) receives a progress record every GWF_BEAT_EVERY steps.
template <int NW, bool USE_LDS, bool MIXED>
__device__ __forceinline__ ReadOut align_read(const DGraph &g, const uint8_t *q, i32 ql, u32 v0, i32 v1, const LBufs &lds,
                              const HBufs &hbm, const Blk<NW> &b, lu64 *vis, GwfBeat *beat, i32 rid) {
	ReadOut o;
	o.ed = -1, o.status = GWF_ST_OK, o.steps = 0, o.states = 0, o.chars = 0, o.spill_step = -1;
	i32 n_a = 1, n_h = 0, hs = 0;
	int64_t chars = 0;
	bool in_lds = USE_LDS;
	u64 lvl = 0;  // closure levels run so far (the per-read visited set's stamps)
	if constexpr (MIXED) {  // mixed in-overlaps: a fresh visited set for the read
		for (u32 t = b.tid(); t <= hbm.vmask; t += Blk<NW>::NT) hbm.VK[t] = EMPTY_KEY, hbm.VS[t] = ~0ull;
		if (b.tid() == 0) *hbm.vfill = 0;
		__builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
	}
	if (b.tid() == 0) {
		if (in_lds) lds.P[0] = mk(v0, 0, -1);
		else hbm.P[0] = mk(v0, 0, -1);
	}
	b.sync();
	for (i32 s = 0;; ++s) {
		o.steps += 1;
		o.states += n_a;
		// the ED is at most ql (v1 < 0), or ql plus the bases deleted on the way to v1
		if (s > ql + 1 + (v1 >= 0 ? (int64_t)g.total_len : 0)) { o.status = GWF_ST_LIMIT; break; }
		if (beat && (s % GWF_BEAT_EVERY) == 0 && b.tid() == 0) beat_put(beat, rid, s, n_a, !in_lds);
		int rc = STEP_SPILL;
		if constexpr (USE_LDS) {
			if (in_lds) {
				const u64 lvl0 = lvl;
				rc = gwf_step<NW, false, MIXED>(g, q, ql, lds, b, vis, v1, n_a, n_h, hs, chars, lvl);
				if (rc == STEP_SPILL) {  // the step input is intact: move it to HBM and redo the step there
					// (with the same closure stamps: its claims repeat those of the attempt)
					lvl = lvl0;
					spill_state(lds, hbm, n_a, n_h, hs, b);
					in_lds = false;
					o.spill_step = s;
				}
			}
		}
		if (rc == STEP_SPILL) rc = gwf_step<NW, true, MIXED>(g, q, ql, hbm, b, vis, v1, n_a, n_h, hs, chars, lvl);
		if (rc == STEP_SPILL) { o.status = GWF_ST_OVERFLOW; break; }
		if (rc == STEP_FOUND) { o.ed = s; break; }
		if (rc == STEP_EMPTY) { o.ed = -1; break; }
	}
	o.chars = chars;  // per thread; summed in write_out
	return o;
}

template <int NW>
__device__ __forceinline__ void write_out(const ReadOut &o, i32 r, u64 t0, i32 *out_ed, i32 *out_status, int64_t *out_counts,
                          const Blk<NW> &b) {
	int64_t c = o.chars;
#pragma unroll
	for (int s = 32; s > 0; s >>= 1) c += __shfl_xor(c, s);
	if constexpr (NW > 1) {
		if (lane_id() == 0) b.redu[b.wid()] = (u64)c;
		b.sync();
		c = 0;
		for (int w = 0; w < NW; ++w) c += (int64_t)b.redu[w];
		b.sync();
	}
	if (b.tid() == 0) {
		out_ed[r] = o.ed;
		out_status[r] = o.status;
		if (out_counts) {
			int64_t *cc = out_counts + GWF_NCOUNT * (size_t)r;
			cc[0] = o.steps;
			cc[1] = o.states;
			cc[2] = c;
			cc[3] = (int64_t)t0;
			cc[4] = (int64_t)wall_clock64();
			cc[5] = o.spill_step;
		}
	}
}

// ------------------------------------------------------------ kernels
// Plain: one wavefront per read, grid = reads, every array in HBM.  Block j
// aligns read read_idx[j] (read_idx == nullptr: read base + j) in a workspace
// of its own size (ws_off[j], ws_a[j]), or, with ws_off == nullptr, in slot j
// of equal workspaces of capacity a_uniform.
template <bool MIXED>
__global__ __launch_bounds__(WAVE) void gwf_plain_kernel(DGraph g, DReads rd, const i32 *read_idx, i32 base, i32 n,
                                                         const u64 *ws_off, const i32 *ws_a, i32 a_uniform,
                                                         u64 *scratch, i32 *out_ed, i32 *out_status,
                                                         int64_t *out_counts) {
	__shared__ u64 lds_sort[SORT_CHUNK];
	__shared__ u64 vis[VIS_SLOTS];
	__shared__ i32 red[2];
	__shared__ u64 redu[2];
	__shared__ i32 xch[4];
	const i32 blk = blockIdx.x;
	if (blk >= n) return;
	const i32 r = read_idx ? read_idx[blk] : base + blk;
	Blk<1> b{(li32 *)red, (lu64 *)redu, (li32 *)xch};
	const i32 a = ws_off ? ws_a[blk] : a_uniform;
	Caps caps;
	caps.m = 2 * a;
	caps.x = caps.f = caps.t = caps.h = a;
	caps.v = gwf_vis_slots(a, MIXED);
	HBufs w = carve(scratch + (ws_off ? ws_off[blk] : (u64)blk * plain_stride(caps)), caps);
	__shared__ i32 vfill;
	w.vfill = (li32 *)&vfill;
	w.sort_lds = (lu64 *)lds_sort;
	w.sort_cap = SORT_CHUNK;
	w.sort_n = SORT_CHUNK;
	const u64 t0 = wall_clock64();
	const LBufs none{};  // (no LDS-resident arrays)
	const ReadOut o = align_read<1, false, MIXED>(g, rd.seq + rd.off[r], rd.ql[r], (u32)rd.v0[r], rd.v1, none, w, b,
	                                              (lu64 *)vis, nullptr, r);
	write_out<1>(o, r, t0, out_ed, out_status, out_counts, b);
}

__device__ __forceinline__ u64 uni64(u64 x) {
	return (u64)(u32)__builtin_amdgcn_readfirstlane((i32)(u32)x) |
	       (u64)(u32)__builtin_amdgcn_readfirstlane((i32)(u32)(x >> 32)) << 32;
}

// Largest power-of-two run that an LDS arena of le elements can sort once
// a read has spilled (the arena is free then).
__device__ __forceinline__ i32 arena_sort_cap(size_t le) {
	i32 sc = 2;
	while ((size_t)sc * 2 <= le && sc < (1 << 14)) sc <<= 1;
	return sc;
}

// Persistent: resident workgroups pull reads from a queue; each read is
// aligned by NW waves with its arrays in LDS until they outgrow it.
// The queue of bin `bin` is order[bin_off[bin] .. bin_off[bin + 1]) (device
// memory, so a device-side ordering pass can fill it); heads[bin] counts the
// reads taken from it.  order == nullptr: the identity order.
// The step is latency-bound: resident waves matter more than a few spilled
// registers.  Capping registers at 5 waves per SIMD (96 VGPRs, a few dozen
// bytes per lane spilled) measured 9-15% faster than the uncapped 4 waves.
template <int NW, bool MIXED>
__global__ __launch_bounds__(WAVE *NW) __attribute__((amdgpu_waves_per_eu(5))) void gwf_persistent_kernel(
    DGraph g, DReads rd, const i32 *order,
                                                                 const i32 *bin_off, i32 bin, i32 *heads, Caps lc,
                                                                 Caps hc, i32 mw, i32 arena_n, u64 *scratch, u64 ws_elems_per_wg,
                                                                 i32 *out_ed, i32 *out_status, int64_t *out_counts,
                                                                 GwfBeat *beats) {
	extern __shared__ u64 arena[];
	__shared__ u64 vis[VIS_SLOTS];
	__shared__ i32 red[2 * NW];
	__shared__ u64 redu[2 * NW];
	__shared__ i32 xch[4];
	__shared__ i32 next_q;
	Blk<NW> b{(li32 *)red, (lu64 *)redu, (li32 *)xch};
	LBufs lb = carve_lds((lu64 *)arena, lc);
	lb.mw = mw;
	HBufs hb = carve(scratch + (size_t)blockIdx.x * ws_elems_per_wg, hc);
	__shared__ i32 vfill;
	hb.vfill = (li32 *)&vfill;
	lb.VK = hb.VK, lb.VS = hb.VS, lb.vmask = hb.vmask, lb.vfill = hb.vfill;  // one visited set in both modes
	const i32 q_lo = bin_off[bin], q_n = bin_off[bin + 1] - q_lo;
	hb.sort_lds = (lu64 *)arena;  // the LDS arena is free once a read has spilled
	hb.sort_cap = arena_sort_cap((size_t)arena_n);
	hb.sort_n = arena_n;
	hb.mw = mw;
	front_in_scratch(hb);
	GwfBeat *beat = beats ? beats + blockIdx.x : nullptr;
	for (;;) {
		if (b.tid() == 0) next_q = atomicAdd(&heads[bin], 1);
		b.sync();
		const i32 qi = __builtin_amdgcn_readfirstlane(next_q);  // uniform: the read's pointers stay scalar
		b.sync();
		if (qi >= q_n) break;
		const i32 r = order ? order[q_lo + qi] : q_lo + qi;
		const u64 t0 = wall_clock64();
		const ReadOut o = align_read<NW, true, MIXED>(g, rd.seq + rd.off[r], rd.ql[r], (u32)rd.v0[r], rd.v1, lb, hb, b,
		                                              (lu64 *)vis, beat, r);
		write_out<NW>(o, r, t0, out_ed, out_status, out_counts, b);
		b.sync();
	}
	if (beat && b.tid() == 0) beat_put(beat, -1, 0, 0, 0);
}

// Device server: resident workgroups claim requests published by the host
// in a ring of host-coherent memory, copy the read into the request's
// device region, align it (LDS first, then the region's HBM workspace) and
// publish the result with a system-scope release, so the host sees every
// read as soon as it is done while the kernel keeps running.  Requests that
// arrive while others run are claimed by whichever workgroup frees up first.
//
// Host memory is read with relaxed system-scope atomic loads (coherent
// without cache invalidation), so idle workgroups polling the ring do not
// disturb the caches of busy ones; the request's device region may have been
// used by a workgroup on another XCD, hence the agent-scope acquire before
// using it and the system-scope release that ends every request.
__device__ __forceinline__ u64 host_ld(const u64 *p) {
	return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

template <int NW, bool MIXED>
__global__ __launch_bounds__(WAVE *NW) __attribute__((amdgpu_waves_per_eu(4))) void gwf_serve_kernel(SrvArgs a) {
	extern __shared__ u64 arena[];
	__shared__ u64 vis[VIS_SLOTS];
	__shared__ i32 red[2 * NW];
	__shared__ u64 redu[2 * NW];
	__shared__ i32 xch[4];
	__shared__ int64_t s_seq, s_claim;
	__shared__ i32 s_polls;
	__shared__ SrvEntry s_e;
	__shared__ i32 vfill;
	constexpr i32 NT = Blk<NW>::NT;
	static_assert(sizeof(SrvEntry) == 64, "SrvEntry is eight u64 words");
	Blk<NW> b{(li32 *)red, (lu64 *)redu, (li32 *)xch};
	LBufs lb = carve_lds((lu64 *)arena, a.lc);
	lb.mw = a.mw;
	const i32 sort_cap = arena_sort_cap(lds_elems(a.lc));
	GwfBeat *beat = a.beat ? a.beat + blockIdx.x : nullptr;
	i32 idle = 0;
	for (;;) {
		if (b.tid() == 0) {
			// Claim by ticket: a fetch-add on the claim counter gives this
			// workgroup the number of the next unclaimed request, with one
			// atomic that always succeeds (compare-exchange claims by many idle
			// workgroups mostly fail and queue up behind each other on the one
			// address).  The workgroup then waits until the host has published
			// its request, polling only the host's tail word: one sleep of
			// 127 x 64 cycles (~3 us) between polls within 4 requests of the
			// tail, else 1 to 32 growing with the time spent idle, 8 times as
			// many from 64 requests ahead (measured: ~2000 idle workgroups
			// polling host memory every ~100 us made a busy workgroup's read
			// 2-5x slower).  On the stop bit it ends holding its ticket (the
			// host puts the counter back to the tail).  After idle_polls empty
			// polls it hands the ticket back, which only the holder of the last
			// ticket issued can do, so tickets return newest first and, once
			// every workgroup has ended by itself, the counter is again the
			// first unclaimed request.
			int64_t got = -1;
			u64 t = ~0ull;  // ticket held
			for (;;) {
				const u64 w = host_ld(&a.ctl->word);
				const u64 tail = w & ~GWF_SRV_STOP;
				if (t == ~0ull) {
					if (w & GWF_SRV_STOP) break;
					t = __hip_atomic_fetch_add(a.claimed, (u64)1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
				}
				if (t < tail) {
					got = (int64_t)t;
					break;
				}
				if (w & GWF_SRV_STOP) break;
				if (++idle > a.idle_polls) {
					u64 last = t + 1;
					if (__hip_atomic_compare_exchange_strong(a.claimed, &last, t, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
					                                         __HIP_MEMORY_SCOPE_AGENT))
						break;
				}
				const u64 ahead = t - tail;
				const i32 naps = ahead < 4 ? 1 : (idle < 40 ? 1 << (idle >> 3) : 32) << (ahead < 64 ? 0 : 3);
				for (i32 z = 0; z < naps; ++z) __builtin_amdgcn_s_sleep(127);
			}
			if (got >= 0) {
				// pairs with the host's release store of the tail: the ring
				// entry and the staged read bytes it published before that
				// store are read after this point
				__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "");
				s_claim = (int64_t)wall_clock64();
				s_polls = idle;
				const u64 *src = (const u64 *)&a.ring[(u64)got & a.mask];
				u64 *dst = (u64 *)&s_e;
#pragma unroll
				for (int t = 0; t < 8; ++t) dst[t] = host_ld(src + t);
				idle = 0;
				if (beat) beat_put(beat, (i32)got, 0, 1, 0);
			}
			s_seq = got;
		}
		b.sync();
		const int64_t seq = s_seq;
		if (seq < 0) break;
		// the entry is uniform: keep it in scalar registers
		SrvEntry e;
		e.q_off = uni64(s_e.q_off);
		e.d_off = uni64(s_e.d_off);
		e.ws_off = uni64(s_e.ws_off);
		e.ql = __builtin_amdgcn_readfirstlane(s_e.ql);
		e.v0 = __builtin_amdgcn_readfirstlane(s_e.v0);
		e.v1 = __builtin_amdgcn_readfirstlane(s_e.v1);
		e.hc.m = __builtin_amdgcn_readfirstlane(s_e.hc.m);
		e.hc.x = __builtin_amdgcn_readfirstlane(s_e.hc.x);
		e.hc.f = __builtin_amdgcn_readfirstlane(s_e.hc.f);
		e.hc.t = __builtin_amdgcn_readfirstlane(s_e.hc.t);
		e.hc.h = __builtin_amdgcn_readfirstlane(s_e.hc.h);
		e.hc.v = __builtin_amdgcn_readfirstlane(s_e.hc.v);
		e.res_slot = (u32)__builtin_amdgcn_readfirstlane((i32)s_e.res_slot);
		uint8_t *dq = a.pool + e.d_off;
		{
			const u64 *src = (const u64 *)(a.stage + e.q_off);
			const i32 n8 = ((e.ql + 15) / 16 + 1) * 2;
			for (i32 t = (i32)b.tid(); t < n8; t += NT) ((u64 *)dq)[t] = host_ld(src + t);
		}
		b.sync();
		__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
		HBufs hb = carve((u64 *)(a.pool + e.d_off + e.ws_off), e.hc);
		hb.vfill = (li32 *)&vfill;
		lb.VK = hb.VK, lb.VS = hb.VS, lb.vmask = hb.vmask, lb.vfill = hb.vfill;  // one visited set in both modes
		hb.sort_lds = (lu64 *)arena;
		hb.sort_cap = sort_cap;
		hb.sort_n = (i32)lds_elems(a.lc);
		hb.mw = a.mw;
		front_in_scratch(hb);
		const u64 t0 = wall_clock64();
		const ReadOut o = align_read<NW, true, MIXED>(a.g, dq, e.ql, (u32)e.v0, e.v1, lb, hb, b, (lu64 *)vis, beat,
		                                              (i32)seq);
		int64_t c = o.chars;
#pragma unroll
		for (int s = 32; s > 0; s >>= 1) c += __shfl_xor(c, s);
		if constexpr (NW > 1) {
			if (lane_id() == 0) b.redu[b.wid()] = (u64)c;
			b.sync();
			c = 0;
			for (int w = 0; w < NW; ++w) c += (int64_t)b.redu[w];
		}
		b.sync();  // every wave is done with the region
		if (b.tid() == 0) {
			SrvResult *r = &a.res[e.res_slot];
			r->ed = o.ed;
			r->status = o.status;
			r->steps = o.steps;
			r->states = o.states;
			r->chars = c;
			r->t0 = (int64_t)t0;
			r->t1 = (int64_t)wall_clock64();
			r->spill_step = o.spill_step;
			r->t_claim = s_claim;
			r->wg = (i32)blockIdx.x;
			r->polls = s_polls;
			// release: the result fields and the workgroup's writes to the
			// request's region are visible before the host sees `done`
			__hip_atomic_store(&r->done, (u64)seq + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
			if (beat) beat_put(beat, -1, 0, 0, 0);
		}
		b.sync();
	}
}

// ------------------------------------------------------------ launchers of this layout
hipError_t launch_plain(const DGraph &g, const DReads &r, const i32 *read_idx, i32 n, const u64 *ws_off, const i32 *ws_a,
                        u64 *scratch, i32 *out_ed, i32 *out_status, int64_t *out_counts, hipStream_t st) {
	if (n <= 0) return hipSuccess;
	if (g.mixed)
		hipLaunchKernelGGL(gwf_plain_kernel<true>, dim3(n), dim3(WAVE), 0, st, g, r, read_idx, 0, n, ws_off, ws_a, 0,
		                   scratch, out_ed, out_status, out_counts);
	else
		hipLaunchKernelGGL(gwf_plain_kernel<false>, dim3(n), dim3(WAVE), 0, st, g, r, read_idx, 0, n, ws_off, ws_a, 0,
		                   scratch, out_ed, out_status, out_counts);
	return hipGetLastError();
}

hipError_t launch_plain_uniform(const DGraph &g, const DReads &r, const i32 *read_idx, i32 base, i32 n, i32 a,
                                u64 *scratch, i32 *out_ed, i32 *out_status, int64_t *out_counts, hipStream_t st) {
	if (n <= 0) return hipSuccess;
	if (g.mixed)
		hipLaunchKernelGGL(gwf_plain_kernel<true>, dim3(n), dim3(WAVE), 0, st, g, r, read_idx, base, n, nullptr, nullptr,
		                   a, scratch, out_ed, out_status, out_counts);
	else
		hipLaunchKernelGGL(gwf_plain_kernel<false>, dim3(n), dim3(WAVE), 0, st, g, r, read_idx, base, n, nullptr, nullptr,
		                   a, scratch, out_ed, out_status, out_counts);
	return hipGetLastError();
}

template <int NW>
static hipError_t launch_persist(const DGraph &g, const DReads &r, const i32 *order, const i32 *bin_off, i32 bin,
                                 i32 *heads, const PersistCfg &cfg, u64 *scratch, size_t ws_bytes, i32 *out_ed,
                                 i32 *out_status, int64_t *out_counts, GwfBeat *beat, hipStream_t st) {
	auto k = g.mixed ? gwf_persistent_kernel<NW, true> : gwf_persistent_kernel<NW, false>;
	hipLaunchKernelGGL(k, dim3(cfg.wgs), dim3(WAVE * NW), cfg.lds_bytes, st, g, r, order, bin_off, bin, heads, cfg.lds,
	                   cfg.hbm, cfg.mw, (i32)(cfg.lds_bytes / 8), scratch, (u64)(ws_bytes / 8), out_ed, out_status,
	                   out_counts, beat);
	return hipGetLastError();
}

hipError_t launch_persistent(const DGraph &g, const DReads &r, const i32 *order, const i32 *bin_off, i32 bin,
                             i32 *heads, const PersistCfg &cfg, u64 *scratch, size_t ws_bytes, i32 *out_ed,
                             i32 *out_status, int64_t *out_counts, GwfBeat *beat, hipStream_t st) {
	switch (cfg.nw) {
	case 1: return launch_persist<1>(g, r, order, bin_off, bin, heads, cfg, scratch, ws_bytes, out_ed, out_status, out_counts, beat, st);
	case 2: return launch_persist<2>(g, r, order, bin_off, bin, heads, cfg, scratch, ws_bytes, out_ed, out_status, out_counts, beat, st);
	case 4: return launch_persist<4>(g, r, order, bin_off, bin, heads, cfg, scratch, ws_bytes, out_ed, out_status, out_counts, beat, st);
	default: return hipErrorInvalidValue;
	}
}

template <int NW>
static hipError_t launch_serve_nw(const SrvArgs &a, const PersistCfg &cfg, hipStream_t st) {
	auto k = a.g.mixed ? gwf_serve_kernel<NW, true> : gwf_serve_kernel<NW, false>;
	hipLaunchKernelGGL(k, dim3(cfg.wgs), dim3(WAVE * NW), cfg.lds_bytes, st, a);
	return hipGetLastError();
}

hipError_t launch_serve(const SrvArgs &a, const PersistCfg &cfg, hipStream_t st) {
	switch (cfg.nw) {
	case 1: return launch_serve_nw<1>(a, cfg, st);
	case 2: return launch_serve_nw<2>(a, cfg, st);
	case 4: return launch_serve_nw<4>(a, cfg, st);
	default: return hipErrorInvalidValue;
	}
}

// Set every dynamic-LDS limit of this layout's kernels (once, outside any graph capture).
hipError_t kernels_init() {
	const int mx = (int)gwf_max_lds_bytes();
	hipError_t e = hipSuccess;
	const void *ks[] = {(const void *)gwf_persistent_kernel<1, false>, (const void *)gwf_persistent_kernel<2, false>,
	                    (const void *)gwf_persistent_kernel<4, false>, (const void *)gwf_serve_kernel<1, false>,
	                    (const void *)gwf_serve_kernel<2, false>, (const void *)gwf_serve_kernel<4, false>,
	                    (const void *)gwf_persistent_kernel<1, true>, (const void *)gwf_persistent_kernel<2, true>,
	                    (const void *)gwf_persistent_kernel<4, true>, (const void *)gwf_serve_kernel<1, true>,
	                    (const void *)gwf_serve_kernel<2, true>, (const void *)gwf_serve_kernel<4, true>};
	for (const void *k : ks)
		if (e == hipSuccess) e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, mx);
	return e;
}

#if GWF_KEY_LAYOUT == 0
// occupancy queries (the register use of the layouts' kernels is the same)
template <int NW>
static int occ(size_t lds_bytes) {
	int nb = 0;
	(void)hipFuncSetAttribute((const void *)gwf_persistent_kernel<NW, false>, hipFuncAttributeMaxDynamicSharedMemorySize,
	                          (int)gwf_max_lds_bytes());
	if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, gwf_persistent_kernel<NW, false>, WAVE * NW, lds_bytes) != hipSuccess)
		return 1;
	return nb;
}

template <int NW>
static int occ_serve(size_t lds_bytes) {
	int nb = 0;
	(void)hipFuncSetAttribute((const void *)gwf_serve_kernel<NW, false>, hipFuncAttributeMaxDynamicSharedMemorySize,
	                          (int)gwf_max_lds_bytes());
	if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, gwf_serve_kernel<NW, false>, WAVE * NW, lds_bytes) != hipSuccess)
		return 1;
	return nb;
}
#endif

}  // namespace GWF_KNS

#if GWF_KEY_LAYOUT == 0
// ------------------------------------------------------------ shared by every layout
size_t gwf_ws_bytes(const Caps &c) { return ((gwf_layout0::ws_elems(c) * 8 + 255) / 256) * 256; }
size_t gwf_lds_bytes(const Caps &c) { return 8 * gwf_layout0::lds_elems(c); }

// ------------------------------------------------------------ device-side binning
__global__ __launch_bounds__(256) void gwf_bin_zero_kernel(i32 *hist, i32 nh, i32 *heads) {
	const i32 t = (i32)(blockIdx.x * 256 + threadIdx.x);
	if (t < nh) hist[t] = 0;
	if (t < GWF_MAX_BINS) heads[t] = 0;
}

__global__ __launch_bounds__(256) void gwf_bin_count_kernel(const i32 *probe_ed, const i32 *probe_status,
                                                           const i32 *ql, i32 n, BinBounds bb, i32 *key, i32 *hist) {
	const i32 i = (i32)(blockIdx.x * 256 + threadIdx.x);
	if (i >= n) return;
	const i32 q = ql[i];
	const i32 pl = q < 256 ? q : 256;
	const float e = (probe_status[i] == GWF_ST_OK && pl > 0) ? (float)(probe_ed[i] > 0 ? probe_ed[i] : 0) / (float)pl
	                                                         : 0.3f;
	const float ed = e * (float)q;
	const i32 est = (i32)(3.2f * ed) + 16;
	i32 bin = bb.n_bins - 1;
	for (i32 j = 0; j < bb.n_bins; ++j)
		if (est <= bb.est_max[j]) { bin = j; break; }
	// cost bucket: four per power of two of ql + 4 ED^1.5.  Durations follow
	// ql x ED more closely than ED^2, and the ED estimated from 256 bases is
	// noisy, so it is weighted below the exact length: replaying measured
	// durations of the uneven-length workload (100k reads, 0.5-20 kb, 1-10%
	// error) through 2560 workgroups, this order ends within 0.2% of the
	// longest-first order, ql + 4 ED^2 3.9% behind it.
	const u64 cost = (u64)q + (u64)(4.0f * ed * sqrtf(ed)) + 1;
	const i32 L = 63 - __clzll((long long)cost);
	const i32 quarter = L > 1 ? (i32)((cost >> (L - 2)) & 3) : 0;
	i32 u = 4 * L + quarter;
	u = u > GWF_COST_BUCKETS - 1 ? GWF_COST_BUCKETS - 1 : u;
	const i32 slot = bin * GWF_COST_BUCKETS + (GWF_COST_BUCKETS - 1 - u);  // longest first within a bin
	key[i] = slot;
	atomicAdd(&hist[slot], 1);
}

// one workgroup: exclusive scan of the (bin, bucket) counts into cursors
__global__ __launch_bounds__(1024) void gwf_bin_scan_kernel(i32 *hist, i32 nh, i32 n_bins, i32 *bin_off) {
	__shared__ i32 sh[1024];
	const i32 t = (i32)threadIdx.x;
	const i32 v = t < nh ? hist[t] : 0;
	sh[t] = v;
	__syncthreads();
	for (i32 off = 1; off < 1024; off <<= 1) {
		const i32 x = t >= off ? sh[t - off] : 0;
		__syncthreads();
		sh[t] += x;
		__syncthreads();
	}
	const i32 excl = sh[t] - v;
	if (t < nh) hist[t] = excl;
	if (t < nh && t % GWF_COST_BUCKETS == 0) bin_off[t / GWF_COST_BUCKETS] = excl;
	if (t == 0) bin_off[n_bins] = sh[1023];
}

__global__ __launch_bounds__(256) void gwf_bin_scatter_kernel(const i32 *key, i32 n, i32 *cursor, i32 *order) {
	const i32 i = (i32)(blockIdx.x * 256 + threadIdx.x);
	if (i >= n) return;
	order[atomicAdd(&cursor[key[i]], 1)] = i;
}

__global__ __launch_bounds__(256) void gwf_fill_kernel(i32 *p, i32 n, i32 v) {
	const i32 t = (i32)(blockIdx.x * 256 + threadIdx.x);
	if (t < n) p[t] = v;
}

hipError_t gwf_launch_fill(i32 *p, i32 n, i32 v, hipStream_t st) {
	if (n <= 0) return hipSuccess;
	hipLaunchKernelGGL(gwf_fill_kernel, dim3((n + 255) / 256), dim3(256), 0, st, p, n, v);
	return hipGetLastError();
}

hipError_t gwf_launch_binning(const i32 *probe_ed, const i32 *probe_status, const i32 *ql, i32 n, const BinBounds &bb,
                              i32 *key, i32 *hist, i32 *order, i32 *bin_off, i32 *heads, hipStream_t st) {
	if (bb.n_bins < 1 || bb.n_bins > GWF_MAX_BINS || n < 0) return hipErrorInvalidValue;
	const i32 nh = bb.n_bins * GWF_COST_BUCKETS;
	hipLaunchKernelGGL(gwf_bin_zero_kernel, dim3((nh + 255) / 256), dim3(256), 0, st, hist, nh, heads);
	if (n > 0)
		hipLaunchKernelGGL(gwf_bin_count_kernel, dim3((n + 255) / 256), dim3(256), 0, st, probe_ed, probe_status, ql,
		                   n, bb, key, hist);
	hipLaunchKernelGGL(gwf_bin_scan_kernel, dim3(1), dim3(1024), 0, st, hist, nh, bb.n_bins, bin_off);
	if (n > 0)
		hipLaunchKernelGGL(gwf_bin_scatter_kernel, dim3((n + 255) / 256), dim3(256), 0, st, key, n, hist, order);
	return hipGetLastError();
}

#define GWF_LAYOUT_DECLS(NS)                                                                                     \
	namespace NS {                                                                                                   \
	hipError_t launch_plain(const DGraph &, const DReads &, const i32 *, i32, const u64 *, const i32 *, u64 *, i32 *, \
	                        i32 *, int64_t *, hipStream_t);                                                           \
	hipError_t launch_plain_uniform(const DGraph &, const DReads &, const i32 *, i32, i32, i32, u64 *, i32 *, i32 *,  \
	                                int64_t *, hipStream_t);                                                          \
	hipError_t launch_persistent(const DGraph &, const DReads &, const i32 *, const i32 *, i32, i32 *,               \
	                             const PersistCfg &, u64 *, size_t, i32 *, i32 *, int64_t *, GwfBeat *, hipStream_t); \
	hipError_t launch_serve(const SrvArgs &, const PersistCfg &, hipStream_t);                                       \
	hipError_t kernels_init();                                                                                       \
	}
GWF_LAYOUT_DECLS(gwf_layout1)
GWF_LAYOUT_DECLS(gwf_layout2)
GWF_LAYOUT_DECLS(gwf_layout3)
static_assert(GWF_N_LAYOUTS == 4, "one namespace per layout");

// the layout named by g.layout (set by the runtime for the graph and reads of a run)
#define GWF_BY_LAYOUT(layout, call)                      \
	switch (layout) {                                    \
	case 0: return gwf_layout0::call;                    \
	case 1: return gwf_layout1::call;                    \
	case 2: return gwf_layout2::call;                    \
	case 3: return gwf_layout3::call;                    \
	default: return hipErrorInvalidValue;                \
	}

hipError_t gwf_launch_plain(const DGraph &g, const DReads &r, const i32 *read_idx, i32 n, const u64 *ws_off,
                            const i32 *ws_a, u64 *scratch, i32 *out_ed, i32 *out_status, int64_t *out_counts,
                            hipStream_t st) {
	GWF_BY_LAYOUT(g.layout, launch_plain(g, r, read_idx, n, ws_off, ws_a, scratch, out_ed, out_status, out_counts, st))
}

hipError_t gwf_launch_plain_uniform(const DGraph &g, const DReads &r, const i32 *read_idx, i32 base, i32 n, i32 a,
                                    u64 *scratch, i32 *out_ed, i32 *out_status, int64_t *out_counts, hipStream_t st) {
	GWF_BY_LAYOUT(g.layout, launch_plain_uniform(g, r, read_idx, base, n, a, scratch, out_ed, out_status, out_counts, st))
}

hipError_t gwf_launch_persistent(const DGraph &g, const DReads &r, const i32 *order, const i32 *bin_off, i32 bin,
                                 i32 *heads, const PersistCfg &cfg, u64 *scratch, size_t ws_bytes, i32 *out_ed,
                                 i32 *out_status, int64_t *out_counts, GwfBeat *beat, hipStream_t st) {
	GWF_BY_LAYOUT(g.layout, launch_persistent(g, r, order, bin_off, bin, heads, cfg, scratch, ws_bytes, out_ed, out_status,
	                                          out_counts, beat, st))
}

hipError_t gwf_launch_serve(const SrvArgs &a, const PersistCfg &cfg, hipStream_t st) {
	GWF_BY_LAYOUT(a.g.layout, launch_serve(a, cfg, st))
}

hipError_t gwf_kernels_init() {
	hipError_t e = gwf_layout0::kernels_init();
	if (e == hipSuccess) e = gwf_layout1::kernels_init();
	if (e == hipSuccess) e = gwf_layout2::kernels_init();
	if (e == hipSuccess) e = gwf_layout3::kernels_init();
	return e;
}

size_t gwf_max_lds_bytes() { return 160 * 1024 - 4096; }

// The largest dynamic LDS size (a multiple of 256 bytes, at least lds_bytes)
// that keeps per_cu workgroups of nw waves resident per CU, leaving `margin`
// bytes of each workgroup's share unallocated.
size_t gwf_persist_grow_lds(int nw, size_t lds_bytes, int per_cu, size_t margin) {
	if (per_cu < 1) return lds_bytes;
	size_t share = (160 * 1024) / (size_t)per_cu;
	if (share <= margin) return lds_bytes;
	size_t cand = (share - margin) / 256 * 256;
	if (cand > gwf_max_lds_bytes()) cand = gwf_max_lds_bytes() / 256 * 256;
	while (cand > lds_bytes && gwf_persist_occupancy(nw, cand) < per_cu) cand -= 256;
	return cand > lds_bytes ? cand : lds_bytes;
}

int gwf_persist_occupancy(int nw, size_t lds_bytes) {
	switch (nw) {
	case 1: return gwf_layout0::occ<1>(lds_bytes);
	case 2: return gwf_layout0::occ<2>(lds_bytes);
	case 4: return gwf_layout0::occ<4>(lds_bytes);
	default: return 1;
	}
}

int gwf_serve_occupancy(int nw, size_t lds_bytes) {
	switch (nw) {
	case 1: return gwf_layout0::occ_serve<1>(lds_bytes);
	case 2: return gwf_layout0::occ_serve<2>(lds_bytes);
	case 4: return gwf_layout0::occ_serve<4>(lds_bytes);
	default: return 1;
	}
}
#endif  // GWF_KEY_LAYOUT == 0
