// gwf_kernels.hip — CDNA4 (gfx950) kernels for GWFA graph edit distance.
//
// One 64-lane wavefront aligns one read.  Per score step s (gwf_ed_extend,
// gwf-ed.cu:1064-1290) the wave:
//   1. extends every wavefront state (gwf_extend1, gwf-ed.cu:806-839): each
//      lane compares 16 bytes with two 8-byte XORs; states still matching are
//      then extended cooperatively, 512 bytes per wave step (8 bytes per lane,
//      a 64-bit ballot and __ffsll pick the first mismatching lane);
//   2. emits the in-vertex successors (the WFA recurrence of
//      gwf_ed_extend_batch, gwf-ed.cu:891-1018) directly in sorted order: the
//      state on diagonal d owns output diagonals d-1 / d / d+1 unless a
//      neighbour owns them, and takes the max over its neighbours' candidates;
//   3. runs the vertex-crossing closure (the A queue, gwf-ed.cu:1106-1255)
//      level by level over a frontier, reading the CSR arcs of each vertex;
//   4. sorts the closure's successors (bitonic in LDS), merges them with the
//      in-vertex successors, keeps the furthest state per diagonal
//      (gwf_diag_dedup, gwf-ed.cu:653-687) and drops diagonals inside the
//      forbidden-interval history (gwf_dedup / gwf_mixed_dedup,
//      gwf-ed.cu:691-804), which is updated with this step's exits.
// All arithmetic is exact integer; every buffer write is bounds-checked and a
// read that outgrows its workspace reports GWF_ST_OVERFLOW (the host re-runs
// it with a larger one) instead of dropping state.
#include "gwf_dev.h"

#define WAVE 64
#define VIS_SLOTS 512     // per-step visited set (w, i+1), LDS
#define SORT_LDS 1024     // largest run sorted in LDS in one piece
#define EMPTY_KEY (~0ull)

// ------------------------------------------------------------ key helpers
__device__ __forceinline__ u64 mk(u32 v, i32 d, i32 k) {
	return (u64)v << 40 | (u64)(u32)(d + GWF_D_BIAS) << 20 | (u64)(u32)(GWF_KINV_MAX - k);
}
__device__ __forceinline__ u32 key_v(u64 x) { return (u32)(x >> 40); }
__device__ __forceinline__ i32 key_d(u64 x) { return (i32)((x >> 20) & 0xFFFFFu) - GWF_D_BIAS; }
__device__ __forceinline__ i32 key_k(u64 x) { return GWF_KINV_MAX - (i32)(x & 0xFFFFFu); }
__device__ __forceinline__ u64 key_diag(u64 x) { return x >> 20; }
__device__ __forceinline__ u64 diag_id(u32 v, i32 d) { return (u64)v << 20 | (u64)(u32)(d + GWF_D_BIAS); }

// ------------------------------------------------------------ wave helpers
__device__ __forceinline__ u32 lane_id() { return __lane_id(); }

__device__ __forceinline__ u32 rank_below(u64 mask) {
	return __builtin_amdgcn_mbcnt_hi((u32)(mask >> 32), __builtin_amdgcn_mbcnt_lo((u32)mask, 0u));
}

__device__ __forceinline__ i32 uni(i32 x) { return __builtin_amdgcn_readfirstlane(x); }

// exclusive prefix sum over the wave; *total receives the wave sum
__device__ __forceinline__ i32 scan_excl(i32 x, i32 *total) {
	const u32 lane = lane_id();
	i32 incl = x;
#pragma unroll
	for (int o = 1; o < WAVE; o <<= 1) {
		i32 y = __shfl_up(incl, o);
		if (lane >= (u32)o) incl += y;
	}
	*total = __shfl(incl, WAVE - 1);
	return incl - x;
}

// inclusive prefix max over the wave (u64)
__device__ __forceinline__ u64 scan_max_incl(u64 x) {
	const u32 lane = lane_id();
#pragma unroll
	for (int o = 1; o < WAVE; o <<= 1) {
		u64 y = __shfl_up(x, o);
		if (lane >= (u32)o) x = x > y ? x : y;
	}
	return x;
}

__device__ __forceinline__ void wsync() { __syncthreads(); }

// unaligned 8-byte load from a buffer padded past its end
__device__ __forceinline__ u64 load8u(const uint8_t *p) {
	const uintptr_t a = (uintptr_t)p;
	const u64 *w = (const u64 *)(a & ~(uintptr_t)7);
	const u32 sh = (u32)(a & 7) * 8;
	const u64 lo = w[0], hi = w[1];
	return sh ? (lo >> sh) | (hi << (64 - sh)) : lo;
}

// first index in a[0..n) with a[i] >= x (lower) or a[i] > x (upper)
__device__ __forceinline__ i32 lower_bound(const u64 *a, i32 n, u64 x) {
	i32 lo = 0, hi = n;
	while (lo < hi) {
		i32 m = (lo + hi) >> 1;
		if (a[m] < x) lo = m + 1; else hi = m;
	}
	return lo;
}
__device__ __forceinline__ i32 upper_bound(const u64 *a, i32 n, u64 x) {
	i32 lo = 0, hi = n;
	while (lo < hi) {
		i32 m = (lo + hi) >> 1;
		if (a[m] <= x) lo = m + 1; else hi = m;
	}
	return lo;
}

// ------------------------------------------------------------ extension
// LCE of graph bytes ts[k+1..] and query bytes qd[k+1..] (qd = q + d) up to
// max_k: the signed-max_k form of gwf_extend1 (gwf-ed.cu:806-839).  Every
// lane passes its own state (act = false lanes pass through); all lanes must
// call it together.
__device__ i32 extend_wave(bool act, const uint8_t *seq, u32 seq_off, const uint8_t *q, i32 d, i32 k,
                           i32 max_k, int64_t *chars) {
	bool more = false;
	if (act) {
		const uint8_t *ts = seq + seq_off;
		const uint8_t *qd = q + d;
#pragma unroll
		for (int it = 0; it < 2; ++it) {
			const i32 avail = max_k - k;
			if (avail <= 0) break;
			const u64 x = load8u(ts + k + 1) ^ load8u(qd + k + 1);
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
		const i32 kk = __shfl(k, src), mk_ = __shfl(max_k, src), dd = __shfl(d, src);
		const u32 so = (u32)__shfl((i32)seq_off, src);
		const uint8_t *ts = seq + so;
		const uint8_t *qd = q + dd;
		i32 base = kk;
		i32 res;
		for (;;) {
			const i32 avail = mk_ - base;              // chars still allowed
			const i32 my = (i32)lane * 8;              // lane's first char (relative)
			i32 m = 8;
			if (my < avail) {
				const u64 x = load8u(ts + base + 1 + my) ^ load8u(qd + base + 1 + my);
				m = x ? (i32)(__builtin_ctzll(x) >> 3) : 8;
			}
			const bool stop = (my < avail) && (m < 8 || my + 8 >= avail);
			const u64 sm = __ballot(stop);
			if (sm) {
				const int f = __ffsll((unsigned long long)sm) - 1;
				const i32 got = __shfl(my + m, f);
				res = base + (got < avail ? got : avail);
				break;
			}
			if (avail <= 0) { res = base; break; }
			base += WAVE * 8;
		}
		if ((int)lane == src) { *chars += res - kk + (res < mk_ ? 1 : 0); k = res; }
	}
	return k;
}

// ------------------------------------------------------------ sorting/merging
// Bitonic sort of lds[0..N) (N a power of two, padded with EMPTY_KEY).
__device__ void bitonic_lds(u64 *lds, i32 N) {
	const u32 lane = lane_id();
	for (i32 size = 2; size <= N; size <<= 1) {
		for (i32 stride = size >> 1; stride > 0; stride >>= 1) {
			for (i32 t = lane; t < (N >> 1); t += WAVE) {
				const i32 i = 2 * t - (t & (stride - 1));
				const i32 j = i + stride;
				const bool up = (i & size) == 0;
				const u64 a = lds[i], b = lds[j];
				if ((a > b) == up) { lds[i] = b; lds[j] = a; }
			}
			wsync();
		}
	}
}

// Merge sorted a[0..na) and b[0..nb) into out (a before b on ties).
__device__ void merge_keys(const u64 *a, i32 na, const u64 *b, i32 nb, u64 *out) {
	const u32 lane = lane_id();
	for (i32 j = lane; j < na; j += WAVE) {
		const u64 x = a[j];
		out[j + lower_bound(b, nb, x)] = x;
	}
	for (i32 t = lane; t < nb; t += WAVE) {
		const u64 x = b[t];
		out[t + upper_bound(a, na, x)] = x;
	}
	wsync();
}

// Sort p[0..n) ascending; lds holds SORT_LDS keys; tmp is a global buffer of n keys.
__device__ void sort_keys(u64 *p, i32 n, u64 *lds, u64 *tmp) {
	const u32 lane = lane_id();
	if (n <= 1) return;
	for (i32 c0 = 0; c0 < n; c0 += SORT_LDS) {  // sort runs of SORT_LDS in LDS
		const i32 m = n - c0 < SORT_LDS ? n - c0 : SORT_LDS;
		i32 N = 2;
		while (N < m) N <<= 1;
		for (i32 t = lane; t < N; t += WAVE) lds[t] = t < m ? p[c0 + t] : EMPTY_KEY;
		wsync();
		bitonic_lds(lds, N);
		for (i32 t = lane; t < m; t += WAVE) p[c0 + t] = lds[t];
		wsync();
	}
	u64 *src = p, *dst = tmp;
	for (i32 w = SORT_LDS; w < n; w <<= 1) {  // pairwise merges in global memory
		for (i32 l = 0; l < n; l += 2 * w) {
			const i32 mid = l + w < n ? l + w : n;
			const i32 r = l + 2 * w < n ? l + 2 * w : n;
			merge_keys(src + l, mid - l, src + mid, r - mid, dst + l);
		}
		u64 *t = src; src = dst; dst = t;
	}
	if (src != p) {
		for (i32 t = lane; t < n; t += WAVE) p[t] = src[t];
		wsync();
	}
}

// Merge interval lists (alo/ahi by alo) and unit intervals b (as [b, b+1))
// into olo/ohi.
__device__ void merge_intervals(const u64 *alo, const u64 *ahi, i32 na, const u64 *b, i32 nb, u64 *olo,
                                u64 *ohi) {
	const u32 lane = lane_id();
	for (i32 j = lane; j < na; j += WAVE) {
		const u64 x = alo[j];
		const i32 p = j + lower_bound(b, nb, x);
		olo[p] = x, ohi[p] = ahi[j];
	}
	for (i32 t = lane; t < nb; t += WAVE) {
		const u64 x = b[t];
		const i32 p = t + upper_bound(alo, na, x);
		olo[p] = x, ohi[p] = x + 1;
	}
	wsync();
}

// Coalesce overlapping or adjacent intervals of a list sorted by lo
// (gwf_intv_merge_adj, gwf-ed.cu:266-296).  Returns the count, or -1 if it
// exceeds cap.
__device__ i32 coalesce(const u64 *lo, const u64 *hi, i32 n, u64 *olo, u64 *ohi, i32 cap) {
	const u32 lane = lane_id();
	u64 carry = 0;      // max hi over previous chunks
	i32 ng = 0;         // groups started so far
	bool ovf = false;
	for (i32 base = 0; base < n; base += WAVE) {
		const i32 j = base + lane;
		const bool act = j < n;
		const u64 l = act ? lo[j] : EMPTY_KEY;
		const u64 h = act ? hi[j] : 0;
		u64 pin = scan_max_incl(h);
		pin = pin > carry ? pin : carry;
		u64 pex = __shfl_up(pin, 1);
		if (lane == 0) pex = carry;
		const bool start = act && (j == 0 || l > pex);
		const u64 nl = (act && j + 1 < n) ? lo[j + 1] : EMPTY_KEY;
		const bool end = act && (j + 1 == n || nl > pin);
		const u64 sm = __ballot(start);
		const i32 g = ng + (i32)rank_below(sm) + (start ? 1 : 0) - 1;
		if (g >= cap && act) ovf = true;
		if (!ovf && start) olo[g] = l;
		if (!ovf && end) ohi[g] = pin;
		ng += __popcll(sm);
		carry = __shfl(pin, WAVE - 1);
	}
	wsync();
	return __ballot(ovf) ? -1 : ng;
}

__device__ __forceinline__ bool forbidden(const u64 *hlo, const u64 *hhi, i32 nh, u64 dg) {
	const i32 p = upper_bound(hlo, nh, dg) - 1;
	return p >= 0 && dg < hhi[p];
}

// ------------------------------------------------------------ visited set
__device__ __forceinline__ u32 vis_hash(u64 x) {
	x ^= x >> 31;
	x *= 0x7fb5d329728ea185ULL;
	x ^= x >> 27;
	return (u32)x;
}
// true if key was absent (and is now present); a full probe window reports
// "absent", which only costs duplicate work.
__device__ __forceinline__ bool vis_insert(u64 *vis, u64 key) {
	u32 h = vis_hash(key) & (VIS_SLOTS - 1);
	for (int p = 0; p < 16; ++p) {
		const u64 old = atomicCAS((unsigned long long *)&vis[h], (unsigned long long)EMPTY_KEY,
		                          (unsigned long long)key);
		if (old == EMPTY_KEY) return true;
		if (old == key) return false;
		h = (h + 1) & (VIS_SLOTS - 1);
	}
	return true;
}

// ------------------------------------------------------------ workspace
struct Ws {
	u64 *A, *S, *M, *X, *F0, *F1, *T, *Tb, *Clo, *Chi, *H0lo, *H0hi, *H1lo, *H1hi, *SG;
};

__host__ __device__ inline size_t ws_elems(const Caps &c) {
	return (size_t)c.a + c.s + (size_t)(c.s + c.x) + c.x + 2 * (size_t)c.f + 2 * (size_t)c.t +
	       2 * (size_t)(c.h + c.t) + 4 * (size_t)c.h + (size_t)(c.x > c.t ? c.x : c.t) + 8;
}

size_t gwf_ws_bytes(const Caps &c) { return ((ws_elems(c) * 8 + 255) / 256) * 256; }

__device__ inline Ws carve(u64 *base, const Caps &c) {
	Ws w;
	u64 *p = base;
	w.A = p; p += c.a;
	w.S = p; p += c.s;
	w.M = p; p += c.s + c.x;
	w.X = p; p += c.x;
	w.F0 = p; p += c.f;
	w.F1 = p; p += c.f;
	w.T = p; p += c.t;
	w.Tb = p; p += c.t;
	w.Clo = p; p += c.h + c.t;
	w.Chi = p; p += c.h + c.t;
	w.H0lo = p; p += c.h;
	w.H0hi = p; p += c.h;
	w.H1lo = p; p += c.h;
	w.H1hi = p; p += c.h;
	w.SG = p;
	return w;
}

// wave-append helper: returns this lane's slot (valid if pred) and advances *n
__device__ __forceinline__ i32 wave_slot(bool pred, i32 *n) {
	const u64 m = __ballot(pred);
	const i32 s = *n + (i32)rank_below(m);
	*n += __popcll(m);
	return s;
}

// ------------------------------------------------------------ one read
// Returns the ED (>= 0), -1 if the wavefront empties, or sets *status.
__device__ i32 align_read(const DGraph &g, const uint8_t *q, i32 ql, u32 v0, const Caps &cap, Ws &w,
                          u64 *lds_sort, u64 *vis, i32 *status, int64_t *n_steps, int64_t *n_states,
                          int64_t *n_chars) {
	const u32 lane = lane_id();
	i32 n_a = 1, n_h = 0;
	u64 *Hlo = w.H0lo, *Hhi = w.H0hi, *Hlo2 = w.H1lo, *Hhi2 = w.H1hi;
	int64_t chars = 0;
	if (lane == 0) w.A[0] = mk(v0, 0, -1);
	wsync();
	*status = GWF_ST_OK;
	for (i32 s = 0;; ++s) {
		*n_steps += 1;
		*n_states += n_a;
		if (s > ql + 1) { *status = GWF_ST_LIMIT; return -1; }
		// ---- 1. extend, detect query end, collect vertex ends
		bool found = false;
		i32 n_f = 0;
		for (i32 base = 0; base < n_a; base += WAVE) {
			const i32 j = base + lane;
			const bool act = j < n_a;
			const u64 key = act ? w.A[j] : 0;
			const u32 v = key_v(key);
			const i32 d = key_d(key);
			i32 k = key_k(key);
			const uint4 vi = act ? g.vinfo[v] : make_uint4(0, 0, 0, 0);
			const i32 vl = (i32)vi.y;
			const i32 max_k = (ql - d < vl ? ql - d : vl) - 1;
			k = extend_wave(act, g.seq, vi.x, q, d, k, max_k, &chars);
			if (act) w.A[j] = mk(v, d, k);
			const bool qend = act && d + k == ql - 1;
			if (__ballot(qend)) found = true;
			const bool vend = act && !qend && k == vl - 1;
			const i32 slot = wave_slot(vend, &n_f);
			if (vend && slot < cap.f) w.F0[slot] = mk(v, d, k);
		}
		if (found) { *n_chars += chars; return s; }
		if (n_f > cap.f) { *status = GWF_ST_OVERFLOW; return -1; }
		wsync();

		// ---- 2. in-vertex successors S (sorted) and exits T1
		i32 n_s = 0, n_t = 0;
		bool ovf = false;
		for (i32 base = 0; base < n_a; base += WAVE) {
			const i32 j = base + lane;
			const bool act = j < n_a;
			const u64 cur = act ? w.A[j] : 0;
			const u64 prv = (act && j > 0) ? w.A[j - 1] : EMPTY_KEY;
			const u64 nxt = (act && j + 1 < n_a) ? w.A[j + 1] : EMPTY_KEY;
			const u32 v = key_v(cur);
			const i32 d = key_d(cur), k = key_k(cur);
			const i32 vl = act ? (i32)g.vinfo[v].y : 0;
			const bool hp = act && j > 0 && key_v(prv) == v;
			const bool hn = act && j + 1 < n_a && key_v(nxt) == v;
			const i32 dp = key_d(prv), kp = key_k(prv), dn = key_d(nxt), kn = key_k(nxt);
			// outputs on diagonals d-1, d, d+1 (max over ins/mis/del candidates)
			const bool e0 = act && !(hp && dp >= d - 2);
			const i32 k0 = k + 1;
			i32 k1 = k + 1;
			if (hp && dp == d - 1 && kp > k1) k1 = kp;
			if (hn && dn == d + 1 && kn + 1 > k1) k1 = kn + 1;
			const bool e1 = act;
			const bool e2 = act && !(hn && dn == d + 1);
			i32 k2 = k;
			if (hn && dn == d + 2 && kn + 1 > k2) k2 = kn + 1;
			const bool s0 = e0 && (d - 1) + k0 < ql && k0 < vl, t0 = e0 && k0 == vl;
			const bool s1 = e1 && d + k1 < ql && k1 < vl, t1 = e1 && k1 == vl;
			const bool s2 = e2 && (d + 1) + k2 < ql && k2 < vl, t2 = e2 && k2 == vl;
			i32 tot;
			i32 p = n_s + scan_excl((i32)s0 + (i32)s1 + (i32)s2, &tot);
			if (n_s + tot > cap.s) ovf = true;
			else {
				if (s0) w.S[p++] = mk(v, d - 1, k0);
				if (s1) w.S[p++] = mk(v, d, k1);
				if (s2) w.S[p++] = mk(v, d + 1, k2);
			}
			n_s += tot;
			p = n_t + scan_excl((i32)t0 + (i32)t1 + (i32)t2, &tot);
			if (n_t + tot > cap.t) ovf = true;
			else {
				if (t0) w.T[p++] = diag_id(v, d - 1);
				if (t1) w.T[p++] = diag_id(v, d);
				if (t2) w.T[p++] = diag_id(v, d + 1);
			}
			n_t += tot;
		}
		if (ovf) { *status = GWF_ST_OVERFLOW; return -1; }
		const i32 n_t1 = n_t;

		// ---- 3. vertex-crossing closure
		for (i32 t = lane; t < VIS_SLOTS; t += WAVE) vis[t] = EMPTY_KEY;
		wsync();
		i32 n_x = 0;
		u64 *fb = w.F0, *nb = w.F1;
		bool closure_origin = false;
		while (n_f > 0) {
			i32 n_na = 0;
			for (i32 base = 0; base < n_f; base += WAVE) {
				const i32 j = base + lane;
				const bool act = j < n_f;
				const u64 key = act ? fb[j] : 0;
				const u32 v = key_v(key);
				const i32 d = key_d(key), k = key_k(key);
				const i32 i = d + k;
				const uint4 vi = act ? g.vinfo[v] : make_uint4(0, 0, 0, 0);
				const bool rec = act && closure_origin;
				const i32 ts = wave_slot(rec, &n_t);
				if (rec && ts < cap.t) w.T[ts] = diag_id(v, d);
				const uint8_t c = act ? q[i + 1] : 0;
				const i32 deg = act ? (i32)vi.w : 0;
				i32 maxdeg = deg;
				for (int o = 32; o > 0; o >>= 1) {
					const i32 y = __shfl_xor(maxdeg, o);
					maxdeg = maxdeg > y ? maxdeg : y;
				}
				i32 n_ext = 0;
				for (i32 e = 0; e < maxdeg; ++e) {
					const bool has = e < deg;
					const u32 wv = has ? g.arc_w[vi.z + e] : 0;
					const i32 ol = has ? g.arc_o[vi.z + e] : 0;
					const u32 woff = has ? g.vinfo[wv].x : 0;
					const bool m = has && g.seq[woff + (u32)ol] == c;
					n_ext += m;
					const bool absent = has && vis_insert(vis, (u64)wv << 32 | (u32)(i + 1));
					const bool pa = absent && m;
					const i32 sa = wave_slot(pa, &n_na);
					if (pa && sa < cap.f) nb[sa] = mk(wv, i + 1 - ol, ol);
					const bool px = absent && !m;
					const u64 xm = __ballot(px);
					const i32 sx = n_x + 2 * (i32)rank_below(xm);
					n_x += 2 * __popcll(xm);
					if (px && sx + 1 < cap.x) {
						w.X[sx] = mk(wv, i - ol, ol);
						w.X[sx + 1] = mk(wv, i + 1 - ol, ol);
					}
				}
				const bool pi = rec && (deg == 0 || n_ext != deg);
				const i32 si = wave_slot(pi, &n_x);
				if (pi && si < cap.x) w.X[si] = mk(v, d + 1, k);
			}
			if (n_na > cap.f || n_x > cap.x || n_t > cap.t) { *status = GWF_ST_OVERFLOW; return -1; }
			wsync();
			// extend the states that entered a vertex with a match
			i32 n_f2 = 0;
			for (i32 base = 0; base < n_na; base += WAVE) {
				const i32 j = base + lane;
				const bool act = j < n_na;
				const u64 key = act ? nb[j] : 0;
				const u32 v = key_v(key);
				const i32 d = key_d(key);
				i32 k = key_k(key);
				const uint4 vi = act ? g.vinfo[v] : make_uint4(0, 0, 0, 0);
				const i32 vl = (i32)vi.y;
				const i32 max_k = (ql - d < vl ? ql - d : vl) - 1;
				k = extend_wave(act, g.seq, vi.x, q, d, k, max_k, &chars);
				const bool qend = act && d + k == ql - 1;
				if (__ballot(qend)) found = true;
				const bool vend = act && !qend && k == vl - 1;
				const bool mid = act && !qend && !vend;
				const i32 sf = wave_slot(vend, &n_f2);
				if (vend && sf < cap.f) fb[sf] = mk(v, d, k);
				const u64 xm = __ballot(mid);
				const i32 sx = n_x + 3 * (i32)rank_below(xm);
				n_x += 3 * __popcll(xm);
				if (mid && sx + 2 < cap.x) {
					w.X[sx] = mk(v, d - 1, k + 1);
					w.X[sx + 1] = mk(v, d, k + 1);
					w.X[sx + 2] = mk(v, d + 1, k);
				}
			}
			if (found) { *n_chars += chars; return s; }
			if (n_f2 > cap.f || n_x > cap.x) { *status = GWF_ST_OVERFLOW; return -1; }
			wsync();
			n_f = n_f2;
			closure_origin = true;
		}

		// ---- 4. sort the closure outputs and this step's closure exits
		sort_keys(w.X, n_x, lds_sort, w.SG);
		sort_keys(w.T + n_t1, n_t - n_t1, lds_sort, w.SG);

		// ---- 5. merge in-vertex and closure successors
		merge_keys(w.S, n_s, w.X, n_x, w.M);
		const i32 n_m = n_s + n_x;

		// ---- 6. forbidden-interval history: H2 = coalesce(H u T)
		merge_keys(w.T, n_t1, w.T + n_t1, n_t - n_t1, w.Tb);
		merge_intervals(Hlo, Hhi, n_h, w.Tb, n_t, w.Clo, w.Chi);
		const i32 nh2 = coalesce(w.Clo, w.Chi, n_h + n_t, Hlo2, Hhi2, cap.h);
		if (nh2 < 0) { *status = GWF_ST_OVERFLOW; return -1; }
		{ u64 *t = Hlo; Hlo = Hlo2; Hlo2 = t; t = Hhi; Hhi = Hhi2; Hhi2 = t; }
		n_h = nh2;

		// ---- 7. keep the furthest state per diagonal, drop forbidden ones
		i32 n_new = 0;
		for (i32 base = 0; base < n_m; base += WAVE) {
			const i32 j = base + lane;
			const bool act = j < n_m;
			const u64 x = act ? w.M[j] : 0;
			const u64 dg = key_diag(x);
			const bool first = act && (j == 0 || key_diag(w.M[j - 1]) != dg);
			const bool keep = first && !forbidden(Hlo, Hhi, n_h, dg);
			const i32 slot = wave_slot(keep, &n_new);
			if (keep && slot < cap.a) w.A[slot] = x;
		}
		if (n_new > cap.a) { *status = GWF_ST_OVERFLOW; return -1; }
		wsync();
		n_a = n_new;
		if (n_a == 0) { *n_chars += chars; return -1; }
	}
}

// ------------------------------------------------------------ kernels
__global__ __launch_bounds__(WAVE) void gwf_plain_kernel(DGraph g, DReads rd, const i32 *read_idx, i32 n,
                                                         Caps caps, u64 *scratch, u64 ws_elems_per_read,
                                                         i32 *out_ed, i32 *out_status, int64_t *out_counts) {
	__shared__ u64 lds_sort[SORT_LDS];
	__shared__ u64 vis[VIS_SLOTS];
	const i32 b = blockIdx.x;
	if (b >= n) return;
	const i32 r = read_idx ? read_idx[b] : b;
	Ws w = carve(scratch + (size_t)b * ws_elems_per_read, caps);
	const u64 t0 = wall_clock64();
	int64_t steps = 0, states = 0, chars = 0;
	i32 st;
	const i32 ed = align_read(g, rd.seq + rd.off[r], rd.ql[r], (u32)rd.v0[r], caps, w, lds_sort, vis, &st,
	                          &steps, &states, &chars);
	// chars is per lane; reduce
	for (int o = 32; o > 0; o >>= 1) chars += __shfl_xor(chars, o);
	if (lane_id() == 0) {
		int64_t *c = out_counts + GWF_NCOUNT * (size_t)r;
		out_ed[r] = ed;
		out_status[r] = st;
		c[0] = steps;
		c[1] = states;
		c[2] = chars;
		c[3] = (int64_t)t0;
		c[4] = (int64_t)wall_clock64();
	}
}

hipError_t gwf_launch_plain(const DGraph &g, const DReads &r, const i32 *read_idx, i32 n, const Caps &caps,
                            u64 *scratch, size_t ws_bytes, i32 *out_ed, i32 *out_status, int64_t *out_counts,
                            hipStream_t st) {
	if (n <= 0) return hipSuccess;
	hipLaunchKernelGGL(gwf_plain_kernel, dim3(n), dim3(WAVE), 0, st, g, r, read_idx, n, caps, scratch,
	                   (u64)(ws_bytes / 8), out_ed, out_status, out_counts);
	return hipGetLastError();
}
