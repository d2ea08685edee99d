/*
 * gwf_oracle.c — TEST INFRASTRUCTURE ONLY (the checker and the timed CPU
 * baseline; never linked into or called by the product).
 *
 * A sequential CPU restatement of the reference's device GWFA routine
 * (gwf-ed.cu:80-1337).  Every block of the reference that runs redundantly on
 * 32 lanes is written here once, as the single-lane program it computes.
 * Where the reference's code has quirks that change results or drop state,
 * this file keeps the intended (exact) semantics instead:
 *   A1 `max_k` is signed (gwf-ed.cu:809 declares it uint32_t);
 *   A3 buffers grow instead of dropping entries (gwf-ed.cu:754-758, 1008,
 *      1054-1061, 1177) and the visited set never reports "present" when full;
 *   A4 no ring/append overflow (gwf-ed.cu:849-851, 882-887, 1263-1270);
 *   A8 the score loop stops when an end is found even if the end offset is -1
 *      (gwf-ed.cu:1327 tests end_off >= 0 and spins on an all-insertion end);
 *   A9 sortedness is tested over every adjacent pair (gwf-ed.cu:219-263 and
 *      372-417 skip pairs across their 32-element double buffer).
 * Everything else (run grouping, the WFA recurrence and its tie order, the
 * A-queue closure with its visited set keyed (w, i+1), the do_dedup skip, the
 * ooo partition sort, keep-first-max-k dedup, interval history and the
 * forbidden-interval filter) follows the reference step by step.
 *
 * Pinning: oracle_dp_ed() below is an independent textbook DP over graph
 * positions; tests compare the two on random graphs/reads and on the shipped
 * fixtures: equal for v1 < 0 whenever no visited-set claim was refused to an
 * arc of another overlap (always so when every vertex's in-arcs share one
 * overlap), never below it otherwise; for v1 >= 0 equal wherever the
 * vertex-end insertion skip cannot fire at v1, never below it otherwise.
 */
#define _GNU_SOURCE
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#include <pthread.h>
#include <unistd.h>
#include "gwf_oracle.h"

#define OR_DIAG_SHIFT 0x40000000 /* gwf-ed.cu:11 */

typedef struct { uint64_t vd; int32_t k; uint32_t xo; } odiag_t; /* gwfa.h:27-31 */

static void *or_grow(void *p, size_t *m, size_t need, size_t elt)
{
	size_t nm;
	if (need <= *m) return p;
	nm = *m ? *m : 64;
	while (nm < need) nm <<= 1;
	p = realloc(p, nm * elt);
	if (!p) { fprintf(stderr, "oracle: out of memory\n"); abort(); }
	*m = nm;
	return p;
}

/* gwf_gen_vd, gwf-ed.cu:80-83 */
static inline uint64_t or_gen_vd(uint32_t v, int32_t d)
{
	return (uint64_t)v << 32 | (uint32_t)(OR_DIAG_SHIFT + d);
}
static inline int32_t or_vd_d(uint64_t vd) { return (int32_t)(uint32_t)vd - OR_DIAG_SHIFT; }

/* Per-read working state: the fields of gwf_edbuf_t (gwfa.h:54-80) and
 * `counters` (gwfa.h:44-52), all growable. */
typedef struct {
	uint64_t *a_vd; int32_t *a_k; uint32_t *a_xo; size_t m_a; int64_t n_a;
	odiag_t *A; size_t m_A; int64_t head_A, n_A;      /* FIFO, reset per step */
	odiag_t *B; size_t m_B; int64_t n_B;
	odiag_t *ooo; size_t m_ooo;
	odiag_t *scr; size_t m_scr;
	uint64_t *tmp0, *tmp1; size_t m_tmp0, m_tmp1; int64_t n_tmp;
	uint64_t *intv0, *intv1; size_t m_intv0, m_intv1; int64_t n_intv;
	uint64_t *swap0, *swap1; size_t m_swap0, m_swap1; int64_t n_swap;
	uint64_t *ha; size_t ha_cap, ha_n;                /* visited set, 0 = empty */
	int32_t *hol;                                     /* overlap of the arc that claimed each key */
	int64_t nc;                                       /* chars compared */
	int64_t na_events, ncross;
	int64_t nlossy;                                   /* claims refused to a different overlap */
} obuf_t;

static void ob_free(obuf_t *b)
{
	free(b->a_vd); free(b->a_k); free(b->a_xo); free(b->A); free(b->B); free(b->ooo);
	free(b->scr); free(b->tmp0); free(b->tmp1); free(b->intv0); free(b->intv1);
	free(b->swap0); free(b->swap1); free(b->ha); free(b->hol);
}

static void ob_reserve_a(obuf_t *b, size_t n)
{
	size_t m0 = b->m_a, m1 = b->m_a, m2 = b->m_a;
	b->a_vd = or_grow(b->a_vd, &m0, n, sizeof(uint64_t));
	b->a_k = or_grow(b->a_k, &m1, n, sizeof(int32_t));
	b->a_xo = or_grow(b->a_xo, &m2, n, sizeof(uint32_t));
	b->m_a = m0;
}

/* static_A_pushp, gwf-ed.cu:841-857 (growable FIFO instead of a ring) */
static void or_A_push(obuf_t *b, uint64_t vd, int32_t k, uint32_t xo)
{
	odiag_t *p;
	b->A = or_grow(b->A, &b->m_A, (size_t)(b->head_A + b->n_A + 1), sizeof(odiag_t));
	p = &b->A[b->head_A + b->n_A];
	p->vd = vd, p->k = k, p->xo = xo;
	++b->n_A;
}

/* static_A_shift, gwf-ed.cu:859-877 */
static odiag_t or_A_shift(obuf_t *b)
{
	odiag_t t = b->A[b->head_A];
	++b->head_A, --b->n_A;
	return t;
}

/* static_B_pushp, gwf-ed.cu:879-888 */
static void or_B_push(obuf_t *b, uint32_t v, int32_t d, int32_t k, uint32_t xo, uint32_t ooo)
{
	odiag_t *p;
	b->B = or_grow(b->B, &b->m_B, (size_t)b->n_B + 1, sizeof(odiag_t));
	p = &b->B[b->n_B++];
	p->vd = or_gen_vd(v, d), p->k = k, p->xo = xo << 1 | ooo;
}

/* gwf_diag_update, gwf-ed.cu:355-370 */
static int or_diag_update(odiag_t *p, uint32_t v, int32_t d, int32_t k, uint32_t x, uint32_t ooo)
{
	const uint64_t vd = or_gen_vd(v, d);
	if (p->vd == vd) {
		p->xo = p->k > k ? p->xo : x << 1 | ooo;
		p->k = p->k > k ? p->k : k;
		return 0;
	}
	return 1;
}

static void or_tmp_push(obuf_t *b, uint64_t vd)
{
	b->tmp0 = or_grow(b->tmp0, &b->m_tmp0, (size_t)b->n_tmp + 1, sizeof(uint64_t));
	b->tmp1 = or_grow(b->tmp1, &b->m_tmp1, (size_t)b->n_tmp + 1, sizeof(uint64_t));
	b->tmp0[b->n_tmp] = vd, b->tmp1[b->n_tmp] = vd + 1;
	++b->n_tmp;
}

/* static_set64_put, gwf-ed.cu:1020-1062, as an exact growable set.  Key 0 is
 * reported present, as in the reference (it is (v0, query position 0), the
 * start itself).  Not in the reference: the set also keeps the overlap of the
 * arc that claimed each key and counts refused claims that came with another
 * overlap (nlossy).  Only such a refusal drops a state that is not a copy of
 * one already reached, so a run with none must give the exact distance
 * (checked against the DP by the tests). */
static uint64_t or_hash64(uint64_t x)
{
	x ^= x >> 33; x *= 0xff51afd7ed558ccdULL;
	x ^= x >> 33; x *= 0xc4ceb9fe1a85ec53ULL;
	x ^= x >> 33;
	return x;
}
static int or_set_put(obuf_t *b, uint64_t key, int32_t ol)
{
	size_t i, mask;
	if (key == 0) { b->nlossy += ol != 0; return 0; }
	if ((b->ha_n + 1) * 2 > b->ha_cap) { /* grow, rehash */
		size_t nc = b->ha_cap ? b->ha_cap << 1 : 1024, j;
		uint64_t *nh = calloc(nc, sizeof(uint64_t));
		int32_t *no = calloc(nc, sizeof(int32_t));
		if (!nh || !no) abort();
		for (j = 0; j < b->ha_cap; ++j) {
			uint64_t x = b->ha[j];
			if (x == 0) continue;
			for (i = or_hash64(x) & (nc - 1); nh[i]; i = (i + 1) & (nc - 1)) {}
			nh[i] = x, no[i] = b->hol[j];
		}
		free(b->ha); free(b->hol);
		b->ha = nh, b->hol = no, b->ha_cap = nc;
	}
	mask = b->ha_cap - 1;
	for (i = or_hash64(key) & mask;; i = (i + 1) & mask) {
		if (b->ha[i] == key) { b->nlossy += b->hol[i] != ol; return 0; }
		if (b->ha[i] == 0) { b->ha[i] = key, b->hol[i] = ol; ++b->ha_n; return 1; }
	}
}

/* gwf_extend1, gwf-ed.cu:806-839, with a signed max_k (A1).  Counts each
 * character pair compared. */
static inline int32_t or_extend1(int32_t d, int32_t k, int32_t vl, const char *ts, int32_t ql,
                                 const char *qs, int64_t *nc)
{
	const int32_t max_k = (ql - d < vl ? ql - d : vl) - 1;
	const char *const ts_ = ts + 1;
	const char *const qs_ = qs + d + 1;
	while (k < max_k) {
		++*nc;
		if (ts_[k] != qs_[k]) break;
		++k;
	}
	return k;
}

/* gwf_ed_extend_batch, gwf-ed.cu:891-1018: extend a run of n consecutive
 * diagonals on one vertex and emit their n+2 successors. */
static void or_extend_batch(const gwf_graph_t *g, int32_t ql, const char *q, int32_t n, int64_t off,
                            obuf_t *b)
{
	const uint32_t v = (uint32_t)(b->a_vd[off] >> 32);
	const int32_t vl = (int32_t)g->len[v];
	const char *ts = g->seq[v];
	uint64_t *vd = b->a_vd + off;
	int32_t *ak = b->a_k + off;
	uint32_t *xo = b->a_xo + off;
	odiag_t *B;
	int32_t j, m;

	for (j = 0; j < n; ++j) { /* 902-910 */
		int32_t k = or_extend1(or_vd_d(vd[j]), ak[j], vl, ts, ql, q, &b->nc);
		xo[j] += (uint32_t)(k - ak[j]) << 2;
		ak[j] = k;
	}
	b->B = or_grow(b->B, &b->m_B, (size_t)(b->n_B + n + 2), sizeof(odiag_t));
	B = b->B + b->n_B;
	B[0].vd = vd[0] - 1; /* 917-920: deletion from the first diagonal */
	B[0].xo = xo[0] + 2;
	B[0].k = ak[0] + 1;
	B[1].vd = vd[0]; /* 924-928 */
	B[1].xo = n == 1 || ak[0] > ak[1] ? xo[0] + 4 : xo[1] + 2;
	B[1].k = (n == 1 || ak[0] > ak[1] ? ak[0] : ak[1]) + 1;
	for (j = 1; j < n - 1; ++j) { /* 935-951 */
		uint32_t x = xo[j - 1] + 2;
		int32_t k = ak[j - 1];
		x = k > ak[j] + 1 ? x : xo[j] + 4;
		k = k > ak[j] + 1 ? k : ak[j] + 1;
		x = k > ak[j + 1] + 1 ? x : xo[j + 1] + 2;
		k = k > ak[j + 1] + 1 ? k : ak[j + 1] + 1;
		B[j + 1].vd = vd[j], B[j + 1].k = k, B[j + 1].xo = x;
	}
	if (n >= 2) { /* 953-959 */
		B[n].vd = vd[n - 1];
		B[n].xo = ak[n - 2] > ak[n - 1] + 1 ? xo[n - 2] + 2 : xo[n - 1] + 4;
		B[n].k = ak[n - 2] > ak[n - 1] + 1 ? ak[n - 2] : ak[n - 1] + 1;
	}
	B[n + 1].vd = vd[n - 1] + 1; /* 960-963 */
	B[n + 1].xo = xo[n - 1] + 2;
	B[n + 1].k = ak[n - 1];

	for (j = 0; j < n; ++j) { /* 965-984: vertex end or query end -> A */
		if (ak[j] == vl - 1 || or_vd_d(vd[j]) + ak[j] == ql - 1) {
			xo[j] |= 1;
			or_A_push(b, vd[j], ak[j], xo[j]);
		}
	}
	for (j = 0, m = 0; j < n + 2; ++j) { /* 987-1016 */
		odiag_t *p = &B[j];
		int32_t d = or_vd_d(p->vd);
		if (d + p->k < ql && p->k < vl) B[m++] = *p;
		else if (p->k == vl) or_tmp_push(b, or_gen_vd(v, d));
	}
	b->n_B += m;
}

/* Stable sort of (key, val) pairs by key; the reference uses a stable
 * odd-even transposition or merge sort (gwf-ed.cu:85-216). */
static void or_sort_pairs(uint64_t *key, uint64_t *val, int64_t n, uint64_t *tk, uint64_t *tv)
{
	int64_t w, i;
	uint64_t *sk = key, *sv = val, *dk = tk, *dv = tv, *t;
	for (w = 1; w < n; w <<= 1) {
		for (i = 0; i < n; i += w << 1) {
			int64_t mid = i + w < n ? i + w : n, right = i + (w << 1) < n ? i + (w << 1) : n;
			int64_t x = i, y = mid, o = i;
			while (x < mid && y < right) {
				if (sk[x] <= sk[y]) dk[o] = sk[x], dv[o++] = sv[x++];
				else dk[o] = sk[y], dv[o++] = sv[y++];
			}
			while (x < mid) dk[o] = sk[x], dv[o++] = sv[x++];
			while (y < right) dk[o] = sk[y], dv[o++] = sv[y++];
		}
		t = sk, sk = dk, dk = t;
		t = sv, sv = dv, dv = t;
	}
	if (sk != key) { memcpy(key, sk, n * sizeof(uint64_t)); memcpy(val, sv, n * sizeof(uint64_t)); }
}

/* gwf_intv_merge_adj, gwf-ed.cu:266-296 */
static int64_t or_intv_merge_adj(int64_t n, uint64_t *vd0, uint64_t *vd1)
{
	int64_t i, k;
	uint64_t st, en;
	if (n == 0) return 0;
	st = vd0[0], en = vd1[0];
	for (i = 1, k = 0; i < n; ++i) {
		if (vd0[i] > en) {
			vd0[k] = st, vd1[k++] = en;
			st = vd0[i], en = vd1[i];
		} else en = en > vd1[i] ? en : vd1[i];
	}
	vd0[k] = st, vd1[k++] = en;
	return k;
}

/* gwf_dedup, gwf-ed.cu:780-804, with gwf_intv_merge2 (300-348),
 * gwf_diag_dedup (653-687), gwf_diag_sort (527-651) and gwf_mixed_dedup
 * (691-747) inlined in order. */
static int64_t or_dedup(obuf_t *b, int64_t n_a)
{
	int64_t i, j, n, st;
	if (b->n_intv + b->n_tmp > 0) {
		int sorted = 1;
		for (i = 1; i < b->n_tmp; ++i)
			if (b->tmp0[i - 1] > b->tmp0[i]) { sorted = 0; break; }
		if (!sorted) {
			size_t m0 = b->m_swap0, m1 = b->m_swap1;
			b->swap0 = or_grow(b->swap0, &m0, (size_t)b->n_tmp, sizeof(uint64_t)); b->m_swap0 = m0;
			b->swap1 = or_grow(b->swap1, &m1, (size_t)b->n_tmp, sizeof(uint64_t)); b->m_swap1 = m1;
			or_sort_pairs(b->tmp0, b->tmp1, b->n_tmp, b->swap0, b->swap1);
		}
		/* static_intv_copy, 751-778 (never dropped here) */
		b->swap0 = or_grow(b->swap0, &b->m_swap0, (size_t)b->n_intv + 1, sizeof(uint64_t));
		b->swap1 = or_grow(b->swap1, &b->m_swap1, (size_t)b->n_intv + 1, sizeof(uint64_t));
		memcpy(b->swap0, b->intv0, b->n_intv * sizeof(uint64_t));
		memcpy(b->swap1, b->intv1, b->n_intv * sizeof(uint64_t));
		b->n_swap = b->n_intv;
		/* gwf_intv_merge2: stable merge, swap before tmp on ties */
		{
			int64_t x = 0, y = 0, o = 0, tot = b->n_swap + b->n_tmp;
			b->intv0 = or_grow(b->intv0, &b->m_intv0, (size_t)tot + 1, sizeof(uint64_t));
			b->intv1 = or_grow(b->intv1, &b->m_intv1, (size_t)tot + 1, sizeof(uint64_t));
			while (x < b->n_swap || y < b->n_tmp) {
				if (x < b->n_swap && (y >= b->n_tmp || b->swap0[x] <= b->tmp0[y]))
					b->intv0[o] = b->swap0[x], b->intv1[o++] = b->swap1[x++];
				else
					b->intv0[o] = b->tmp0[y], b->intv1[o++] = b->tmp1[y++];
			}
			b->n_intv = or_intv_merge_adj(tot, b->intv0, b->intv1);
		}
	}

	/* gwf_diag_dedup: sort if needed */
	{
		int sorted = 1;
		for (i = 1; i < n_a; ++i)
			if (b->a_vd[i - 1] > b->a_vd[i]) { sorted = 0; break; }
		if (!sorted) { /* gwf_diag_sort: stable partition by ooo, sort c, merge */
			int64_t n_c = 0, n_bb, nb = 0, nc = 0;
			odiag_t *bb, *cc;
			for (i = 0; i < n_a; ++i) n_c += b->a_xo[i] & 1;
			n_bb = n_a - n_c;
			b->ooo = or_grow(b->ooo, &b->m_ooo, (size_t)n_a + 1, sizeof(odiag_t));
			bb = b->ooo, cc = bb + n_bb;
			for (i = 0; i < n_a; ++i) {
				odiag_t x;
				x.vd = b->a_vd[i], x.k = b->a_k[i], x.xo = b->a_xo[i];
				if (x.xo & 1) cc[nc++] = x;
				else bb[nb++] = x;
			}
			/* stable merge sort of c by vd (sort_gwf_diags, 503-524) */
			{
				int64_t w, s0;
				odiag_t *src = cc, *dst, *t;
				b->scr = or_grow(b->scr, &b->m_scr, (size_t)n_c + 1, sizeof(odiag_t));
				dst = b->scr;
				for (w = 1; w < n_c; w <<= 1) {
					for (s0 = 0; s0 < n_c; s0 += w << 1) {
						int64_t mid = s0 + w < n_c ? s0 + w : n_c;
						int64_t right = s0 + (w << 1) < n_c ? s0 + (w << 1) : n_c;
						int64_t x = s0, y = mid, o = s0;
						while (x < mid && y < right) dst[o++] = src[x].vd <= src[y].vd ? src[x++] : src[y++];
						while (x < mid) dst[o++] = src[x++];
						while (y < right) dst[o++] = src[y++];
					}
					t = src, src = dst, dst = t;
				}
				if (src != cc) memcpy(cc, src, n_c * sizeof(odiag_t));
			}
			for (i = 0; i < n_c; ++i) cc[i].xo &= 0xfffffffeU; /* 596-608 */
			{ /* merge path, b before c on ties (611-650) */
				int64_t x = 0, y = 0, o = 0;
				while (x < n_bb || y < n_c) {
					odiag_t best = (x < n_bb && (y >= n_c || bb[x].vd <= cc[y].vd)) ? bb[x++] : cc[y++];
					b->a_vd[o] = best.vd, b->a_k[o] = best.k, b->a_xo[o] = best.xo;
					++o;
				}
			}
		}
	}
	/* keep the first max-k entry of every vd (659-684) */
	for (i = 1, st = 0, n = 0; i <= n_a; ++i) {
		if (i == n_a || b->a_vd[i] != b->a_vd[st]) {
			int64_t max_j = st;
			for (j = st + 1; j < i; ++j)
				if (b->a_k[max_j] < b->a_k[j]) max_j = j;
			b->a_vd[n] = b->a_vd[max_j], b->a_k[n] = b->a_k[max_j], b->a_xo[n] = b->a_xo[max_j];
			++n;
			st = i;
		}
	}
	n_a = n;

	if (b->n_intv > 0) { /* gwf_mixed_dedup: drop diagonals inside forbidden intervals */
		int64_t w = 0;
		for (i = 0; i < n_a; ++i) {
			uint64_t x = b->a_vd[i];
			int keep = 1;
			int64_t lo = 0, hi = b->n_intv - 1;
			while (lo <= hi) {
				int64_t mid = (lo + hi) / 2;
				if (x >= b->intv0[mid] && x < b->intv1[mid]) { keep = 0; break; }
				if (x < b->intv0[mid]) hi = mid - 1;
				else lo = mid + 1;
			}
			if (keep) { b->a_vd[w] = x, b->a_k[w] = b->a_k[i], b->a_xo[w] = b->a_xo[i]; ++w; }
		}
		n_a = w;
	}
	return n_a;
}

/* gwf_ed_extend, gwf-ed.cu:1064-1290: one score step.  Returns 1 when an end
 * was found (end_v/end_off set), 0 otherwise; b->n_a is updated. */
static int or_ed_extend(obuf_t *b, const gwf_graph_t *g, int32_t ql, const char *q, int32_t v1,
                        int32_t *end_v, int32_t *end_off)
{
	int64_t i, x, n = b->n_a;
	int do_dedup = 1;

	b->n_A = 0, b->head_A = 0;
	b->n_B = 0;
	*end_v = *end_off = -1;
	b->n_tmp = 0;

	for (x = 0, i = 1; i <= n; ++i) { /* 1086-1096: runs of consecutive vd */
		if (i == n || b->a_vd[i] != b->a_vd[i - 1] + 1) {
			or_extend_batch(g, ql, q, (int32_t)(i - x), x, b);
			x = i;
		}
	}
	if (b->n_A == 0) do_dedup = 0; /* 1098-1099 */

	while (b->n_A > 0) { /* 1106-1255: A-queue closure */
		odiag_t t;
		uint32_t x0, ooo, v;
		int32_t d, k, ii, vl;

		t = or_A_shift(b);
		++b->na_events;
		ooo = t.xo & 1;
		v = (uint32_t)(t.vd >> 32);
		d = or_vd_d(t.vd);
		vl = (int32_t)g->len[v];
		k = or_extend1(d, t.k, vl, g->seq[v], ql, q, &b->nc);
		ii = k + d;
		x0 = (t.xo >> 1) + ((uint32_t)(k - t.k) << 1);

		if (k + 1 < vl && ii + 1 < ql) { /* 1141-1158: mid vertex, mid query */
			int push1 = 1, push2 = 1;
			if (b->n_B >= 2) push1 = or_diag_update(&b->B[b->n_B - 2], v, d - 1, k + 1, x0 + 1, ooo);
			if (b->n_B >= 1) push2 = or_diag_update(&b->B[b->n_B - 1], v, d, k + 1, x0 + 2, ooo);
			if (push1) or_B_push(b, v, d - 1, k + 1, x0 + 1, 1);
			if (push2 || push1) or_B_push(b, v, d, k + 1, x0 + 2, 1);
			or_B_push(b, v, d + 1, k, x0 + 1, ooo);
		} else if (ii + 1 < ql) { /* 1161-1227: vertex end, mid query */
			int32_t ov = (int32_t)(g->aux[v] >> 32);
			int32_t nv = (int32_t)g->aux[v];
			int32_t n_ext = 0, j;
			or_tmp_push(b, or_gen_vd(v, d));
			for (j = 0; j < nv; ++j) {
				uint32_t w = (uint32_t)g->arc[ov + j].a;
				int32_t ol = g->arc[ov + j].o;
				int absent = or_set_put(b, (uint64_t)w << 32 | (uint32_t)(ii + 1), ol);
				++b->ncross;
				if (q[ii + 1] == g->seq[w][ol]) {
					++n_ext;
					if (absent) or_A_push(b, or_gen_vd(w, ii + 1 - ol), ol, (x0 + 2) << 1 | 1);
				} else if (absent) {
					or_B_push(b, w, ii - ol, ol, x0 + 1, 1);
					or_B_push(b, w, ii + 1 - ol, ol, x0 + 2, 1);
				}
			}
			if (nv == 0 || n_ext != nv) or_B_push(b, v, d + 1, k, x0 + 1, 1);
		} else if (v1 < 0 || ((int32_t)v == v1 && k + 1 == vl)) { /* 1231-1236 */
			*end_v = (int32_t)v, *end_off = k;
			b->n_A = 0;
			return 1;
		} else if (k + 1 < vl) { /* 1237-1241 */
			or_B_push(b, v, d - 1, k + 1, x0 + 1, ooo);
		} else if ((int32_t)v != v1) { /* 1244-1254 */
			int32_t ov = (int32_t)(g->aux[v] >> 32), nv = (int32_t)g->aux[v], j;
			for (j = 0; j < nv; ++j) {
				uint32_t w = (uint32_t)g->arc[ov + j].a;
				int32_t ol = g->arc[ov + j].o;
				or_B_push(b, w, ii - ol, ol, x0 + 1, 1);
			}
		}
	}

	/* 1261-1276: a <- B */
	ob_reserve_a(b, (size_t)b->n_B + 1);
	for (i = 0; i < b->n_B; ++i)
		b->a_vd[i] = b->B[i].vd, b->a_k[i] = b->B[i].k, b->a_xo[i] = b->B[i].xo;
	b->n_a = b->n_B;
	b->n_B = 0;
	if (do_dedup) b->n_a = or_dedup(b, b->n_a); /* 1281-1282 */
	return 0;
}

/* gwf_ed, gwf-ed.cu:1292-1337 */
int32_t oracle_gwf_ed(const gwf_graph_t *g, int32_t ql, const char *q, int32_t v0, int32_t v1,
                      uint32_t max_lag, oracle_stats_t *st)
{
	obuf_t b;
	int32_t s = 0, end_v = -1, end_off = -1;
	int64_t steps = 0, states = 0, max_na = 0;
	(void)max_lag; /* the reference's prune is commented out (gwf-ed.cu:1286-1287) */
	if (!g || !q || ql < 0 || v0 < 0 || (uint32_t)v0 >= g->n_vtx || !g->aux) return -2;
	memset(&b, 0, sizeof(b));
	ob_reserve_a(&b, 16);
	b.a_vd[0] = or_gen_vd((uint32_t)v0, 0), b.a_k[0] = -1, b.a_xo[0] = 0;
	b.n_a = 1;
	while (b.n_a > 0) {
		int found;
		++steps, states += b.n_a;
		if (b.n_a > max_na) max_na = b.n_a;
		found = or_ed_extend(&b, g, ql, q, v1, &end_v, &end_off);
		if (found || b.n_a == 0) break; /* A8: the reference tests end_off >= 0 */
		++s;
	}
	if (st) {
		st->steps += steps, st->states += states, st->chars += b.nc;
		if (max_na > st->max_na) st->max_na = max_na;
		st->a_events += b.na_events, st->crossings += b.ncross, st->vis_lossy += b.nlossy;
	}
	ob_free(&b);
	return end_v >= 0 ? s : -1;
}

/* ---- brute-force DP checker (independent of the code above) ------------- */

int32_t oracle_dp_ed(const gwf_graph_t *g, int32_t ql, const char *q, int32_t v0, int32_t v1)
{
	/* Position y = (v, p) has index base[v] + p.  D[y][j] = min cost of
	 * aligning q[0..j) along a path from the start of v0 whose last graph base
	 * is y.  The virtual start S has D[S][j] = j.  Predecessors of (v, p):
	 * (v, p-1) when p > 0; (u, len(u)-1) for every arc u->v with overlap p;
	 * S when v == v0 and p == 0. */
	const int32_t INF = 0x3fffffff;
	uint32_t nv = g->n_vtx, v;
	int64_t *base, tot = 0, y;
	int32_t *prev, *cur, j, best;
	uint32_t *pred_n; int64_t *pred_off, *pred, np = 0;
	if (ql < 0 || v0 < 0 || (uint32_t)v0 >= nv) return -2;
	base = malloc((nv + 1) * sizeof(int64_t));
	for (v = 0; v < nv; ++v) base[v] = tot, tot += g->len[v];
	base[nv] = tot;
	/* cross-vertex predecessors: count, then fill (CSR over positions) */
	pred_n = calloc(tot + 1, sizeof(uint32_t));
	for (v = 0; v < g->n_arc; ++v) {
		uint32_t u = (uint32_t)(g->arc[v].a >> 32), w = (uint32_t)g->arc[v].a;
		int32_t o = g->arc[v].o;
		if (u >= nv || w >= nv || o < 0 || (uint32_t)o >= g->len[w] || g->len[u] == 0) continue;
		++pred_n[base[w] + o];
	}
	pred_off = malloc((tot + 1) * sizeof(int64_t));
	for (y = 0; y < tot; ++y) pred_off[y] = np, np += pred_n[y];
	pred_off[tot] = np;
	pred = malloc((np + 1) * sizeof(int64_t));
	memset(pred_n, 0, (tot + 1) * sizeof(uint32_t));
	for (v = 0; v < g->n_arc; ++v) {
		uint32_t u = (uint32_t)(g->arc[v].a >> 32), w = (uint32_t)g->arc[v].a;
		int32_t o = g->arc[v].o;
		if (u >= nv || w >= nv || o < 0 || (uint32_t)o >= g->len[w] || g->len[u] == 0) continue;
		y = base[w] + o;
		pred[pred_off[y] + pred_n[y]++] = base[u] + g->len[u] - 1;
	}
	prev = malloc((tot + 1) * sizeof(int32_t));
	cur = malloc((tot + 1) * sizeof(int32_t));
	for (j = 0; j <= ql; ++j) {
		int changed;
		char c = j > 0 ? q[j - 1] : 0;
		int32_t dS = j, dSprev = j - 1;
		/* insertion and (mis)match terms */
		for (v = 0; v < nv; ++v) {
			int64_t p;
			for (p = 0; p < (int64_t)g->len[v]; ++p) {
				int32_t bst = INF;
				y = base[v] + p;
				if (j > 0) {
					int32_t sub = g->seq[v][p] != c;
					int64_t t;
					if (prev[y] + 1 < bst) bst = prev[y] + 1;
					if (p > 0 && prev[y - 1] + sub < bst) bst = prev[y - 1] + sub;
					if (p == 0 && (int32_t)v == v0 && dSprev + sub < bst) bst = dSprev + sub;
					for (t = pred_off[y]; t < pred_off[y + 1]; ++t)
						if (prev[pred[t]] + sub < bst) bst = prev[pred[t]] + sub;
				}
				cur[y] = bst;
			}
		}
		/* deletion closure within the column (unit weights; iterate to a fixpoint) */
		do {
			changed = 0;
			for (v = 0; v < nv; ++v) {
				int64_t p;
				for (p = 0; p < (int64_t)g->len[v]; ++p) {
					int32_t bst = cur[base[v] + p];
					int64_t t;
					y = base[v] + p;
					if (p > 0 && cur[y - 1] + 1 < bst) bst = cur[y - 1] + 1;
					if (p == 0 && (int32_t)v == v0 && dS + 1 < bst) bst = dS + 1;
					for (t = pred_off[y]; t < pred_off[y + 1]; ++t)
						if (cur[pred[t]] + 1 < bst) bst = cur[pred[t]] + 1;
					if (bst < cur[y]) cur[y] = bst, changed = 1;
				}
			}
		} while (changed);
		{ int32_t *t = prev; prev = cur; cur = t; }
	}
	/* prev holds column ql */
	if (v1 < 0) {
		best = ql; /* all insertions from S */
		for (y = 0; y < tot; ++y) if (prev[y] < best) best = prev[y];
	} else {
		best = (uint32_t)v1 < nv && g->len[v1] > 0 ? prev[base[v1] + g->len[v1] - 1] : INF;
		if (best >= INF) best = -1;
	}
	free(base); free(pred_n); free(pred_off); free(pred); free(prev); free(cur);
	return best;
}

/* ---- graph index restatement (main.cu:246-270) -------------------------- */

static int or_arc_cmp(const void *x, const void *y)
{
	const gwf_arc_t *a = (const gwf_arc_t *)x, *b = (const gwf_arc_t *)y;
	return a->a < b->a ? -1 : a->a > b->a ? 1 : (a->o < b->o ? -1 : a->o > b->o);
}

void oracle_ed_index(gwf_graph_t *g, uint64_t *aux)
{
	uint32_t i, st;
	memset(aux, 0, g->n_vtx * sizeof(uint64_t));
	qsort(g->arc, g->n_arc, sizeof(gwf_arc_t), or_arc_cmp);
	for (st = 0, i = 1; i <= g->n_arc; ++i) {
		if (i == g->n_arc || g->arc[i].a >> 32 != g->arc[st].a >> 32) {
			uint32_t v = (uint32_t)(g->arc[st].a >> 32);
			if (v < g->n_vtx) aux[v] = (uint64_t)st << 32 | (i - st);
			st = i;
		}
	}
}

/* ---- multi-threaded batch runner (CPU baseline) ------------------------- */

typedef struct {
	const gwf_graph_t *g;
	int32_t n, v1;
	const char *seqs; const int64_t *off; const int32_t *ql, *v0;
	int32_t *out;
	int64_t next;
	pthread_mutex_t mu;
	oracle_stats_t st;
} or_pool_t;

static void *or_worker(void *arg)
{
	or_pool_t *p = (or_pool_t *)arg;
	oracle_stats_t st;
	memset(&st, 0, sizeof(st));
	for (;;) {
		int64_t i = __atomic_fetch_add(&p->next, 1, __ATOMIC_RELAXED);
		if (i >= p->n) break;
		p->out[i] = oracle_gwf_ed(p->g, p->ql[i], p->seqs + p->off[i], p->v0 ? p->v0[i] : 0, p->v1, 0, &st);
	}
	pthread_mutex_lock(&p->mu);
	p->st.steps += st.steps, p->st.states += st.states, p->st.chars += st.chars;
	if (st.max_na > p->st.max_na) p->st.max_na = st.max_na;
	p->st.a_events += st.a_events, p->st.crossings += st.crossings, p->st.vis_lossy += st.vis_lossy;
	pthread_mutex_unlock(&p->mu);
	return 0;
}

int oracle_batch(const gwf_graph_t *g, int32_t n, const char *seqs, const int64_t *off,
                 const int32_t *ql, const int32_t *v0, int32_t v1, int32_t *out,
                 int n_threads, oracle_stats_t *st)
{
	or_pool_t p;
	pthread_t *th;
	int t;
	if (n_threads <= 0) n_threads = (int)sysconf(_SC_NPROCESSORS_ONLN);
	if (n_threads < 1) n_threads = 1;
	memset(&p, 0, sizeof(p));
	p.g = g, p.n = n, p.v1 = v1, p.seqs = seqs, p.off = off, p.ql = ql, p.v0 = v0, p.out = out;
	pthread_mutex_init(&p.mu, 0);
	th = malloc(n_threads * sizeof(pthread_t));
	for (t = 0; t < n_threads; ++t) pthread_create(&th[t], 0, or_worker, &p);
	for (t = 0; t < n_threads; ++t) pthread_join(th[t], 0);
	free(th);
	pthread_mutex_destroy(&p.mu);
	if (st) {
		st->steps += p.st.steps, st->states += p.st.states, st->chars += p.st.chars;
		if (p.st.max_na > st->max_na) st->max_na = p.st.max_na;
		st->a_events += p.st.a_events, st->crossings += p.st.crossings, st->vis_lossy += p.st.vis_lossy;
	}
	return n_threads;
}
