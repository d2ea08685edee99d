// gwf_runtime.hip — the C-ABI of libgwfa_hip.so (include/gwfa.h): graph
// index, device-resident graph and reads, batched alignment with automatic
// re-runs of reads that outgrow their workspace, and the host-callable gwf_ed.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <functional>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <thread>
#include <string>
#include <unordered_map>
#include <vector>

#include "gwf_dev.h"
#include "gwf_policy.h"
#include "gwfa.h"

namespace {

thread_local std::string g_err;

void set_err(const char *fmt, ...) {
	char buf[512];
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(buf, sizeof(buf), fmt, ap);
	va_end(ap);
	g_err = buf;
}

#define HIPCHK(x)                                                                  \
	do {                                                                           \
		hipError_t e_ = (x);                                                       \
		if (e_ != hipSuccess) {                                                    \
			set_err("%s failed: %s", #x, hipGetErrorString(e_));                   \
			return GWF_ERR_HIP;                                                    \
		}                                                                          \
	} while (0)

constexpr i32 kProbeLen = 256;    // prefix aligned to estimate each read's error rate
constexpr int kPlainStreams = 16; // streams of the plain launches (as the reference's driver, main.cu:22)

int device_ready() {
	int n = 0;
	if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
		set_err("no HIP device visible");
		return GWF_ERR_NODEV;
	}
	return GWF_OK;
}

}  // namespace

uint64_t gwf_next_generation() {
	static std::atomic<uint64_t> next{1};
	return next.fetch_add(1, std::memory_order_relaxed);
}

namespace {

// Run context: stream, events and device buffers reused across gwf_ed_run
// calls (grown on demand), so steady-state runs allocate nothing.  Runs are
// serialised by mu.
struct RunCtx {
	std::mutex mu;
	int dev = -1;
	int wall_khz = 0;
	int n_cu = 0;
	hipStream_t s = nullptr;
	hipEvent_t e0 = nullptr, e1 = nullptr;
	i32 *ed = nullptr, *status = nullptr, *idx = nullptr, *head = nullptr;
	int64_t *cnt = nullptr;
	u64 *woff = nullptr;  // plain launches: per-read workspace offsets (elements) and sizes
	i32 *wa = nullptr;
	size_t n_cap = 0;
	u64 *scratch = nullptr;
	size_t scratch_bytes = 0;
	GwfBeat *beat = nullptr;  // host-coherent heartbeat records
	size_t beat_n = 0;
	// binned persistent runs
	i32 *pe = nullptr, *ps = nullptr, *key = nullptr, *ord = nullptr, *hist = nullptr, *boff = nullptr;
	size_t pe_c = 0, ps_c = 0, key_c = 0, ord_c = 0;
	hipStream_t sb[GWF_MAX_BINS] = {};
	hipEvent_t ev_fork = nullptr, ev_join[GWF_MAX_BINS] = {};
	u64 *scr[GWF_MAX_BINS] = {};
	size_t scr_c[GWF_MAX_BINS] = {};
	// graph of the last binned run, replayed while nothing it captured changes
	hipGraphExec_t gexec = nullptr;
	std::vector<uintptr_t> gkey;
	// plain launches: one workspace arena per stream
	hipStream_t pst[kPlainStreams] = {};
	hipEvent_t ev_pj[kPlainStreams] = {};
	u64 *pscr[kPlainStreams] = {};
	size_t pscr_c[kPlainStreams] = {};
};
RunCtx g_run;

template <class T>
hipError_t grow(T **p, size_t *cap, size_t need) {
	if (*p && *cap >= need) return hipSuccess;
	if (*p) (void)hipFree(*p);
	*p = nullptr;
	*cap = 0;
	hipError_t e = hipMalloc((void **)p, need);
	if (e == hipSuccess) *cap = need;
	return e;
}

int ctx_prepare(RunCtx &c, size_t n) {
	int dev = 0;
	HIPCHK(hipGetDevice(&dev));
	if (c.dev != dev) {  // first use (device buffers of another device are dropped)
		c.ed = c.status = c.idx = c.head = c.wa = nullptr;
		c.cnt = nullptr;
		c.woff = nullptr;
		c.scratch = nullptr;
		c.n_cap = c.scratch_bytes = 0;
		HIPCHK(hipStreamCreate(&c.s));
		HIPCHK(hipEventCreate(&c.e0));
		HIPCHK(hipEventCreate(&c.e1));
		HIPCHK(hipDeviceGetAttribute(&c.wall_khz, hipDeviceAttributeWallClockRate, dev));
		HIPCHK(hipDeviceGetAttribute(&c.n_cu, hipDeviceAttributeMultiprocessorCount, dev));
		HIPCHK(hipMalloc((void **)&c.head, 256));
		HIPCHK(hipMalloc((void **)&c.hist, sizeof(i32) * GWF_MAX_BINS * GWF_COST_BUCKETS));
		HIPCHK(hipMalloc((void **)&c.boff, 256));
		for (int b = 0; b < GWF_MAX_BINS; ++b) {
			HIPCHK(hipStreamCreate(&c.sb[b]));
			HIPCHK(hipEventCreateWithFlags(&c.ev_join[b], hipEventDisableTiming));
		}
		HIPCHK(hipEventCreateWithFlags(&c.ev_fork, hipEventDisableTiming));
		for (int k = 0; k < kPlainStreams; ++k) {
			HIPCHK(hipStreamCreate(&c.pst[k]));
			HIPCHK(hipEventCreateWithFlags(&c.ev_pj[k], hipEventDisableTiming));
			c.pscr[k] = nullptr, c.pscr_c[k] = 0;
		}
		HIPCHK(gwf_kernels_init());
		c.pe = c.ps = c.key = c.ord = nullptr;
		c.pe_c = c.ps_c = c.key_c = c.ord_c = 0;
		for (int b = 0; b < GWF_MAX_BINS; ++b) c.scr[b] = nullptr, c.scr_c[b] = 0;
		c.gexec = nullptr;
		c.gkey.clear();
		c.dev = dev;
	}
	if (c.n_cap < n) {
		const size_t m = std::max<size_t>(n, 1024);
		size_t c0 = c.n_cap * 4, c1 = c.n_cap * 4, c2 = c.n_cap * 4, c3 = c.n_cap * 8 * GWF_NCOUNT;
		size_t c4 = c.n_cap * 8, c5 = c.n_cap * 4;
		c.n_cap = 0;
		HIPCHK(grow(&c.ed, &c0, 4 * m));
		HIPCHK(grow(&c.status, &c1, 4 * m));
		HIPCHK(grow(&c.idx, &c2, 4 * m));
		HIPCHK(grow(&c.cnt, &c3, 8 * GWF_NCOUNT * m));
		HIPCHK(grow(&c.woff, &c4, 8 * m));
		HIPCHK(grow(&c.wa, &c5, 4 * m));
		c.n_cap = m;
	}
	return GWF_OK;
}

// device bytes of the workspaces the context holds (reusable by the next run)
size_t ctx_held(const RunCtx &c) {
	size_t h = c.scratch_bytes;
	for (int b = 0; b < GWF_MAX_BINS; ++b) h += c.scr_c[b];
	for (int k = 0; k < kPlainStreams; ++k) h += c.pscr_c[k];
	return h;
}

int64_t env_int(const char *name, int64_t dflt) {
	const char *v = getenv(name);
	return v && *v ? atoll(v) : dflt;
}

// HBM workspace of a read of length ql: the reference's per-read sizing
// (main.cu:498-504) plus headroom, times `scale` after overflows.
// GWF_DEBUG_CAP_A overrides the base (tests force the re-run path with it).
// mixed: the graph has vertices whose in-arcs carry different overlaps (the
// workspace then holds the read's visited set too).
Caps caps_for(i32 ql, int scale, bool mixed) {
	const int64_t dbg = env_int("GWF_DEBUG_CAP_A", 0);
	int64_t a = dbg > 0 ? dbg : (int64_t)(1.6 * ql) + 256;
	a *= scale;
	if (a > (1 << 25)) a = 1 << 25;
	Caps c;
	c.m = (i32)(2 * a);
	c.x = (i32)a;
	c.f = (i32)a;
	c.t = (i32)a;
	c.h = (i32)a;
	c.v = gwf_vis_slots((i32)a, mixed);
	return c;
}

// LDS capacities for a wavefront of up to ~m states.
Caps lds_caps(i32 m) {
	Caps c;
	c.m = m;
	c.x = std::max(64, (m / 2 + 63) / 64 * 64);
	c.f = std::max(64, m / 4);
	c.t = std::max(64, m / 2);
	c.h = std::max(64, m / 4);
	c.v = 0;  // the per-read visited set, when there is one, stays in HBM
	return c;
}

// Launch parameters for waves-per-read nw, LDS wavefront capacity m and
// sort variant mw (GWF_PERSIST_NW / _M / _WGS override, for tests).
PersistCfg make_cfg(const RunCtx &cx, i32 nw, i32 m, i32 mw, i32 max_ql, bool mixed, bool env_override = true) {
	PersistCfg c;
	if (env_override) {
		nw = (i32)env_int("GWF_PERSIST_NW", nw);
		m = (i32)env_int("GWF_PERSIST_M", m);
	}
	c.nw = nw;
	c.mw = mw;
	c.lds = lds_caps(m);
	c.lds_bytes = gwf_lds_bytes(c.lds);
	c.hbm = caps_for(max_ql, 1, mixed);
	const int per_cu = gwf_persist_occupancy(nw, c.lds_bytes);
	// LDS left over at this occupancy becomes staging room for spilled reads.
	// The occupancy query accepts sizes at which MI355X keeps one workgroup
	// per CU out (2 waves x 10 per CU: 13.3 KB dynamic measured 9 resident,
	// 12 KB 10), hence the 4 KB margin on each workgroup's share.
	c.lds_bytes = gwf_persist_grow_lds(nw, c.lds_bytes, per_cu, (size_t)env_int("GWF_LDS_MARGIN", 4096));
	c.wgs = std::max(1, per_cu) * std::max(1, cx.n_cu);
	if (env_override) c.wgs = (i32)env_int("GWF_PERSIST_WGS", c.wgs);
	return c;
}

// The tuning table: per bin of predicted wavefront size, waves per read,
// LDS capacity and sort variant.  Built-in values are measured on MI355X
// (gwf_autotune replaces them for a workload).
std::mutex g_tune_mu;
gwf_tune_t g_tune;
u64 g_tune_version = 1;

gwf_tune_t builtin_tune() {
	gwf_tune_t t;
	memset(&t, 0, sizeof(t));
	const i32 est[] = {160, 480, 1200, 3000, 1 << 30};
	const i32 nw[] = {1, 2, 4, 4, 4};
	const i32 m[] = {256, 512, 1024, 1024, 1024};
	t.n_bins = 5;
	for (int b = 0; b < 5; ++b) t.bin[b] = {est[b], nw[b], m[b], 0};
	return t;
}

int tune_check(const gwf_tune_t &t) {
	if (t.n_bins < 1 || t.n_bins > GWF_MAX_BINS) return GWF_ERR_ARG;
	for (int b = 0; b < t.n_bins; ++b) {
		const gwf_tune_bin_t &x = t.bin[b];
		if (x.nw != 1 && x.nw != 2 && x.nw != 4) return GWF_ERR_ARG;
		if (x.lds_states < 64 || gwf_lds_bytes(lds_caps(x.lds_states)) > gwf_max_lds_bytes()) return GWF_ERR_ARG;
		if (x.merge_runs < 0 || x.merge_runs > 64) return GWF_ERR_ARG;
		if (b > 0 && x.est_max <= t.bin[b - 1].est_max) return GWF_ERR_ARG;
	}
	return GWF_OK;
}

// Text form, one bin per line:
//   # comment
//   bin <est_max> <nw> <lds_states> <merge_runs>
int gwf_tune_read_file(const char *path, gwf_tune_t *out) {
	FILE *f = fopen(path, "r");
	if (!f) return GWF_ERR_ARG;
	gwf_tune_t t;
	memset(&t, 0, sizeof(t));
	char line[256];
	int rc = GWF_OK;
	while (fgets(line, sizeof line, f)) {
		if (line[0] == '#' || line[0] == '\n') continue;
		gwf_tune_bin_t x;
		if (sscanf(line, "bin %d %d %d %d", &x.est_max, &x.nw, &x.lds_states, &x.merge_runs) != 4 ||
		    t.n_bins >= GWF_MAX_BINS) {
			rc = GWF_ERR_ARG;
			break;
		}
		t.bin[t.n_bins++] = x;
	}
	fclose(f);
	if (rc == GWF_OK) rc = tune_check(t);
	if (rc == GWF_OK) *out = t;
	return rc;
}

gwf_tune_t tune_current(u64 *version) {
	std::lock_guard<std::mutex> lk(g_tune_mu);
	if (g_tune.n_bins == 0) {
		g_tune = builtin_tune();
		if (const char *f = getenv("GWF_TUNE_FILE")) {
			gwf_tune_t t;
			if (gwf_tune_read_file(f, &t) == GWF_OK && tune_check(t) == GWF_OK) g_tune = t;
		}
	}
	if (version) *version = g_tune_version;
	return g_tune;
}

// gwf_ed's device copies, one per graph pointer.  An entry is dropped by
// gwf_ed_index and gwf_cleanup on that pointer (re-indexing rebuilds arc and
// aux, so the upload no longer describes the graph), and is also replaced
// when the graph's counts or array pointers no longer match the ones it was
// uploaded from.
struct CachedGraph {
	gwf_dgraph_t *dg = nullptr;
	u32 n_vtx = 0, n_arc = 0;
	const void *len = nullptr, *seq = nullptr, *arc = nullptr, *aux = nullptr;
	bool matches(const gwf_graph_t *g) const {
		return n_vtx == g->n_vtx && n_arc == g->n_arc && len == g->len && seq == g->seq && arc == g->arc &&
		       aux == g->aux;
	}
};
std::mutex g_cache_mu;
std::unordered_map<const gwf_graph_t *, CachedGraph> g_cache;

void cache_drop(const gwf_graph_t *g) {
	std::lock_guard<std::mutex> lk(g_cache_mu);
	auto it = g_cache.find(g);
	if (it != g_cache.end()) {
		gwf_dgraph_free(it->second.dg);
		g_cache.erase(it);
	}
}

}  // namespace

int64_t gwf_env_int(const char *name, int64_t dflt) { return env_int(name, dflt); }
Caps gwf_caps_for(i32 ql, int scale, bool mixed) { return caps_for(ql, scale, mixed); }
Caps gwf_lds_caps(i32 m) { return lds_caps(m); }
void gwf_set_error(const char *fmt, ...) {
	char buf[512];
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(buf, sizeof(buf), fmt, ap);
	va_end(ap);
	g_err = buf;
}

extern "C" {

const char *gwf_last_error(void) { return g_err.c_str(); }

// main.cu:246-270: sort arcs by a = src << 32 | dst, aux[v] = first << 32 | count.
void gwf_ed_index(gwf_graph_t *g) {
	if (!g) return;
	cache_drop(g);
	free(g->aux);
	g->aux = (uint64_t *)calloc(g->n_vtx ? g->n_vtx : 1, sizeof(uint64_t));
	if (!g->aux || g->n_arc == 0) return;
	std::stable_sort(g->arc, g->arc + g->n_arc, [](const gwf_arc_t &x, const gwf_arc_t &y) { return x.a < y.a; });
	uint32_t st = 0;
	for (uint32_t i = 1; i <= g->n_arc; ++i) {
		if (i == g->n_arc || g->arc[i].a >> 32 != g->arc[st].a >> 32) {
			uint32_t v = (uint32_t)(g->arc[st].a >> 32);
			if (v < g->n_vtx) g->aux[v] = (uint64_t)st << 32 | (i - st);
			st = i;
		}
	}
}

void gwf_cleanup(gwf_graph_t *g) {
	if (!g) return;
	cache_drop(g);
	free(g->aux);
	g->aux = nullptr;
}

int gwf_dgraph_upload(const gwf_graph_t *g, gwf_dgraph_t **out) {
	if (!g || !out) { set_err("gwf_dgraph_upload: NULL argument"); return GWF_ERR_ARG; }
	*out = nullptr;
	if (!g->aux && g->n_arc) { set_err("gwf_dgraph_upload: graph not indexed (call gwf_ed_index)"); return GWF_ERR_ARG; }
	if (g->n_vtx == 0 || g->n_vtx > gwf_layout_max_vtx(3)) { set_err("gwf_dgraph_upload: n_vtx %u out of range", g->n_vtx); return GWF_ERR_ARG; }
	int rc = device_ready();
	if (rc) return rc;
	const u32 nv = g->n_vtx, na = g->n_arc;
	std::vector<uint4> vinfo(2 * (size_t)nv);
	std::vector<i32> in_ol(nv, -1);
	bool mixed = false;
	uint64_t total = 0;
	u32 max_vlen = 0;
	for (u32 v = 0; v < nv; ++v) {
		if (g->len[v] == 0 || g->len[v] > (u32)GWF_MAX_VLEN) {
			set_err("gwf_dgraph_upload: vertex %u has length %u (supported: 1..%d)", v, g->len[v], GWF_MAX_VLEN);
			return GWF_ERR_ARG;
		}
		total += g->len[v];
		max_vlen = std::max(max_vlen, g->len[v]);
	}
	if (gwf_pick_layout(nv, max_vlen, 0) < 0) {
		set_err("gwf_dgraph_upload: no state layout holds %u vertices with vertex length %u", nv, max_vlen);
		return GWF_ERR_ARG;
	}
	if (total + kSeqPad >= (1ull << 32)) { set_err("gwf_dgraph_upload: graph sequence exceeds 4 GiB"); return GWF_ERR_ARG; }
	std::vector<uint4> arcs(na ? na : 1, make_uint4(0, 0, 0, 0));
	std::vector<u32> seq_off(nv);
	{
		uint64_t o = 0;
		for (u32 v = 0; v < nv; ++v) seq_off[v] = (u32)o, o += g->len[v];
	}
	// in-arc counts: a target with one in-arc is flagged in its arc record
	std::vector<u32> indeg(nv, 0);
	for (u32 i = 0; i < na; ++i) {
		const u32 w = (u32)g->arc[i].a;
		if (w < nv) ++indeg[w];
	}
	for (u32 i = 0; i < na; ++i) {
		const u32 u = (u32)(g->arc[i].a >> 32), w = (u32)g->arc[i].a;
		const i32 o = g->arc[i].o;
		if (u >= nv || w >= nv || o < 0 || (u32)o >= g->len[w]) {
			set_err("gwf_dgraph_upload: arc %u (%u->%u, overlap %d) is invalid", i, u, w, o);
			return GWF_ERR_ARG;
		}
		// The visited set keys a vertex entry by (target, query position), as
		// gwf-ed.cu:1196 does.  With one overlap per target, every entry under
		// a key is the same state; with several, the reference keeps the entry
		// its FIFO closure reaches first, so such graphs are aligned with the
		// closure in that order and one visited set per read (DGraph::mixed).
		if (in_ol[w] >= 0 && in_ol[w] != o) mixed = true;
		in_ol[w] = o;
		arcs[i] = make_uint4(w, (u32)o | (indeg[w] == 1 ? GWF_ARC_SINGLE_IN : 0u) | (u32)(uint8_t)g->seq[w][o] << 24,
		                     seq_off[w], g->len[w]);
	}
	uint64_t so = 0;
	for (u32 v = 0; v < nv; ++v) {
		const uint64_t ax = g->aux ? g->aux[v] : 0;
		const u32 a0 = (u32)(ax >> 32), deg = (u32)ax;
		vinfo[2 * (size_t)v] = make_uint4((u32)so, g->len[v], a0, deg);
		vinfo[2 * (size_t)v + 1] = make_uint4(deg > 0 ? arcs[a0].x : 0, deg > 0 ? arcs[a0].y : 0,
		                                      deg > 1 ? arcs[a0 + 1].x : 0, deg > 1 ? arcs[a0 + 1].y : 0);
		so += g->len[v];
	}
	const size_t b_vi = sizeof(uint4) * vinfo.size(), b_seq = (size_t)total + kSeqPad, b_arc = sizeof(uint4) * arcs.size();
	auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
	const size_t bytes = al(b_vi) + al(b_seq) + al(b_arc);
	std::vector<uint8_t> seqh(b_seq, 0);
	so = 0;
	for (u32 v = 0; v < nv; ++v) {
		memcpy(&seqh[so], g->seq[v], g->len[v]);
		so += g->len[v];
	}
	gwf_dgraph_t *dg = new gwf_dgraph_t();
	if (hipMalloc(&dg->mem, bytes) != hipSuccess) {
		delete dg;
		set_err("gwf_dgraph_upload: hipMalloc(%zu) failed", bytes);
		return GWF_ERR_NOMEM;
	}
	uint8_t *p = (uint8_t *)dg->mem;
	dg->g.vinfo = (const uint4 *)p;
	hipError_t e = hipMemcpy(p, vinfo.data(), b_vi, hipMemcpyHostToDevice);
	p += al(b_vi);
	dg->g.seq = p;
	if (e == hipSuccess) e = hipMemcpy(p, seqh.data(), b_seq, hipMemcpyHostToDevice);
	p += al(b_seq);
	dg->g.arc = (const uint4 *)p;
	if (e == hipSuccess) e = hipMemcpy(p, arcs.data(), b_arc, hipMemcpyHostToDevice);
	dg->g.n_vtx = nv;
	dg->g.total_len = (u32)total;
	dg->g.mixed = mixed ? 1u : 0u;
	dg->g.layout = 0;
	dg->max_vlen = max_vlen;
	dg->n_vtx = nv;
	dg->gen = gwf_next_generation();
	if (e != hipSuccess) {
		set_err("gwf_dgraph_upload: hipMemcpy failed: %s", hipGetErrorString(e));
		gwf_dgraph_free(dg);
		return GWF_ERR_HIP;
	}
	*out = dg;
	return GWF_OK;
}

void gwf_dgraph_free(gwf_dgraph_t *dg) {
	if (!dg) return;
	if (dg->mem) (void)hipFree(dg->mem);
	delete dg;
}

int gwf_dreads_upload(int32_t n, const char *seqs, const int64_t *off, const int32_t *ql, const int32_t *v0,
                      gwf_dreads_t **out) {
	if (!out || n < 0 || (n > 0 && (!seqs || !off || !ql))) { set_err("gwf_dreads_upload: bad argument"); return GWF_ERR_ARG; }
	*out = nullptr;
	int rc = device_ready();
	if (rc) return rc;
	gwf_dreads_t *dr = new gwf_dreads_t();
	dr->ql.resize(n), dr->v0.resize(n);
	std::vector<int64_t> noff(n ? n : 1);
	int64_t total = 0;
	for (i32 i = 0; i < n; ++i) {
		if (ql[i] < 0 || ql[i] > GWF_MAX_QLEN) {
			delete dr;
			set_err("gwf_dreads_upload: read %d has length %d (supported: 0..%d)", i, ql[i], GWF_MAX_QLEN);
			return GWF_ERR_ARG;
		}
		dr->ql[i] = ql[i];
		dr->v0[i] = v0 ? v0[i] : 0;
		if (dr->v0[i] < 0) { delete dr; set_err("gwf_dreads_upload: read %d has v0 %d", i, dr->v0[i]); return GWF_ERR_ARG; }
		noff[i] = total;
		total += (ql[i] + 8) & ~7;  // every read starts 8-byte aligned (the extension's loads need it)
	}
	std::vector<uint8_t> buf((size_t)total + kSeqPad, 0);
	for (i32 i = 0; i < n; ++i) memcpy(&buf[noff[i]], seqs + off[i], ql[i]);
	auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
	const size_t b_seq = buf.size(), b_off = 8 * noff.size(), b_ql = 4 * (size_t)(n ? n : 1);
	const size_t bytes = al(b_seq) + al(b_off) + 3 * al(b_ql);
	if (hipMalloc(&dr->mem, bytes) != hipSuccess) {
		delete dr;
		set_err("gwf_dreads_upload: hipMalloc(%zu) failed", bytes);
		return GWF_ERR_NOMEM;
	}
	uint8_t *p = (uint8_t *)dr->mem;
	hipError_t e = hipMemcpy(p, buf.data(), b_seq, hipMemcpyHostToDevice);
	dr->r.seq = p;
	p += al(b_seq);
	if (e == hipSuccess) e = hipMemcpy(p, noff.data(), b_off, hipMemcpyHostToDevice);
	dr->r.off = (const int64_t *)p;
	p += al(b_off);
	if (e == hipSuccess && n) e = hipMemcpy(p, dr->ql.data(), 4 * (size_t)n, hipMemcpyHostToDevice);
	dr->r.ql = (const i32 *)p;
	p += al(b_ql);
	if (e == hipSuccess && n) e = hipMemcpy(p, dr->v0.data(), 4 * (size_t)n, hipMemcpyHostToDevice);
	dr->r.v0 = (const i32 *)p;
	p += al(b_ql);
	{
		std::vector<i32> qp(n ? n : 1);
		for (i32 i = 0; i < n; ++i) qp[i] = std::min(dr->ql[i], kProbeLen);
		if (e == hipSuccess && n) e = hipMemcpy(p, qp.data(), 4 * (size_t)n, hipMemcpyHostToDevice);
	}
	dr->ql_probe = (const i32 *)p;
	dr->r.n = n;
	dr->r.v1 = -1;
	dr->gen = gwf_next_generation();
	if (e != hipSuccess) {
		set_err("gwf_dreads_upload: hipMemcpy failed: %s", hipGetErrorString(e));
		gwf_dreads_free(dr);
		return GWF_ERR_HIP;
	}
	*out = dr;
	return GWF_OK;
}

uint64_t gwf_dgraph_generation(const gwf_dgraph_t *dg) { return dg ? dg->gen : 0; }
uint64_t gwf_dreads_generation(const gwf_dreads_t *dr) { return dr ? dr->gen : 0; }

void gwf_dreads_free(gwf_dreads_t *dr) {
	if (!dr) return;
	if (dr->mem) (void)hipFree(dr->mem);
	delete dr;
}

extern "C++" int gwf_run_graph(const gwf_dgraph_t *dg, const gwf_dreads_t *dr, const char *what, DGraph *out) {
	i32 maxql = 0;
	for (i32 q : dr->ql) maxql = std::max(maxql, q);
	int l = gwf_pick_layout(dg->n_vtx, dg->max_vlen, maxql);
	// GWF_KEY_LAYOUT=n: layout n instead of the first that fits (tests)
	const int64_t force = env_int("GWF_KEY_LAYOUT", -1);
	if (force >= 0) {
		l = (int)force;
		if (force >= GWF_N_LAYOUTS || dg->n_vtx > gwf_layout_max_vtx(l) || dg->max_vlen > gwf_layout_max_vlen(l) ||
		    maxql > gwf_layout_max_ql(l)) {
			set_err("%s: state layout %lld (GWF_KEY_LAYOUT) does not hold %u vertices, vertex length %u and read "
			        "length %d", what, (long long)force, dg->n_vtx, dg->max_vlen, maxql);
			return GWF_ERR_ARG;
		}
	}
	if (l < 0) {
		set_err("%s: no state layout holds %u vertices, vertex length %u and read length %d together", what, dg->n_vtx,
		        dg->max_vlen, maxql);
		return GWF_ERR_ARG;
	}
	*out = dg->g;
	out->layout = (u32)l;
	return GWF_OK;
}

namespace {

struct RunAcc {  // statistics of one gwf_ed_run call
	gwf_run_stats_t st;
	hipError_t err = hipSuccess;
	const char *what = "";
};

#define ACC(x)                                              \
	do {                                                    \
		hipError_t e_ = (x);                                \
		if (e_ != hipSuccess) {                             \
			acc.err = e_;                                   \
			acc.what = #x;                                  \
			return false;                                   \
		}                                                   \
	} while (0)

// Heartbeat (GWF_HEARTBEAT_MS > 0): while a launch runs, print one line per
// period to stderr from the workgroups' progress records (main.cu:48-149's
// progress slots, read from host-coherent memory instead of a debug build).
struct Heartbeat {
	GwfBeat *beat = nullptr;
	i32 n = 0;
	int period_ms = 0;
	const char *what = "";
};

void heartbeat_wait(RunCtx &cx, const Heartbeat &hb) {
	const auto t0 = std::chrono::steady_clock::now();
	auto last = t0;
	std::vector<u64> prev(hb.n, 0);
	std::vector<int> still(hb.n, 0);
	while (hipEventQuery(cx.e1) == hipErrorNotReady) {
		const auto now = std::chrono::steady_clock::now();
		if (std::chrono::duration<double, std::milli>(now - last).count() >= hb.period_ms) {
			i32 busy = 0, hbm = 0, stuck = 0, max_step = 0, max_na = 0;
			for (i32 w = 0; w < hb.n; ++w) {
				const volatile GwfBeat &b = hb.beat[w];
				if (b.read < 0) continue;
				++busy;
				hbm += b.in_hbm != 0;
				max_step = std::max(max_step, (i32)b.step);
				max_na = std::max(max_na, (i32)b.n_a);
				still[w] = b.beat == prev[w] ? still[w] + 1 : 0;
				stuck += still[w] >= 5;
				prev[w] = b.beat;
			}
			fprintf(stderr, "[gwf heartbeat] %s t=%.1fs workgroups busy=%d/%d in_hbm=%d max_step=%d max_n_a=%d silent>=5=%d\n",
			        hb.what, std::chrono::duration<double>(now - t0).count(), busy, hb.n, hbm, max_step, max_na, stuck);
			fflush(stderr);
			last = now;
		}
		std::this_thread::sleep_for(std::chrono::milliseconds(1));
	}
}

// launches: alignment kernels the call starts (0 for a graph whose kernels
// were counted when it was planned)
bool timed_launch(RunCtx &cx, RunAcc &acc, const std::function<hipError_t()> &launch, const Heartbeat *hb = nullptr,
                  int launches = 1) {
	ACC(hipEventRecord(cx.e0, cx.s));
	ACC(launch());
	ACC(hipEventRecord(cx.e1, cx.s));
	if (hb && hb->beat) heartbeat_wait(cx, *hb);
	ACC(hipEventSynchronize(cx.e1));
	float ms = 0;
	ACC(hipEventElapsedTime(&ms, cx.e0, cx.e1));
	acc.st.kernel_ms += ms;
	acc.st.n_launch += launches;
	return true;
}

// Plain kernel over `todo`, in the reference driver's launch shape
// (main.cu:498-735): every read gets a workspace of its own length's size
// (main.cu:498-520 sizes scratch per read), reads are packed in order into
// batches, and each batch is one launch of one wavefront per read.  When all
// workspaces fit the budget that is a single launch; otherwise each of
// kPlainStreams streams owns an arena of budget / kPlainStreams bytes, the
// batches go round-robin over the streams, and a stream's next batch reuses
// its arena after its previous launch (stream order), so launches on
// different streams overlap.  A read whose workspace alone exceeds an arena
// gets a batch (and an arena) of its own.
bool run_plain(RunCtx &cx, RunAcc &acc, const DGraph &g, const DReads &r, const std::vector<i32> &qlh,
               const std::vector<i32> &todo, int scale, size_t budget) {
	struct Batch {
		size_t pos;
		i32 nb;
		int stream;
	};
	const size_t nt = todo.size();
	std::vector<u64> off(nt);
	std::vector<i32> wa(nt);
	std::vector<size_t> wsb(nt);
	size_t total = 0;
	for (size_t j = 0; j < nt; ++j) {
		const Caps c = caps_for(qlh[todo[j]], scale, g.mixed != 0);
		wa[j] = c.x;
		wsb[j] = gwf_ws_bytes(c);
		total += wsb[j];
	}
	const size_t arena = total <= budget ? total : std::max<size_t>(budget / kPlainStreams, 1);
	std::vector<Batch> bt;
	size_t need[kPlainStreams] = {};
	for (size_t pos = 0; pos < nt;) {
		size_t end = pos, used = 0;
		while (end < nt && (end == pos || used + wsb[end] <= arena)) {
			off[end] = used / 8;
			used += wsb[end++];
		}
		const int k = (int)(bt.size() % kPlainStreams);
		bt.push_back({pos, (i32)(end - pos), k});
		need[k] = std::max(need[k], used);
		pos = end;
	}
	for (int k = 0; k < kPlainStreams; ++k)
		if (need[k]) ACC(grow(&cx.pscr[k], &cx.pscr_c[k], need[k]));
	ACC(hipMemcpyAsync(cx.idx, todo.data(), 4 * nt, hipMemcpyHostToDevice, cx.s));
	ACC(hipMemcpyAsync(cx.woff, off.data(), 8 * nt, hipMemcpyHostToDevice, cx.s));
	ACC(hipMemcpyAsync(cx.wa, wa.data(), 4 * nt, hipMemcpyHostToDevice, cx.s));
	const int used = (int)std::min<size_t>(bt.size(), kPlainStreams);
	return timed_launch(cx, acc, [&]() -> hipError_t {
		hipError_t e = hipEventRecord(cx.ev_fork, cx.s);
		for (int k = 0; k < used && e == hipSuccess; ++k) e = hipStreamWaitEvent(cx.pst[k], cx.ev_fork, 0);
		for (const Batch &x : bt) {
			if (e != hipSuccess) break;
			e = gwf_launch_plain(g, r, cx.idx + x.pos, x.nb, cx.woff + x.pos, cx.wa + x.pos, cx.pscr[x.stream],
			                     cx.ed, cx.status, cx.cnt, cx.pst[x.stream]);
		}
		for (int k = 0; k < used && e == hipSuccess; ++k) {
			e = hipEventRecord(cx.ev_pj[k], cx.pst[k]);
			if (e == hipSuccess) e = hipStreamWaitEvent(cx.s, cx.ev_pj[k], 0);
		}
		return e;
	}, nullptr, (int)bt.size());
}

// The binned run, all on the device: probe every read's prefix, order the
// reads into per-bin longest-first queues, then one persistent launch per
// bin with that bin's tuned configuration (on forked streams when their
// workspaces fit the budget, so each bin's tail overlaps the next bin).
struct BinnedPlan {
	DGraph g;  // the graph, with the run's key layout
	PersistCfg probe;
	PersistCfg bin[GWF_MAX_BINS];
	BinBounds bb;
	i32 n = 0;
	i32 v1 = -1;
	bool concurrent = false;
	// one bin and no more reads than resident workgroups: every read starts at
	// launch whatever the order, so the probe and the binning are skipped
	bool direct = false;
	GwfBeat *beat = nullptr;  // heartbeat records: beat_stride per bin (probe: bin n_bins)
	i32 beat_stride = 0;
};

// The longest read length (at most maxql) whose HBM workspace, times wgs
// workgroups, fits the budget (at least 1: the workspace of a read of length 1).
i32 bin_ws_ql(i32 maxql, i32 wgs, size_t budget, bool mixed) {
	auto fits = [&](i32 q) { return gwf_ws_bytes(caps_for(q, 1, mixed)) * (size_t)std::max(wgs, 1) <= budget; };
	if (fits(maxql)) return maxql;
	i32 lo = 1, hi = maxql;  // fits(lo) or lo == 1; !fits(hi)
	while (hi - lo > 1) {
		const i32 mid = lo + (hi - lo) / 2;
		if (fits(mid)) lo = mid;
		else hi = mid;
	}
	return lo;
}

hipError_t enqueue_binned(RunCtx &cx, const gwf_dgraph_t *dg, const gwf_dreads_t *dr, const BinnedPlan &p) {
	hipError_t e;
#define ENQ(x)                            \
	do {                                  \
		if ((e = (x)) != hipSuccess) return e; \
	} while (0)
	DReads rp = dr->r, rr = dr->r;
	rp.ql = dr->ql_probe;
	rp.v1 = -1;  // the probe only estimates the error rate
	rr.v1 = p.v1;
	ENQ(gwf_launch_fill(cx.boff + 16, 1, 0, cx.s));
	ENQ(gwf_launch_fill(cx.boff + 17, 1, p.n, cx.s));
	ENQ(gwf_launch_fill(cx.head + 16, 1, 0, cx.s));
	if (p.direct)  // reads in input order from the same one-range queue the probe uses
		return gwf_launch_persistent(p.g, rr, nullptr, cx.boff + 16, 0, cx.head + 16, p.bin[0], cx.scr[0],
		                             gwf_ws_bytes(p.bin[0].hbm), cx.ed, cx.status, cx.cnt, p.beat, cx.s);
	ENQ(gwf_launch_persistent(p.g, rp, nullptr, cx.boff + 16, 0, cx.head + 16, p.probe, cx.scratch,
	                          gwf_ws_bytes(p.probe.hbm), cx.pe, cx.ps, cx.cnt,
	                          p.beat ? p.beat + (size_t)p.bb.n_bins * p.beat_stride : nullptr, cx.s));
	ENQ(gwf_launch_binning(cx.pe, cx.ps, dr->r.ql, p.n, p.bb, cx.key, cx.hist, cx.ord, cx.boff, cx.head, cx.s));
	if (p.concurrent) {
		ENQ(hipEventRecord(cx.ev_fork, cx.s));
		for (int b = 0; b < p.bb.n_bins; ++b) ENQ(hipStreamWaitEvent(cx.sb[b], cx.ev_fork, 0));
	}
	for (int b = p.bb.n_bins - 1; b >= 0; --b) {  // costliest bin first
		hipStream_t st = p.concurrent ? cx.sb[b] : cx.s;
		u64 *scr = p.concurrent ? cx.scr[b] : cx.scr[0];
		ENQ(gwf_launch_persistent(p.g, rr, cx.ord, cx.boff, b, cx.head, p.bin[b], scr,
		                          gwf_ws_bytes(p.bin[b].hbm), cx.ed, cx.status, cx.cnt,
		                          p.beat ? p.beat + (size_t)b * p.beat_stride : nullptr, st));
	}
	if (p.concurrent) {
		for (int b = 0; b < p.bb.n_bins; ++b) {
			ENQ(hipEventRecord(cx.ev_join[b], cx.sb[b]));
			ENQ(hipStreamWaitEvent(cx.s, cx.ev_join[b], 0));
		}
	}
#undef ENQ
	return hipSuccess;
}

bool run_binned(RunCtx &cx, RunAcc &acc, const gwf_dgraph_t *dg, const DGraph &g, const gwf_dreads_t *dr, i32 v1, i32 maxql,
                size_t budget, bool graph_ok) {
	const i32 n = dr->r.n;
	u64 tv = 0;
	const gwf_tune_t t = tune_current(&tv);
	BinnedPlan p;
	p.g = g;
	p.n = n;
	p.v1 = v1;
	const bool mixed = g.mixed != 0;
	p.probe = make_cfg(cx, 1, 256, 8, std::min(maxql, kProbeLen), mixed);
	p.bb.n_bins = t.n_bins;
	size_t ws_sum = 0, ws_max = 0;
	for (int b = 0; b < t.n_bins; ++b) {
		p.bb.est_max[b] = t.bin[b].est_max;
		p.bin[b] = make_cfg(cx, t.bin[b].nw, t.bin[b].lds_states, t.bin[b].merge_runs, maxql, mixed);
		// Every resident workgroup holds one HBM workspace.  Sized for the
		// longest read they would exceed the budget when a few reads are far
		// longer than the rest; the workspaces are then sized for the longest
		// length that fits, and a read whose state outgrows them reports an
		// overflow and is re-run by the budgeted plain launches below.
		p.bin[b].hbm = caps_for(bin_ws_ql(maxql, p.bin[b].wgs, budget, mixed), 1, mixed);
		const size_t w = gwf_ws_bytes(p.bin[b].hbm) * (size_t)p.bin[b].wgs;
		ws_sum += w;
		ws_max = std::max(ws_max, w);
	}
	p.concurrent = ws_sum <= budget && env_int("GWF_BIN_STREAMS", 1) != 0;
	p.direct = t.n_bins == 1 && n <= p.bin[0].wgs && env_int("GWF_DIRECT", 1) != 0;
	ACC(grow(&cx.scratch, &cx.scratch_bytes, gwf_ws_bytes(p.probe.hbm) * (size_t)p.probe.wgs));
	ACC(grow(&cx.pe, &cx.pe_c, 4 * (size_t)n));
	ACC(grow(&cx.ps, &cx.ps_c, 4 * (size_t)n));
	ACC(grow(&cx.key, &cx.key_c, 4 * (size_t)n));
	ACC(grow(&cx.ord, &cx.ord_c, 4 * (size_t)n));
	for (int b = 0; b < (p.concurrent ? t.n_bins : 1); ++b) {
		const size_t w = p.concurrent ? gwf_ws_bytes(p.bin[b].hbm) * (size_t)p.bin[b].wgs : ws_max;
		ACC(grow(&cx.scr[b], &cx.scr_c[b], w));
	}
	Heartbeat hb;
	hb.period_ms = (int)env_int("GWF_HEARTBEAT_MS", 0);
	if (hb.period_ms > 0) {  // heartbeat runs are not captured
		i32 stride = p.probe.wgs;
		for (int b = 0; b < t.n_bins; ++b) stride = std::max(stride, p.bin[b].wgs);
		const size_t need = (size_t)stride * (t.n_bins + 1);
		if (need > cx.beat_n) {
			if (cx.beat) (void)hipHostFree(cx.beat);
			cx.beat = nullptr;
			cx.beat_n = 0;
			ACC(hipHostMalloc((void **)&cx.beat, sizeof(GwfBeat) * need, hipHostMallocCoherent | hipHostMallocMapped));
			cx.beat_n = need;
		}
		for (size_t w = 0; w < need; ++w) cx.beat[w].read = -1, cx.beat[w].beat = 0;
		p.beat = cx.beat;
		p.beat_stride = stride;
		hb.beat = cx.beat;
		hb.n = (i32)need;
		hb.what = "binned persistent";
	}
	const bool use_graph = graph_ok && env_int("GWF_GRAPH", 1) != 0 && !p.beat;
	acc.st.n_launch += p.direct ? 1 : 1 + t.n_bins;
	if (!use_graph) {
		ACC(hipEventRecord(cx.e0, cx.s));
		ACC(enqueue_binned(cx, dg, dr, p));
		ACC(hipEventRecord(cx.e1, cx.s));
		if (hb.beat) heartbeat_wait(cx, hb);
		ACC(hipEventSynchronize(cx.e1));
		float ms = 0;
		ACC(hipEventElapsedTime(&ms, cx.e0, cx.e1));
		acc.st.kernel_ms += ms;
		return true;
	}
	// everything the captured launches point at or were sized by.  The
	// handles enter by generation (unique per upload) and by the device
	// memory the launches read, never by host address: a handle freed and
	// re-created at the same address is a different key.
	std::vector<uintptr_t> key = {(uintptr_t)dg->gen, (uintptr_t)dr->gen, (uintptr_t)dg->mem, (uintptr_t)dr->mem,
	                              (uintptr_t)dr->ql_probe, (uintptr_t)n, (uintptr_t)tv, (uintptr_t)maxql, (uintptr_t)(v1 + 1),
	                              (uintptr_t)p.concurrent, (uintptr_t)p.direct, (uintptr_t)cx.ed, (uintptr_t)cx.status, (uintptr_t)cx.cnt,
	                              (uintptr_t)cx.scratch, (uintptr_t)cx.pe, (uintptr_t)cx.ps, (uintptr_t)cx.key,
	                              (uintptr_t)cx.ord, (uintptr_t)p.g.layout};
	for (int b = 0; b < GWF_MAX_BINS; ++b) key.push_back((uintptr_t)cx.scr[b]);
	for (int b = 0; b < t.n_bins; ++b) {
		key.push_back((uintptr_t)p.bin[b].nw);
		key.push_back((uintptr_t)p.bin[b].wgs);
		key.push_back((uintptr_t)p.bin[b].lds_bytes);
		key.push_back((uintptr_t)p.bin[b].hbm.m);
	}
	key.push_back((uintptr_t)p.probe.hbm.m);
	if (!cx.gexec || cx.gkey != key) {
		if (cx.gexec) (void)hipGraphExecDestroy(cx.gexec);
		cx.gexec = nullptr;
		cx.gkey.clear();
		ACC(hipStreamBeginCapture(cx.s, hipStreamCaptureModeThreadLocal));
		const hipError_t e = enqueue_binned(cx, dg, dr, p);
		hipGraph_t g = nullptr;
		const hipError_t e2 = hipStreamEndCapture(cx.s, &g);
		ACC(e);
		ACC(e2);
		const hipError_t e3 = hipGraphInstantiate(&cx.gexec, g, nullptr, nullptr, 0);
		(void)hipGraphDestroy(g);
		ACC(e3);
		cx.gkey = key;
	}
	return timed_launch(cx, acc, [&] { return hipGraphLaunch(cx.gexec, cx.s); }, nullptr, 0);
}

#undef ACC

}  // namespace

int gwf_ed_run(gwf_dgraph_t *dg, gwf_dreads_t *dr, int32_t v1, uint32_t max_lag, int32_t kernel, int32_t *out,
               gwf_run_stats_t *stats) {
	return gwf_ed_run_timed(dg, dr, v1, max_lag, kernel, out, stats, nullptr, nullptr);
}

int gwf_ed_run_timed(gwf_dgraph_t *dg, gwf_dreads_t *dr, int32_t v1, uint32_t max_lag, int32_t kernel, int32_t *out,
                     gwf_run_stats_t *stats, int64_t *t_start_ns, int64_t *t_end_ns) {
	(void)max_lag;  // accepted and ignored, as in the reference (gwf-ed.cu:1286-1287)
	const auto t0 = std::chrono::steady_clock::now();
	if (!dg || !dr || (!out && dr->r.n > 0)) { set_err("gwf_ed_run: NULL argument"); return GWF_ERR_ARG; }
	if (v1 >= 0 && (u32)v1 >= dg->n_vtx) { set_err("gwf_ed_run: v1 %d >= n_vtx %u", v1, dg->n_vtx); return GWF_ERR_ARG; }
	if (kernel != GWF_KERNEL_PLAIN && kernel != GWF_KERNEL_PERSISTENT) {
		set_err("gwf_ed_run: unknown kernel %d", kernel);
		return GWF_ERR_ARG;
	}
	const i32 n = dr->r.n;
	for (i32 i = 0; i < n; ++i)
		if ((u32)dr->v0[i] >= dg->n_vtx) { set_err("gwf_ed_run: read %d has v0 %d >= n_vtx %u", i, dr->v0[i], dg->n_vtx); return GWF_ERR_ARG; }
	DGraph g;
	if (int rc = gwf_run_graph(dg, dr, "gwf_ed_run", &g)) return rc;
	RunAcc acc;
	memset(&acc.st, 0, sizeof(acc.st));
	if (n == 0) { if (stats) *stats = acc.st; return GWF_OK; }

	RunCtx &cx = g_run;
	std::lock_guard<std::mutex> lock(cx.mu);
	int rc = ctx_prepare(cx, (size_t)n);
	if (rc) return rc;
	std::vector<i32> status(n);
	std::vector<int64_t> cnt(GWF_NCOUNT * (size_t)n);
	size_t free_b = 0, tot_b = 0;
	if (hipMemGetInfo(&free_b, &tot_b) != hipSuccess) { set_err("hipMemGetInfo failed"); return GWF_ERR_HIP; }
	// device bytes the run's workspaces may take (GWF_BUDGET_BYTES overrides
	// both, for tests): the persistent launches' per-workgroup workspaces, and
	// the plain launches' per-read ones (the reference's pool takes 90% of
	// the free memory, main.cu:541; half of it here)
	const size_t held = ctx_held(cx);
	const size_t budget = (size_t)env_int("GWF_BUDGET_BYTES", (int64_t)std::min<size_t>((free_b + held) / 2, (size_t)32 << 30));
	const size_t plain_budget =
	    (size_t)env_int("GWF_BUDGET_BYTES", (int64_t)std::min<size_t>((free_b + held) / 2, (size_t)128 << 30));
	std::vector<i32> todo;
	int scale = 1;
	bool ok = true;

	if (kernel == GWF_KERNEL_PERSISTENT) {
		i32 maxql = 0;
		for (i32 q : dr->ql) maxql = std::max(maxql, q);
		ok = run_binned(cx, acc, dg, g, dr, v1, maxql, budget, true);
		if (ok) ok = hipMemcpy(status.data(), cx.status, 4 * (size_t)n, hipMemcpyDeviceToHost) == hipSuccess;
		for (i32 i = 0; i < n && ok; ++i)
			if (status[i] == GWF_ST_OVERFLOW) todo.push_back(i);
			else if (status[i] != GWF_ST_OK) { set_err("gwf_ed_run: read %d ended with internal status %d", i, status[i]); rc = GWF_ERR_STATE; }
		if (!todo.empty()) {
			acc.st.n_retry += (int64_t)todo.size();
			scale = 4;
		}
	} else {
		todo.resize(n);
		for (i32 i = 0; i < n; ++i) todo[i] = i;
	}
	// plain kernel (and re-runs of reads that overflowed their HBM workspace)
	while (ok && rc == GWF_OK && !todo.empty()) {
		DReads rv = dr->r;
		rv.v1 = v1;
		ok = run_plain(cx, acc, g, rv, dr->ql, todo, scale, plain_budget);
		if (!ok) break;
		ok = hipMemcpy(status.data(), cx.status, 4 * (size_t)n, hipMemcpyDeviceToHost) == hipSuccess;
		std::vector<i32> again;
		for (i32 r : todo) {
			if (status[r] == GWF_ST_OVERFLOW) again.push_back(r);
			else if (status[r] != GWF_ST_OK) { set_err("gwf_ed_run: read %d ended with internal status %d", r, status[r]); rc = GWF_ERR_STATE; }
		}
		if (!again.empty()) {
			acc.st.n_retry += (int64_t)again.size();
			scale *= 4;
			if (caps_for(1, scale, false).x >= (1 << 25) && scale > (1 << 12)) {
				set_err("gwf_ed_run: %zu reads exceed the maximum workspace", again.size());
				rc = GWF_ERR_CAPACITY;
			}
		}
		todo.swap(again);
	}
	if (!ok) {
		set_err("gwf_ed_run: %s failed: %s", acc.what, hipGetErrorString(acc.err));
		rc = GWF_ERR_HIP;
	}
	if (rc == GWF_OK) {
		if (hipMemcpy(out, cx.ed, 4 * (size_t)n, hipMemcpyDeviceToHost) != hipSuccess ||
		    hipMemcpy(cnt.data(), cx.cnt, 8 * GWF_NCOUNT * (size_t)n, hipMemcpyDeviceToHost) != hipSuccess) {
			set_err("gwf_ed_run: result copy failed");
			rc = GWF_ERR_HIP;
		}
	}
	if (rc == GWF_OK) {
		int64_t tmin = INT64_MAX;
		for (i32 i = 0; i < n; ++i) {
			const int64_t *c = &cnt[GWF_NCOUNT * (size_t)i];
			acc.st.sum_steps += c[0];
			acc.st.sum_states += c[1];
			acc.st.sum_chars += c[2];
			acc.st.n_spill += c[5] >= 0;
			acc.st.sum_hbm_steps += c[5] >= 0 ? c[0] - c[5] : 0;
			tmin = std::min(tmin, c[3]);
		}
		const double ns_per_tick = cx.wall_khz > 0 ? 1e6 / cx.wall_khz : 10.0;
		for (i32 i = 0; i < n; ++i) {
			const int64_t *c = &cnt[GWF_NCOUNT * (size_t)i];
			if (t_start_ns) t_start_ns[i] = (int64_t)((c[3] - tmin) * ns_per_tick);
			if (t_end_ns) t_end_ns[i] = (int64_t)((c[4] - tmin) * ns_per_tick);
		}
	}
	acc.st.wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
	if (stats) *stats = acc.st;
	return rc;
}

int gwf_ed_batch(gwf_dgraph_t *dg, int32_t n, const char *const *q, const int32_t *ql, const int32_t *v0, int32_t v1,
                 uint32_t max_lag, int32_t kernel, int32_t *out, gwf_run_stats_t *stats) {
	if (n < 0 || (n > 0 && (!q || !ql))) { set_err("gwf_ed_batch: bad argument"); return GWF_ERR_ARG; }
	std::string buf;
	std::vector<int64_t> off(n ? n : 1);
	for (i32 i = 0; i < n; ++i) {
		if (ql[i] < 0) { set_err("gwf_ed_batch: negative length"); return GWF_ERR_ARG; }
		off[i] = (int64_t)buf.size();
		buf.append(q[i], (size_t)ql[i]);
		buf.push_back('\0');
	}
	gwf_dreads_t *dr = nullptr;
	int rc = gwf_dreads_upload(n, buf.data(), off.data(), ql, v0, &dr);
	if (rc) return rc;
	rc = gwf_ed_run(dg, dr, v1, max_lag, kernel, out, stats);
	gwf_dreads_free(dr);
	return rc;
}

int gwf_autotune(gwf_dgraph_t *dg, gwf_dreads_t *dr, int32_t max_evals, gwf_tune_t *out, char *report,
                 int64_t report_len) {
	static_assert(GWF_MAX_BINS == GWF_TUNE_MAX_BINS, "bin tables agree");
	if (!dg || !dr || max_evals < 0) { set_err("gwf_autotune: bad argument"); return GWF_ERR_ARG; }
	if (max_evals == 0) max_evals = 64;
	const i32 n = dr->r.n;
	std::string rep;
	char line[320];
	gwf_tune_t best_t = tune_current(nullptr);
	if (n > 0) {
		RunCtx &cx = g_run;
		std::lock_guard<std::mutex> lock(cx.mu);
		int rc = ctx_prepare(cx, (size_t)n);
		if (rc) return rc;
		i32 maxql = 0;
		for (i32 q : dr->ql) maxql = std::max(maxql, q);
		DGraph g;
		if (int rc2 = gwf_run_graph(dg, dr, "gwf_autotune", &g)) return rc2;
		size_t free_b = 0, tot_b = 0;
		if (hipMemGetInfo(&free_b, &tot_b) != hipSuccess) { set_err("hipMemGetInfo failed"); return GWF_ERR_HIP; }
		const size_t budget = (size_t)env_int("GWF_BUDGET_BYTES", (int64_t)std::min<size_t>((free_b + ctx_held(cx)) / 2, (size_t)32 << 30));
		RunAcc acc;
		memset(&acc.st, 0, sizeof(acc.st));
		int evals = 0;
		std::vector<i32> status(n);
		// one end-to-end binned run of dr with table t (eager, not captured)
		auto eval = [&](const gwf_tune_t &t, const char *what) -> double {
			if (evals >= max_evals) return 1e30;
			{
				std::lock_guard<std::mutex> lk(g_tune_mu);
				g_tune = t;
				g_tune_version += 1;
			}
			const double before = acc.st.kernel_ms;
			if (!run_binned(cx, acc, dg, g, dr, -1, maxql, budget, false)) return -1.0;
			if (hipMemcpy(status.data(), cx.status, 4 * (size_t)n, hipMemcpyDeviceToHost) != hipSuccess) return -1.0;
			i32 over = 0;
			for (i32 x : status) over += x != GWF_ST_OK;
			const double ms = acc.st.kernel_ms - before + (over ? 1e6 : 0.0);  // overflowing tables lose
			++evals;
			std::string desc;
			for (int b = 0; b < t.n_bins; ++b) {
				char d[64];
				snprintf(d, sizeof d, "%s(%d,%d,%d)", b ? " " : "", t.bin[b].nw, t.bin[b].lds_states, t.bin[b].merge_runs);
				desc += d;
			}
			snprintf(line, sizeof line, "eval %2d %-14s %9.3f ms  bins (nw,lds,mw): %s\n", evals, what, ms, desc.c_str());
			rep += line;
			return ms;
		};
		auto uniform = [](gwf_tune_t t, const gwf_tune_bin_t &c) {
			for (int b = 0; b < t.n_bins; ++b) {
				const i32 e = t.bin[b].est_max;
				t.bin[b] = c;
				t.bin[b].est_max = e;
			}
			return t;
		};
		const gwf_tune_t base = builtin_tune();
		static const gwf_tune_bin_t cands[] = {{0, 4, 1024, 0}, {0, 4, 2048, 0}, {0, 4, 512, 0}, {0, 4, 256, 0},
		                                       {0, 2, 512, 0},  {0, 2, 256, 0},  {0, 2, 128, 0}, {0, 1, 256, 0},
		                                       {0, 1, 128, 0}};
		gwf_tune_t one;  // a single bin: one queue, one launch
		memset(&one, 0, sizeof(one));
		one.n_bins = 1;
		one.bin[0].est_max = 1 << 30;
		const double keep = 0.98;  // a change must gain 2% (run-to-run noise is ~1-2%)
		// 1. the best single configuration, as one bin
		double best = 1e30;
		gwf_tune_bin_t best_c = cands[0];
		(void)eval(uniform(one, cands[0]), "warm-up");
		for (const gwf_tune_bin_t &c : cands) {
			const double ms = eval(uniform(one, c), "one bin");
			if (ms < 0) { set_err("gwf_autotune: run failed: %s", hipGetErrorString(acc.err)); return GWF_ERR_HIP; }
			if (ms < best) best = ms, best_t = uniform(one, c), best_c = c;
		}
		// 2. the built-in bins, all with that configuration, then per bin,
		// costliest first, each other configuration for that bin alone
		gwf_tune_t cur = uniform(base, best_c);
		double cur_ms = eval(cur, "bins");
		for (int b = base.n_bins - 1; b >= 0 && cur_ms >= 0; --b) {
			for (const gwf_tune_bin_t &c : cands) {
				if (c.nw == cur.bin[b].nw && c.lds_states == cur.bin[b].lds_states) continue;
				gwf_tune_t t = cur;
				t.bin[b].nw = c.nw, t.bin[b].lds_states = c.lds_states;
				const double ms = eval(t, "bin config");
				if (ms >= 0 && ms < cur_ms * keep) cur_ms = ms, cur = t;
			}
		}
		if (cur_ms >= 0 && cur_ms < best * keep) {  // confirm: bins beat one bin on a second run too
			const double again = eval(cur, "confirm bins");
			const double one_again = eval(best_t, "confirm one");
			if (again >= 0 && one_again >= 0 && again < one_again * keep) best = again, best_t = cur;
		}
		// 3. the sort variant, for the whole table
		static const i32 mws[] = {2, 4, 8};
		for (i32 mw : mws) {
			gwf_tune_t t = best_t;
			for (int b = 0; b < t.n_bins; ++b) t.bin[b].merge_runs = mw;
			const double ms = eval(t, "sort variant");
			if (ms >= 0 && ms < best * keep) best = ms, best_t = t;
		}
		snprintf(line, sizeof line, "chose %.3f ms after %d runs of %d reads\n", best, evals, n);
		rep += line;
	}
	if (gwf_tune_set(&best_t) != GWF_OK) return GWF_ERR_STATE;
	if (out) *out = best_t;
	if (report && report_len > 0) {
		const size_t k = std::min((size_t)report_len - 1, rep.size());
		memcpy(report, rep.data(), k);
		report[k] = 0;
	}
	return GWF_OK;
}

int gwf_tune_get(gwf_tune_t *out) {
	if (!out) { set_err("gwf_tune_get: NULL argument"); return GWF_ERR_ARG; }
	*out = tune_current(nullptr);
	return GWF_OK;
}

int gwf_tune_set(const gwf_tune_t *t) {
	const gwf_tune_t x = t ? *t : builtin_tune();
	if (tune_check(x) != GWF_OK) { set_err("gwf_tune_set: invalid table"); return GWF_ERR_ARG; }
	std::lock_guard<std::mutex> lk(g_tune_mu);
	g_tune = x;
	g_tune_version += 1;
	return GWF_OK;
}

int gwf_tune_save(const char *path) {
	if (!path) { set_err("gwf_tune_save: NULL path"); return GWF_ERR_ARG; }
	const gwf_tune_t t = tune_current(nullptr);
	FILE *f = fopen(path, "w");
	if (!f) { set_err("gwf_tune_save: cannot write %s", path); return GWF_ERR_ARG; }
	fprintf(f, "# gwfa tuning table: bin <est_max> <nw> <lds_states> <merge_runs>\n");
	for (int b = 0; b < t.n_bins; ++b)
		fprintf(f, "bin %d %d %d %d\n", t.bin[b].est_max, t.bin[b].nw, t.bin[b].lds_states, t.bin[b].merge_runs);
	fclose(f);
	return GWF_OK;
}

int gwf_tune_load(const char *path) {
	if (!path) { set_err("gwf_tune_load: NULL path"); return GWF_ERR_ARG; }
	gwf_tune_t t;
	if (gwf_tune_read_file(path, &t) != GWF_OK) { set_err("gwf_tune_load: %s is not a valid table", path); return GWF_ERR_ARG; }
	return gwf_tune_set(&t);
}

int32_t gwf_ed(const gwf_graph_t *g, int32_t ql, const char *q, int32_t v0, int32_t v1, uint32_t max_lag,
               gwf_edbuf_t *buf, uint64_t *diag_shared) {
	(void)buf;
	(void)diag_shared;
	if (!g || (!q && ql > 0) || ql < 0) { set_err("gwf_ed: bad argument"); return GWF_ERR_ARG; }
	if (v0 < 0 || (uint32_t)v0 >= g->n_vtx) { set_err("gwf_ed: v0 %d out of range", v0); return GWF_ERR_ARG; }
	gwf_dgraph_t *dg = nullptr;
	{
		std::lock_guard<std::mutex> lk(g_cache_mu);
		auto it = g_cache.find(g);
		if (it != g_cache.end() && !it->second.matches(g)) {
			gwf_dgraph_free(it->second.dg);
			g_cache.erase(it);
			it = g_cache.end();
		}
		if (it != g_cache.end()) dg = it->second.dg;
		else {
			int rc = gwf_dgraph_upload(g, &dg);
			if (rc) return rc;
			CachedGraph c;
			c.dg = dg;
			c.n_vtx = g->n_vtx, c.n_arc = g->n_arc;
			c.len = g->len, c.seq = g->seq, c.arc = g->arc, c.aux = g->aux;
			g_cache.emplace(g, c);
		}
	}
	const char *qq = q ? q : "";
	int32_t out = -1;
	int rc = gwf_ed_batch(dg, 1, &qq, &ql, &v0, v1, max_lag, GWF_KERNEL_PLAIN, &out, nullptr);
	return rc ? rc : out;
}

}  // extern "C"
