// gwf_runtime.hip — the C-ABI of libgwfa_hip.so (include/gwfa.h): graph
// index, device-resident graph and reads, batched alignment with automatic
// re-runs of reads that outgrow their workspace, and the host-callable gwf_ed.
#include <algorithm>
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

int device_ready() {
	int n = 0;
	if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
		set_err("no HIP device visible");
		return GWF_ERR_NODEV;
	}
	return GWF_OK;
}

}  // namespace

struct gwf_dreads_s {
	DReads r;
	const i32 *ql_probe = nullptr;  // device: min(ql, kProbeLen) per read
	void *mem = nullptr;
	std::vector<i32> ql, v0;
};

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
	size_t n_cap = 0;
	u64 *scratch = nullptr;
	size_t scratch_bytes = 0;
	GwfBeat *beat = nullptr;  // host-coherent heartbeat records
	size_t beat_n = 0;
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
		c.ed = c.status = c.idx = c.head = nullptr;
		c.cnt = nullptr;
		c.scratch = nullptr;
		c.n_cap = c.scratch_bytes = 0;
		HIPCHK(hipStreamCreate(&c.s));
		HIPCHK(hipEventCreate(&c.e0));
		HIPCHK(hipEventCreate(&c.e1));
		HIPCHK(hipDeviceGetAttribute(&c.wall_khz, hipDeviceAttributeWallClockRate, dev));
		HIPCHK(hipDeviceGetAttribute(&c.n_cu, hipDeviceAttributeMultiprocessorCount, dev));
		HIPCHK(hipMalloc((void **)&c.head, 256));
		c.dev = dev;
	}
	if (c.n_cap < n) {
		const size_t m = std::max<size_t>(n, 1024);
		size_t c0 = c.n_cap * 4, c1 = c.n_cap * 4, c2 = c.n_cap * 4, c3 = c.n_cap * 8 * GWF_NCOUNT;
		c.n_cap = 0;
		HIPCHK(grow(&c.ed, &c0, 4 * m));
		HIPCHK(grow(&c.status, &c1, 4 * m));
		HIPCHK(grow(&c.idx, &c2, 4 * m));
		HIPCHK(grow(&c.cnt, &c3, 8 * GWF_NCOUNT * m));
		c.n_cap = m;
	}
	return GWF_OK;
}

int64_t env_int(const char *name, int64_t dflt) {
	const char *v = getenv(name);
	return v && *v ? atoll(v) : dflt;
}

// HBM workspace of a read of length ql: the reference's per-read sizing
// (main.cu:498-504) plus headroom, times `scale` after overflows.
// GWF_DEBUG_CAP_A overrides the base (tests force the re-run path with it).
Caps caps_for(i32 ql, int scale) {
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
	return c;
}

// The tuning table: waves per read and LDS wavefront capacity for a bin of
// reads, from the largest expected wavefront of the bin.  Overridable with
// GWF_PERSIST_NW / GWF_PERSIST_M.
PersistCfg choose_cfg(const RunCtx &cx, int64_t est_states, i32 max_ql) {
	PersistCfg c;
	// Measured on MI355X (config 2 sweep): LDS capacity beyond ~1024 states
	// costs more in resident reads than it saves in HBM spills.
	i32 m = est_states <= 192 ? 256 : est_states <= 512 ? 512 : 1024;
	i32 nw = m <= 256 ? 1 : m <= 512 ? 2 : 4;
	nw = (i32)env_int("GWF_PERSIST_NW", nw);
	m = (i32)env_int("GWF_PERSIST_M", m);
	c.nw = nw;
	c.lds = lds_caps(m);
	c.lds_bytes = gwf_lds_bytes(c.lds);
	c.hbm = caps_for(max_ql, 1);
	const int per_cu = gwf_persist_occupancy(nw, c.lds_bytes);
	c.wgs = std::max(1, per_cu) * std::max(1, cx.n_cu);
	c.wgs = (i32)env_int("GWF_PERSIST_WGS", c.wgs);
	return c;
}

std::mutex g_cache_mu;
std::unordered_map<const gwf_graph_t *, gwf_dgraph_t *> g_cache;

}  // namespace

int64_t gwf_env_int(const char *name, int64_t dflt) { return env_int(name, dflt); }
Caps gwf_caps_for(i32 ql, int scale) { return caps_for(ql, scale); }
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
	{
		std::lock_guard<std::mutex> lk(g_cache_mu);
		auto it = g_cache.find(g);
		if (it != g_cache.end()) {
			gwf_dgraph_free(it->second);
			g_cache.erase(it);
		}
	}
	free(g->aux);
	g->aux = nullptr;
}

int gwf_dgraph_upload(const gwf_graph_t *g, gwf_dgraph_t **out) {
	if (!g || !out) { set_err("gwf_dgraph_upload: NULL argument"); return GWF_ERR_ARG; }
	*out = nullptr;
	if (!g->aux && g->n_arc) { set_err("gwf_dgraph_upload: graph not indexed (call gwf_ed_index)"); return GWF_ERR_ARG; }
	if (g->n_vtx == 0 || g->n_vtx >= GWF_MAX_VTX) { set_err("gwf_dgraph_upload: n_vtx %u out of range", g->n_vtx); return GWF_ERR_ARG; }
	int rc = device_ready();
	if (rc) return rc;
	const u32 nv = g->n_vtx, na = g->n_arc;
	std::vector<uint4> vinfo(nv);
	std::vector<i32> in_ol(nv, -1);
	uint64_t total = 0;
	for (u32 v = 0; v < nv; ++v) {
		if (g->len[v] == 0 || g->len[v] > (u32)GWF_MAX_VLEN) {
			set_err("gwf_dgraph_upload: vertex %u has length %u (supported: 1..%d)", v, g->len[v], GWF_MAX_VLEN);
			return GWF_ERR_ARG;
		}
		total += g->len[v];
	}
	if (total + kSeqPad >= (1ull << 32)) { set_err("gwf_dgraph_upload: graph sequence exceeds 4 GiB"); return GWF_ERR_ARG; }
	std::vector<uint4> arcs(na ? na : 1, make_uint4(0, 0, 0, 0));
	std::vector<u32> seq_off(nv);
	{
		uint64_t o = 0;
		for (u32 v = 0; v < nv; ++v) seq_off[v] = (u32)o, o += g->len[v];
	}
	for (u32 i = 0; i < na; ++i) {
		const u32 u = (u32)(g->arc[i].a >> 32), w = (u32)g->arc[i].a;
		const i32 o = g->arc[i].o;
		if (u >= nv || w >= nv || o < 0 || (u32)o >= g->len[w]) {
			set_err("gwf_dgraph_upload: arc %u (%u->%u, overlap %d) is invalid", i, u, w, o);
			return GWF_ERR_ARG;
		}
		if (in_ol[w] >= 0 && in_ol[w] != o) {
			set_err("gwf_dgraph_upload: vertex %u has in-arcs with different overlaps (%d, %d); "
			        "the GPU path needs one overlap per target vertex", w, in_ol[w], o);
			return GWF_ERR_ARG;
		}
		in_ol[w] = o;
		arcs[i] = make_uint4(w, (u32)o | (u32)(uint8_t)g->seq[w][o] << 24, seq_off[w], g->len[w]);
	}
	uint64_t so = 0;
	for (u32 v = 0; v < nv; ++v) {
		const uint64_t ax = g->aux ? g->aux[v] : 0;
		vinfo[v] = make_uint4((u32)so, g->len[v], (u32)(ax >> 32), (u32)ax);
		so += g->len[v];
	}
	const size_t b_vi = sizeof(uint4) * nv, b_seq = (size_t)total + kSeqPad, b_arc = sizeof(uint4) * arcs.size();
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
	dg->n_vtx = nv;
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
		total += ql[i] + 1;
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
	if (e != hipSuccess) {
		set_err("gwf_dreads_upload: hipMemcpy failed: %s", hipGetErrorString(e));
		gwf_dreads_free(dr);
		return GWF_ERR_HIP;
	}
	*out = dr;
	return GWF_OK;
}

void gwf_dreads_free(gwf_dreads_t *dr) {
	if (!dr) return;
	if (dr->mem) (void)hipFree(dr->mem);
	delete dr;
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

bool timed_launch(RunCtx &cx, RunAcc &acc, const std::function<hipError_t()> &launch, const Heartbeat *hb = nullptr) {
	ACC(hipEventRecord(cx.e0, cx.s));
	ACC(launch());
	ACC(hipEventRecord(cx.e1, cx.s));
	if (hb && hb->beat) heartbeat_wait(cx, *hb);
	ACC(hipEventSynchronize(cx.e1));
	float ms = 0;
	ACC(hipEventElapsedTime(&ms, cx.e0, cx.e1));
	acc.st.kernel_ms += ms;
	acc.st.n_launch += 1;
	return true;
}

// Plain kernel over `todo` in batches that fit the HBM budget.
bool run_plain(RunCtx &cx, RunAcc &acc, const gwf_dgraph_t *dg, const DReads &r, const std::vector<i32> &qlh,
               const std::vector<i32> &todo, int scale, size_t budget) {
	size_t pos = 0;
	while (pos < todo.size()) {
		i32 maxql = 0;
		size_t end = pos, ws = 0;
		while (end < todo.size()) {
			const i32 q = std::max(maxql, qlh[todo[end]]);
			const size_t wsb = gwf_ws_bytes(caps_for(q, scale));
			if (end > pos && wsb * (end - pos + 1) > budget) break;
			maxql = q, ws = wsb, ++end;
		}
		const Caps caps = caps_for(maxql, scale);
		const i32 nb = (i32)(end - pos);
		ACC(grow(&cx.scratch, &cx.scratch_bytes, ws * nb));
		ACC(hipMemcpyAsync(cx.idx, todo.data() + pos, 4 * (size_t)nb, hipMemcpyHostToDevice, cx.s));
		if (!timed_launch(cx, acc, [&] {
			    return gwf_launch_plain(dg->g, r, cx.idx, nb, caps, cx.scratch, ws, cx.ed, cx.status, cx.cnt, cx.s);
		    }))
			return false;
		pos = end;
	}
	return true;
}

// Persistent kernel over `order` (a queue drained by the resident workgroups).
bool run_persistent(RunCtx &cx, RunAcc &acc, const gwf_dgraph_t *dg, const DReads &r, const std::vector<i32> &order,
                    const PersistCfg &cfg) {
	const i32 n = (i32)order.size();
	if (n == 0) return true;
	const size_t ws = gwf_ws_bytes(cfg.hbm);
	ACC(grow(&cx.scratch, &cx.scratch_bytes, ws * (size_t)cfg.wgs));
	ACC(hipMemcpyAsync(cx.idx, order.data(), 4 * (size_t)n, hipMemcpyHostToDevice, cx.s));
	ACC(hipMemsetAsync(cx.head, 0, sizeof(i32), cx.s));
	Heartbeat hb;
	hb.period_ms = (int)env_int("GWF_HEARTBEAT_MS", 0);
	if (hb.period_ms > 0) {
		if ((size_t)cfg.wgs > cx.beat_n) {
			if (cx.beat) (void)hipHostFree(cx.beat);
			cx.beat = nullptr;
			cx.beat_n = 0;
			ACC(hipHostMalloc((void **)&cx.beat, sizeof(GwfBeat) * cfg.wgs, hipHostMallocCoherent | hipHostMallocMapped));
			cx.beat_n = cfg.wgs;
		}
		for (i32 w = 0; w < cfg.wgs; ++w) cx.beat[w].read = -1, cx.beat[w].beat = 0;
		hb.beat = cx.beat;
		hb.n = cfg.wgs;
		hb.what = "persistent";
	}
	return timed_launch(cx, acc, [&] {
		return gwf_launch_persistent(dg->g, r, cx.idx, n, cx.head, cfg, cx.scratch, ws, cx.ed, cx.status, cx.cnt,
		                             hb.beat, cx.s);
	}, &hb);
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
	if (v1 >= 0) { set_err("gwf_ed_run: v1 >= 0 is not supported on the GPU path"); return GWF_ERR_ARG; }
	if (kernel != GWF_KERNEL_PLAIN && kernel != GWF_KERNEL_PERSISTENT) {
		set_err("gwf_ed_run: unknown kernel %d", kernel);
		return GWF_ERR_ARG;
	}
	const i32 n = dr->r.n;
	for (i32 i = 0; i < n; ++i)
		if ((u32)dr->v0[i] >= dg->n_vtx) { set_err("gwf_ed_run: read %d has v0 %d >= n_vtx %u", i, dr->v0[i], dg->n_vtx); return GWF_ERR_ARG; }
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
	const size_t budget = std::min<size_t>((free_b + cx.scratch_bytes) / 2, (size_t)32 << 30);
	std::vector<i32> todo;
	int scale = 1;
	bool ok = true;

	if (kernel == GWF_KERNEL_PERSISTENT) {
		// probe: align the first kProbeLen bases of every read to estimate its error rate
		std::vector<i32> ident(n), probe(n);
		for (i32 i = 0; i < n; ++i) ident[i] = i;
		DReads rp = dr->r;
		rp.ql = dr->ql_probe;
		i32 maxql = 0;
		for (i32 q : dr->ql) maxql = std::max(maxql, q);
		PersistCfg pc = choose_cfg(cx, 64, std::min(maxql, kProbeLen));
		ok = run_persistent(cx, acc, dg, rp, ident, pc);
		if (ok) ok = hipMemcpy(probe.data(), cx.ed, 4 * (size_t)n, hipMemcpyDeviceToHost) == hipSuccess;
		if (ok) ok = hipMemcpy(status.data(), cx.status, 4 * (size_t)n, hipMemcpyDeviceToHost) == hipSuccess;
		// cost ~ ql + c * ED^2 (the wavefront of score s holds ~2s+1 diagonals)
		std::vector<double> cost(n);
		int64_t est_max = 0;
		for (i32 i = 0; i < n && ok; ++i) {
			const i32 pl = std::min(dr->ql[i], kProbeLen);
			const double e = (status[i] == GWF_ST_OK && pl > 0) ? (double)std::max(probe[i], 0) / pl : 0.3;
			const double ed = e * dr->ql[i];
			cost[i] = dr->ql[i] + 4.0 * ed * ed;
			est_max = std::max<int64_t>(est_max, (int64_t)(3.2 * ed) + 16);
		}
		std::vector<i32> order(ident);
		std::stable_sort(order.begin(), order.end(), [&](i32 x, i32 y) { return cost[x] > cost[y]; });
		if (ok) {
			const PersistCfg cfg = choose_cfg(cx, est_max, maxql);
			ok = run_persistent(cx, acc, dg, dr->r, order, cfg);
		}
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
		ok = run_plain(cx, acc, dg, dr->r, dr->ql, todo, scale, budget);
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
			if (caps_for(1, scale).x >= (1 << 25) && scale > (1 << 12)) {
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
		if (it != g_cache.end()) dg = it->second;
		else {
			int rc = gwf_dgraph_upload(g, &dg);
			if (rc) return rc;
			g_cache.emplace(g, dg);
		}
	}
	const char *qq = q ? q : "";
	int32_t out = -1;
	int rc = gwf_ed_batch(dg, 1, &qq, &ql, &v0, v1, max_lag, GWF_KERNEL_PLAIN, &out, nullptr);
	return rc ? rc : out;
}

}  // extern "C"
