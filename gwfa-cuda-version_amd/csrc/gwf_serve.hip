// gwf_serve.hip — continuous-batching server (gwf_serve_* in include/gwfa.h).
//
// One host thread per server owns the schedule:
//   * requests become visible at their release time (submit time by default);
//   * admission is FIFO: the head request is published when a ring slot is
//     free and its predicted footprint (device copy of the read + the HBM
//     workspace its length calls for) fits the remaining device budget, and
//     its bytes fit the pinned staging arena;
//   * publishing = write the ring entry and the read's bytes into
//     host-coherent memory, then advance the ring's tail word;
//   * the device loop (gwf_serve_kernel) claims requests as workgroups free
//     up, so new requests overlap with the ones already running;
//   * each result is picked up as soon as its workgroup publishes it, its
//     reservation is returned, and a workspace overflow is re-admitted at the
//     head of the queue with a 4x larger workspace;
//   * when nothing is queued or running for park_ms the device loop is
//     stopped (the GPU is released) and relaunched by the next request.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <queue>
#include <string>
#include <thread>
#include <vector>

#include "gwf_dev.h"
#include "gwf_policy.h"

namespace {

int64_t clock_ns() {
	return std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now().time_since_epoch())
	    .count();
}

u64 round_up(u64 x, u64 a) { return (x + a - 1) / a * a; }

// Best-fit allocator over [0, size) with coalescing of free neighbours.
class Arena {
public:
	void init(u64 size, u64 gran) {
		gran_ = gran;
		size_ = size / gran * gran;
		by_off_.clear();
		by_size_.clear();
		used_ = 0;
		if (size_) insert(0, size_);
	}
	u64 round(u64 len) const { return round_up(len ? len : 1, gran_); }
	bool alloc(u64 len, u64 *off) {
		len = round(len);
		auto it = by_size_.lower_bound(len);
		if (it == by_size_.end()) return false;
		const u64 o = it->second, sz = it->first;
		by_size_.erase(it);
		by_off_.erase(o);
		if (sz > len) insert(o + len, sz - len);
		used_ += len;
		*off = o;
		return true;
	}
	void release(u64 off, u64 len) {
		len = round(len);
		used_ -= len;
		auto nx = by_off_.lower_bound(off);
		if (nx != by_off_.end() && off + len == nx->first) {
			len += nx->second;
			erase(nx->first, nx->second);
		}
		auto pv = by_off_.lower_bound(off);
		if (pv != by_off_.begin()) {
			--pv;
			if (pv->first + pv->second == off) {
				const u64 po = pv->first, ps = pv->second;
				erase(po, ps);
				off = po;
				len += ps;
			}
		}
		insert(off, len);
	}
	u64 used() const { return used_; }
	u64 size() const { return size_; }

private:
	void insert(u64 o, u64 s) {
		by_off_[o] = s;
		by_size_.emplace(s, o);
	}
	void erase(u64 o, u64 s) {
		by_off_.erase(o);
		auto r = by_size_.equal_range(s);
		for (auto it = r.first; it != r.second; ++it)
			if (it->second == o) {
				by_size_.erase(it);
				break;
			}
	}
	u64 gran_ = 256, size_ = 0, used_ = 0;
	std::map<u64, u64> by_off_;
	std::multimap<u64, u64> by_size_;
};

struct Req {
	int64_t id = 0;
	std::string bytes;
	i32 ql = 0, v0 = 0, v1 = -1;
	int64_t t_arrival = 0;
	int scale = 1;
	i32 attempts = 0;
};

struct Flight {
	std::unique_ptr<Req> r;
	u64 seq = 0;
	u64 d_off = 0, d_len = 0, q_off = 0, q_len = 0;
	int64_t t_admit = 0;
	bool busy = false;
};

struct ByRelease {
	bool operator()(const std::unique_ptr<Req> &a, const std::unique_ptr<Req> &b) const {
		return a->t_arrival != b->t_arrival ? a->t_arrival > b->t_arrival : a->id > b->id;
	}
};

template <class T>
hipError_t host_coherent(T **h, T **d, size_t n) {
	*h = nullptr;
	*d = nullptr;
	hipError_t e = hipHostMalloc((void **)h, sizeof(T) * (n ? n : 1), hipHostMallocCoherent | hipHostMallocMapped);
	if (e != hipSuccess) return e;
	memset((void *)*h, 0, sizeof(T) * (n ? n : 1));
	return hipHostGetDevicePointer((void **)d, (void *)*h, 0);
}

}  // namespace

struct gwf_serve_s {
	// configuration
	int dev = 0;
	const gwf_dgraph_t *dg = nullptr;
	DGraph g{};      // dg's graph with the server's state layout
	i32 max_ql = 0;  // the longest request that layout holds
	gwf_serve_opts_t o{};
	PersistCfg cfg{};
	double ns_per_tick = 10.0;
	hipStream_t stream = nullptr;
	// host-coherent memory (h: host view, d: device view)
	SrvCtl *ctl_h = nullptr, *ctl_d = nullptr;
	SrvEntry *ring_h = nullptr, *ring_d = nullptr;
	SrvResult *res_h = nullptr, *res_d = nullptr;
	uint8_t *stage_h = nullptr, *stage_d = nullptr;
	GwfBeat *beat_h = nullptr, *beat_d = nullptr;
	// device memory
	u64 *claimed = nullptr;
	uint8_t *pool = nullptr;
	// shared with callers (under mu)
	std::mutex mu;
	std::condition_variable cv_in, cv_out;
	std::vector<std::unique_ptr<Req>> incoming;
	std::deque<gwf_serve_result_t> results;
	int64_t next_id = 0;
	bool quit = false;
	int fatal = GWF_OK;
	std::string fatal_msg;
	gwf_serve_stats_t st{};
	// server thread only
	std::thread th;
	Arena pool_a, stage_a;
	std::priority_queue<std::unique_ptr<Req>, std::vector<std::unique_ptr<Req>>, ByRelease> pending;
	std::deque<std::unique_ptr<Req>> ready;
	std::vector<Flight> slots;       // per result slot
	std::vector<u32> free_slots;     // result slots not in use
	std::vector<u32> inflight;       // result slots in use
	u64 tail = 0;
	bool running = false;
	int64_t idle_since = 0;
	// GWF_SERVE_TRACE=path: per-attempt records written at stop (diagnostics)
	std::string trace_path;
	std::vector<std::string> trace;
	int64_t loop_iters = 0, loop_ns_collect = 0, loop_ns_admit = 0, loop_ns_max = 0;
	std::vector<u64> beat_prev;
	std::vector<int64_t> beat_since;
	int64_t last_watch = 0;

	void loop();
	void finish(std::unique_ptr<Req> r, const SrvResult *res, int status, int64_t t_admit, int64_t now);
	bool admit_one(int64_t now);
	void collect(int64_t now);
	bool device_check(int64_t now);
	void watchdog(int64_t now);
	void park();
	void fail(int rc, const char *msg);
};

void gwf_serve_s::fail(int rc, const char *msg) {
	std::lock_guard<std::mutex> lk(mu);
	if (fatal == GWF_OK) {
		fatal = rc;
		fatal_msg = msg;
	}
}

void gwf_serve_s::finish(std::unique_ptr<Req> r, const SrvResult *res, int status, int64_t t_admit, int64_t now) {
	gwf_serve_result_t x{};
	x.id = r->id;
	x.ed = res ? res->ed : GWF_ED_NONE;
	x.status = status;
	x.t_arrival_ns = r->t_arrival;
	x.t_admit_ns = t_admit;
	x.t_done_ns = now;
	x.service_ns = res ? (int64_t)((res->t1 - res->t0) * ns_per_tick) : 0;
	x.steps = res ? res->steps : 0;
	x.states = res ? res->states : 0;
	x.attempts = r->attempts;
	x.spilled = res && res->spill_step >= 0;
	std::lock_guard<std::mutex> lk(mu);
	results.push_back(x);
	st.completed += 1;
	if (status == GWF_OK) st.bases_done += r->ql;
	else st.rejected += 1;
	cv_out.notify_all();
}

// Publish the head of `ready` if a slot, budget and staging space are free.
bool gwf_serve_s::admit_one(int64_t now) {
	Req &r = *ready.front();
	const Caps hc = gwf_caps_for(r.ql, r.scale, g.mixed != 0);
	const u64 q_len = round_up((u64)r.ql, 16) + 16;
	const u64 ws_off = round_up(round_up((u64)r.ql, 16) + kSeqPad, 256);
	const u64 d_len = ws_off + gwf_ws_bytes(hc);
	if (d_len > pool_a.size() || q_len > stage_a.size()) {
		std::unique_ptr<Req> x = std::move(ready.front());
		ready.pop_front();
		finish(std::move(x), nullptr, GWF_ERR_BUDGET, 0, now);
		return true;
	}
	if (free_slots.empty()) return false;  // ring_slots requests in flight
	u64 d_off = 0, q_off = 0;
	if (!pool_a.alloc(d_len, &d_off)) return false;
	if (!stage_a.alloc(q_len, &q_off)) {
		pool_a.release(d_off, d_len);
		return false;
	}
	uint8_t *qs = stage_h + q_off;
	memcpy(qs, r.bytes.data(), (size_t)r.ql);
	memset(qs + r.ql, 0, q_len - (u64)r.ql);
	const u32 slot = free_slots.back();
	free_slots.pop_back();
	// fewer than ring_slots requests are in flight, so ring[tail & mask] was
	// claimed long ago (every unclaimed request is in flight)
	SrvEntry &e = ring_h[tail & ((u64)o.ring_slots - 1)];
	e.q_off = q_off;
	e.d_off = d_off;
	e.ws_off = ws_off;
	e.ql = r.ql;
	e.v0 = r.v0;
	e.v1 = r.v1;
	e.hc = hc;
	e.res_slot = slot;
	Flight &f = slots[slot];
	f.r = std::move(ready.front());
	ready.pop_front();
	f.r->attempts += 1;
	f.seq = tail;
	f.d_off = d_off, f.d_len = d_len, f.q_off = q_off, f.q_len = q_len;
	f.t_admit = now;
	f.busy = true;
	inflight.push_back(slot);
	tail += 1;
	return true;
}

void gwf_serve_s::collect(int64_t now) {
	for (size_t i = 0; i < inflight.size();) {
		const u32 slot = inflight[i];
		Flight &f = slots[slot];
		const u64 done = __atomic_load_n(&res_h[slot].done, __ATOMIC_ACQUIRE);
		if (done != f.seq + 1) {
			++i;
			continue;
		}
		SrvResult res;
		memcpy(&res, (const void *)&res_h[slot], sizeof(res));
		if (!trace_path.empty()) {
			char line[320];
			snprintf(line, sizeof line, "%lld,%d,%d,%d,%lld,%lld,%lld,%lld,%lld,%lld,%d,%d,%lld\n", (long long)f.r->id,
			         f.r->ql, f.r->attempts, res.status, (long long)f.r->t_arrival, (long long)f.t_admit,
			         (long long)now, (long long)res.t_claim, (long long)res.t0, (long long)res.t1, res.wg, res.polls,
			         (long long)res.spill_step);
			trace.emplace_back(line);
		}
		pool_a.release(f.d_off, f.d_len);
		stage_a.release(f.q_off, f.q_len);
		std::unique_ptr<Req> r = std::move(f.r);
		f.busy = false;
		free_slots.push_back(slot);
		inflight[i] = inflight.back();
		inflight.pop_back();
		if (res.status == GWF_ST_OVERFLOW) {
			r->scale *= 4;
			if (gwf_caps_for(1, r->scale, false).x >= (1 << 25) && r->scale > (1 << 12)) {
				finish(std::move(r), &res, GWF_ERR_CAPACITY, f.t_admit, now);
			} else {
				{
					std::lock_guard<std::mutex> lk(mu);
					st.readmitted += 1;
				}
				ready.push_front(std::move(r));
			}
		} else {
			finish(std::move(r), &res, res.status == GWF_ST_OK ? GWF_OK : GWF_ERR_STATE, f.t_admit, now);
		}
	}
}

// Launch the device loop when work is published and it is not running;
// notice when it has exited by itself.  false on a device error.
bool gwf_serve_s::device_check(int64_t now) {
	(void)now;
	if (running) {
		const hipError_t q = hipStreamQuery(stream);
		if (q == hipSuccess) running = false;
		else if (q != hipErrorNotReady) {
			char m[256];
			snprintf(m, sizeof m, "gwf_serve: device loop failed: %s", hipGetErrorString(q));
			fail(GWF_ERR_HIP, m);
			return false;
		}
	}
	if (!running && !inflight.empty()) {
		SrvArgs a;
		a.g = g;
		a.ring = ring_d;
		a.res = res_d;
		a.ctl = ctl_d;
		a.claimed = claimed;
		a.stage = stage_d;
		a.pool = pool;
		a.beat = beat_d;
		a.mask = (u64)o.ring_slots - 1;
		a.lc = cfg.lds;
		a.mw = cfg.mw;
		// ~5 s of empty polls: the loop ends by itself if the host stops feeding it
		a.idle_polls = gwf_env_int("GWF_SERVE_IDLE_POLLS", 40000);
		for (i32 w = 0; w < cfg.wgs; ++w) beat_h[w].read = -1;
		const hipError_t e = gwf_launch_serve(a, cfg, stream);
		if (e != hipSuccess) {
			char m[256];
			snprintf(m, sizeof m, "gwf_serve: launch failed: %s", hipGetErrorString(e));
			fail(GWF_ERR_HIP, m);
			return false;
		}
		running = true;
		std::lock_guard<std::mutex> lk(mu);
		st.launches += 1;
	}
	return true;
}

void gwf_serve_s::park() {
	if (!running) return;
	__atomic_store_n(&ctl_h->word, tail | GWF_SRV_STOP, __ATOMIC_RELEASE);
	hipError_t e = hipStreamSynchronize(stream);
	__atomic_store_n(&ctl_h->word, tail, __ATOMIC_RELEASE);
	// every published request is done; the workgroups ended holding tickets
	// past the tail, so the next loop's tickets start at the tail again
	const u64 next = tail;
	if (e == hipSuccess) e = hipMemcpy(claimed, &next, sizeof next, hipMemcpyHostToDevice);
	running = false;
	if (e != hipSuccess) {
		char m[256];
		snprintf(m, sizeof m, "gwf_serve: device loop failed: %s", hipGetErrorString(e));
		fail(GWF_ERR_HIP, m);
	}
}

// Heartbeat watchdog: a busy workgroup whose record has not changed for
// watchdog_ms is counted as stalled (and reported when GWF_HEARTBEAT_MS is set).
void gwf_serve_s::watchdog(int64_t now) {
	if (now - last_watch < 200000000) return;
	last_watch = now;
	int64_t stalled = 0;
	for (i32 w = 0; w < cfg.wgs; ++w) {
		const volatile GwfBeat &b = beat_h[w];
		const u64 bt = b.beat;
		if (b.read < 0 || bt != beat_prev[w]) {
			beat_prev[w] = bt;
			beat_since[w] = now;
			continue;
		}
		if (now - beat_since[w] > (int64_t)o.watchdog_ms * 1000000) {
			++stalled;
			if (gwf_env_int("GWF_HEARTBEAT_MS", 0) > 0)
				fprintf(stderr, "[gwf serve watchdog] workgroup %d silent %.1f s on request %d step %d n_a %d\n", w,
				        (now - beat_since[w]) / 1e9, (int)b.read, (int)b.step, (int)b.n_a);
		}
	}
	std::lock_guard<std::mutex> lk(mu);
	st.stalled = stalled;
}

void gwf_serve_s::loop() {
	(void)hipSetDevice(dev);
	std::vector<std::unique_ptr<Req>> in;
	for (;;) {
		bool quitting;
		{
			std::lock_guard<std::mutex> lk(mu);
			in.swap(incoming);
			quitting = quit || fatal != GWF_OK;
		}
		for (auto &r : in) pending.push(std::move(r));
		in.clear();
		int64_t now = clock_ns();
		while (!pending.empty() && pending.top()->t_arrival <= now) {
			ready.push_back(std::move(const_cast<std::unique_ptr<Req> &>(pending.top())));
			pending.pop();
		}
		if (quitting) {  // requests not yet admitted are dropped
			size_t dropped = ready.size() + pending.size();
			ready.clear();
			while (!pending.empty()) pending.pop();
			if (dropped) {
				std::lock_guard<std::mutex> lk(mu);
				st.submitted -= (int64_t)dropped;
			}
		}
		const u64 tail0 = tail;
		while (!ready.empty() && admit_one(now)) {
		}
		if (tail != tail0) __atomic_store_n(&ctl_h->word, tail, __ATOMIC_RELEASE);
		const int64_t t_adm = clock_ns();
		const bool ok = device_check(now);
		collect(now);
		{
			const int64_t t_end = clock_ns();
			loop_iters += 1;
			loop_ns_admit += t_adm - now;
			loop_ns_collect += t_end - t_adm;
			loop_ns_max = std::max(loop_ns_max, t_end - now);
		}
		if (running) watchdog(now);
		{
			std::lock_guard<std::mutex> lk(mu);
			st.queued = (int64_t)(ready.size() + pending.size());
			st.inflight = (int64_t)inflight.size();
			st.inflight_peak = std::max(st.inflight_peak, st.inflight);
			st.budget_used = (int64_t)pool_a.used();
			st.budget_peak = std::max(st.budget_peak, st.budget_used);
		}
		if (!ok) {  // device error: every request on the device fails
			for (u32 slot : inflight) {
				slots[slot].busy = false;
				free_slots.push_back(slot);
				finish(std::move(slots[slot].r), nullptr, GWF_ERR_HIP, slots[slot].t_admit, now);
			}
			inflight.clear();
			running = false;
		}
		const bool idle = inflight.empty() && ready.empty();
		if (idle && running) {
			if (idle_since == 0) idle_since = now;
			else if (quitting || now - idle_since > (int64_t)o.park_ms * 1000000) park();
		} else if (!idle) {
			idle_since = 0;
		}
		if (quitting && idle && pending.empty()) {
			park();
			break;
		}
		if (idle) {
			std::unique_lock<std::mutex> lk(mu);
			int64_t wake = running ? now + (int64_t)o.park_ms * 1000000 : INT64_MAX;
			if (!pending.empty()) wake = std::min(wake, pending.top()->t_arrival);
			if (incoming.empty() && !quit) {
				if (wake == INT64_MAX) cv_in.wait(lk);
				else
					cv_in.wait_for(lk, std::chrono::nanoseconds(std::max<int64_t>(wake - clock_ns(), 0)));
			}
		} else {
			std::this_thread::sleep_for(std::chrono::microseconds(20));
		}
	}
}

extern "C" {

int64_t gwf_serve_clock_ns(void) { return clock_ns(); }

int gwf_serve_start(gwf_dgraph_t *dg, const gwf_serve_opts_t *opts, gwf_serve_t **out) {
	if (!dg || !out) {
		gwf_set_error("gwf_serve_start: NULL argument");
		return GWF_ERR_ARG;
	}
	*out = nullptr;
	int ndev = 0;
	if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
		gwf_set_error("no HIP device visible");
		return GWF_ERR_NODEV;
	}
	std::unique_ptr<gwf_serve_t> s(new gwf_serve_t());
	gwf_serve_opts_t o{};
	if (opts) o = *opts;
	if (o.state_budget_bytes <= 0) o.state_budget_bytes = (int64_t)1 << 30;
	if (o.staging_bytes <= 0) o.staging_bytes = (int64_t)64 << 20;
	if (o.ring_slots <= 0) o.ring_slots = 4096;
	// measured on the uneven-length workload (40k requests, 16 GiB budget):
	// 4 waves x 1024 states (4 workgroups per CU) 367 ms, 2 x 256 434 ms,
	// 2 x 128 451 ms
	if (o.nw <= 0) o.nw = 4;
	if (o.lds_states <= 0) o.lds_states = 1024;
	if (o.park_ms <= 0) o.park_ms = 20;
	if (o.watchdog_ms <= 0) o.watchdog_ms = 10000;
	if ((o.ring_slots & (o.ring_slots - 1)) || o.ring_slots > (1 << 24) ||
	    (o.nw != 1 && o.nw != 2 && o.nw != 4) || o.lds_states < 64) {
		gwf_set_error("gwf_serve_start: bad options (ring_slots a power of two, nw 1/2/4, lds_states >= 64)");
		return GWF_ERR_ARG;
	}
	s->o = o;
	s->dg = dg;
	// the state layout that holds the graph and the longest requests
	s->g = dg->g;
	s->max_ql = -1;
	for (int l = 0; l < GWF_N_LAYOUTS; ++l)
		if (dg->n_vtx <= gwf_layout_max_vtx(l) && dg->max_vlen <= gwf_layout_max_vlen(l) &&
		    gwf_layout_max_ql(l) > s->max_ql) {  // the upload checked that one layout holds the graph
			s->g.layout = (u32)l;
			s->max_ql = gwf_layout_max_ql(l);
		}
	hipError_t e = hipGetDevice(&s->dev);
	int khz = 0, n_cu = 0;
	if (e == hipSuccess) e = hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, s->dev);
	if (e == hipSuccess) e = hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, s->dev);
	if (khz > 0) s->ns_per_tick = 1e6 / khz;
	PersistCfg &c = s->cfg;
	c.nw = o.nw;
	c.lds = gwf_lds_caps(o.lds_states);
	c.lds_bytes = gwf_lds_bytes(c.lds);
	if (c.lds_bytes > gwf_max_lds_bytes()) {
		gwf_set_error("gwf_serve_start: lds_states %d needs %zu bytes of LDS (max %zu)", o.lds_states, c.lds_bytes,
		              gwf_max_lds_bytes());
		return GWF_ERR_ARG;
	}
	c.wgs = o.wgs > 0 ? o.wgs : std::max(1, gwf_serve_occupancy(c.nw, c.lds_bytes)) * std::max(1, n_cu);
	c.hbm = gwf_caps_for(1, 1, false);  // unused: each request brings its own workspace capacities
	c.mw = 0;
	if (e == hipSuccess) e = gwf_kernels_init();
	if (e == hipSuccess) e = hipStreamCreate(&s->stream);
	if (e == hipSuccess) e = host_coherent(&s->ctl_h, &s->ctl_d, 1);
	if (e == hipSuccess) e = host_coherent(&s->ring_h, &s->ring_d, (size_t)o.ring_slots);
	if (e == hipSuccess) e = host_coherent(&s->res_h, &s->res_d, (size_t)o.ring_slots);
	if (e == hipSuccess) e = host_coherent(&s->stage_h, &s->stage_d, (size_t)o.staging_bytes);
	if (e == hipSuccess) e = host_coherent(&s->beat_h, &s->beat_d, (size_t)c.wgs);
	if (e == hipSuccess) e = hipMalloc((void **)&s->claimed, 256);
	if (e == hipSuccess) e = hipMemset(s->claimed, 0, 256);
	if (e == hipSuccess) e = hipMalloc((void **)&s->pool, (size_t)o.state_budget_bytes);
	if (e != hipSuccess) {
		gwf_set_error("gwf_serve_start: allocation failed: %s", hipGetErrorString(e));
		gwf_serve_stop(s.release());
		return GWF_ERR_NOMEM;
	}
	for (i32 w = 0; w < c.wgs; ++w) s->beat_h[w].read = -1;
	s->pool_a.init((u64)o.state_budget_bytes, 256);
	s->stage_a.init((u64)o.staging_bytes, 16);
	s->slots.resize((size_t)o.ring_slots);
	for (i32 i = o.ring_slots - 1; i >= 0; --i) s->free_slots.push_back((u32)i);
	s->beat_prev.assign((size_t)c.wgs, 0);
	s->beat_since.assign((size_t)c.wgs, 0);
	s->st.budget_bytes = o.state_budget_bytes;
	if (const char *tp = getenv("GWF_SERVE_TRACE")) s->trace_path = tp;
	gwf_serve_t *p = s.release();
	p->th = std::thread([p] { p->loop(); });
	*out = p;
	return GWF_OK;
}

int gwf_serve_submit(gwf_serve_t *s, int32_t n, const char *seqs, const int64_t *off, const int32_t *ql,
                     const int32_t *v0, const int64_t *t_release_ns, int64_t *ids) {
	return gwf_serve_submit_v1(s, n, seqs, off, ql, v0, nullptr, t_release_ns, ids);
}

int gwf_serve_submit_v1(gwf_serve_t *s, int32_t n, const char *seqs, const int64_t *off, const int32_t *ql,
                        const int32_t *v0, const int32_t *v1, const int64_t *t_release_ns, int64_t *ids) {
	if (!s || n < 0 || (n > 0 && (!seqs || !off || !ql))) {
		gwf_set_error("gwf_serve_submit: bad argument");
		return GWF_ERR_ARG;
	}
	const int64_t now = clock_ns();
	std::vector<std::unique_ptr<Req>> batch;
	batch.reserve((size_t)n);
	for (i32 i = 0; i < n; ++i) {
		if (ql[i] < 0 || ql[i] > s->max_ql) {
			gwf_set_error("gwf_serve_submit: request %d has length %d (supported on this graph: 0..%d)", i, ql[i],
			              s->max_ql);
			return GWF_ERR_ARG;
		}
		const i32 v = v0 ? v0[i] : 0;
		if (v < 0 || (u32)v >= s->dg->n_vtx) {
			gwf_set_error("gwf_serve_submit: request %d has v0 %d (n_vtx %u)", i, v, s->dg->n_vtx);
			return GWF_ERR_ARG;
		}
		const i32 e = v1 && v1[i] >= 0 ? v1[i] : -1;
		if (e >= 0 && (u32)e >= s->dg->n_vtx) {
			gwf_set_error("gwf_serve_submit: request %d has v1 %d (n_vtx %u)", i, e, s->dg->n_vtx);
			return GWF_ERR_ARG;
		}
		std::unique_ptr<Req> r(new Req());
		r->bytes.assign(seqs + off[i], (size_t)ql[i]);
		r->ql = ql[i];
		r->v0 = v;
		r->v1 = e;
		r->t_arrival = t_release_ns ? std::max(now, t_release_ns[i]) : now;
		batch.push_back(std::move(r));
	}
	std::lock_guard<std::mutex> lk(s->mu);
	if (s->quit) {
		gwf_set_error("gwf_serve_submit: server is stopping");
		return GWF_ERR_STATE;
	}
	if (s->fatal != GWF_OK) {
		gwf_set_error("%s", s->fatal_msg.c_str());
		return s->fatal;
	}
	for (i32 i = 0; i < n; ++i) {
		batch[i]->id = s->next_id++;
		if (ids) ids[i] = batch[i]->id;
		s->incoming.push_back(std::move(batch[i]));
	}
	s->st.submitted += n;
	s->cv_in.notify_one();
	return GWF_OK;
}

int gwf_serve_poll(gwf_serve_t *s, gwf_serve_result_t *out, int32_t max, int32_t wait_ms) {
	if (!s || max < 0 || (max > 0 && !out)) {
		gwf_set_error("gwf_serve_poll: bad argument");
		return GWF_ERR_ARG;
	}
	std::unique_lock<std::mutex> lk(s->mu);
	if (s->results.empty() && wait_ms > 0)
		s->cv_out.wait_for(lk, std::chrono::milliseconds(wait_ms), [s] { return !s->results.empty(); });
	int32_t k = 0;
	while (k < max && !s->results.empty()) {
		out[k++] = s->results.front();
		s->results.pop_front();
	}
	return k;
}

int gwf_serve_drain(gwf_serve_t *s, int32_t timeout_ms) {
	if (!s) {
		gwf_set_error("gwf_serve_drain: NULL server");
		return GWF_ERR_ARG;
	}
	std::unique_lock<std::mutex> lk(s->mu);
	auto done = [s] { return s->st.completed >= s->st.submitted || s->fatal != GWF_OK; };
	if (timeout_ms < 0) s->cv_out.wait(lk, done);
	else if (!s->cv_out.wait_for(lk, std::chrono::milliseconds(timeout_ms), done)) {
		gwf_set_error("gwf_serve_drain: %lld of %lld requests still open after %d ms",
		              (long long)(s->st.submitted - s->st.completed), (long long)s->st.submitted, timeout_ms);
		return GWF_ERR_STATE;
	}
	if (s->fatal != GWF_OK) {
		gwf_set_error("%s", s->fatal_msg.c_str());
		return s->fatal;
	}
	return GWF_OK;
}

int gwf_serve_stats(gwf_serve_t *s, gwf_serve_stats_t *out) {
	if (!s || !out) {
		gwf_set_error("gwf_serve_stats: NULL argument");
		return GWF_ERR_ARG;
	}
	std::lock_guard<std::mutex> lk(s->mu);
	*out = s->st;
	return GWF_OK;
}

int gwf_serve_stop(gwf_serve_t *s) {
	if (!s) return GWF_OK;
	if (s->th.joinable()) {
		{
			std::lock_guard<std::mutex> lk(s->mu);
			s->quit = true;
			s->cv_in.notify_all();
		}
		s->th.join();
	}
	int rc;
	{
		std::lock_guard<std::mutex> lk(s->mu);
		rc = s->fatal;
		if (rc != GWF_OK) gwf_set_error("%s", s->fatal_msg.c_str());
	}
	if (s->stream) (void)hipStreamSynchronize(s->stream);
	if (!s->trace_path.empty()) {
		if (FILE *f = fopen(s->trace_path.c_str(), "w")) {
			fprintf(f, "# ns_per_tick=%.6f wgs=%d loop_iters=%lld admit_ns=%lld collect_ns=%lld max_iter_ns=%lld\n",
			        s->ns_per_tick, s->cfg.wgs, (long long)s->loop_iters, (long long)s->loop_ns_admit,
			        (long long)s->loop_ns_collect, (long long)s->loop_ns_max);
			fprintf(f, "id,ql,attempt,status,t_arrival,t_admit,t_done,d_claim,d_t0,d_t1,wg,polls,spill_step\n");
			for (const std::string &l : s->trace) fputs(l.c_str(), f);
			fclose(f);
		}
	}
	if (s->pool) (void)hipFree(s->pool);
	if (s->claimed) (void)hipFree(s->claimed);
	if (s->ctl_h) (void)hipHostFree(s->ctl_h);
	if (s->ring_h) (void)hipHostFree(s->ring_h);
	if (s->res_h) (void)hipHostFree(s->res_h);
	if (s->stage_h) (void)hipHostFree(s->stage_h);
	if (s->beat_h) (void)hipHostFree(s->beat_h);
	if (s->stream) (void)hipStreamDestroy(s->stream);
	delete s;
	return rc;
}

}  // extern "C"
