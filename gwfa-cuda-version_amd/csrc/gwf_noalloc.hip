// gwf_noalloc.hip — gwf_ed_run_noalloc: the plain-kernel gwf_ed_run with
// every device buffer a call needs (EDs, statuses, the lists of reads to
// re-run, every read's HBM workspace) carved from one file-scope device array,
// so that a call allocates nothing.  For callers that cannot allocate device
// memory in their loop.
//
// Each device holds its own copy of the array for the life of the process;
// host code reaches it only through the device address hipGetSymbolAddress
// gives for the current device (the array's name in host code is a host-side
// stand-in, not a device address).  Calls are serialised by a process-wide
// mutex: they all share the array of their device, which keeps the last
// call's contents.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <mutex>

#include "gwf_dev.h"
#include "gwf_policy.h"
#include "gwfa.h"

static_assert(GWF_NOALLOC_SCRATCH_BYTES % 256 == 0, "scratch is whole 256-byte blocks");
static __device__ __align__(256) u64 scratch[GWF_NOALLOC_SCRATCH_BYTES / 8];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
u64 *g_scratch[kMaxDevices];  // device address of `scratch` on each device (null: not looked up yet)

// Reads of the list src[0..n) (src == nullptr: reads lo .. lo + n) whose
// status is an overflow are appended to dst; hdr[0] counts them and hdr[1]
// receives their longest length.
__global__ __launch_bounds__(256) void gwf_noalloc_collect(const i32 *status, const i32 *ql, const i32 *src, i32 lo,
                                                          i32 n, i32 *dst, i32 *hdr) {
	const i32 j = (i32)(blockIdx.x * 256 + threadIdx.x);
	if (j >= n) return;
	const i32 r = src ? src[j] : lo + j;
	if (status[r] == GWF_ST_OVERFLOW) {
		dst[atomicAdd(&hdr[0], 1)] = r;
		atomicMax(&hdr[1], ql[r]);
	}
}

hipError_t collect(const i32 *status, const i32 *ql, const i32 *src, i32 lo, i32 n, i32 *dst, i32 *hdr, i32 out[2]) {
	hipError_t e = gwf_launch_fill(hdr, 2, 0, nullptr);
	if (e == hipSuccess && n > 0) {
		hipLaunchKernelGGL(gwf_noalloc_collect, dim3((n + 255) / 256), dim3(256), 0, nullptr, status, ql, src, lo, n,
		                   dst, hdr);
		e = hipGetLastError();
	}
	if (e == hipSuccess) e = hipMemcpy(out, hdr, 2 * sizeof(i32), hipMemcpyDeviceToHost);
	return e;
}

}  // namespace

extern "C" int gwf_ed_run_noalloc(gwf_dgraph_t *dg, gwf_dreads_t *dr, int32_t v1, uint32_t max_lag, int32_t *out,
                                  gwf_run_stats_t *stats) {
	(void)max_lag;  // accepted and ignored, as in the reference (gwf-ed.cu:1286-1287)
	const auto t0 = std::chrono::steady_clock::now();
	gwf_run_stats_t st;
	memset(&st, 0, sizeof(st));
	if (!dg || !dr || (!out && dr->r.n > 0)) { gwf_set_error("gwf_ed_run_noalloc: NULL argument"); return GWF_ERR_ARG; }
	if (v1 >= 0 && (u32)v1 >= dg->n_vtx) { gwf_set_error("gwf_ed_run_noalloc: v1 %d >= n_vtx %u", v1, dg->n_vtx); return GWF_ERR_ARG; }
	const i32 n = dr->r.n;
	if (n > GWF_NOALLOC_MAX_READS) {
		gwf_set_error("gwf_ed_run_noalloc: %d reads; at most %d per call", n, GWF_NOALLOC_MAX_READS);
		return GWF_ERR_CAPACITY;
	}
	for (i32 i = 0; i < n; ++i)
		if ((u32)dr->v0[i] >= dg->n_vtx) {
			gwf_set_error("gwf_ed_run_noalloc: read %d has v0 %d >= n_vtx %u", i, dr->v0[i], dg->n_vtx);
			return GWF_ERR_ARG;
		}
	if (n == 0) { if (stats) *stats = st; return GWF_OK; }
	DGraph g;
	if (int rc = gwf_run_graph(dg, dr, "gwf_ed_run_noalloc", &g)) return rc;

	std::lock_guard<std::mutex> lk(g_mu);
	int dev = 0;
	if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
		gwf_set_error("gwf_ed_run_noalloc: no usable HIP device");
		return GWF_ERR_NODEV;
	}
	if (!g_scratch[dev]) {  // once per device, with that device current
		void *p = nullptr;
		const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
		if (e != hipSuccess || !p) {
			gwf_set_error("gwf_ed_run_noalloc: hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
			return GWF_ERR_HIP;
		}
		g_scratch[dev] = (u64 *)p;
	}
	// scratch layout: EDs, statuses, two lists of reads to re-run (n each),
	// the list header, then the workspaces
	uint8_t *base = (uint8_t *)g_scratch[dev];
	const size_t ln = ((size_t)n * 4 + 255) / 256 * 256;
	i32 *ed = (i32 *)base, *status = (i32 *)(base + ln);
	i32 *list[2] = {(i32 *)(base + 2 * ln), (i32 *)(base + 3 * ln)};
	i32 *hdr = (i32 *)(base + 4 * ln);
	u64 *ws = (u64 *)(base + 4 * ln + 256);
	const size_t ws_bytes = GWF_NOALLOC_SCRATCH_BYTES - (4 * ln + 256);

	DReads rv = dr->r;
	rv.v1 = v1;
	const std::vector<i32> &qlh = dr->ql;
	int rc = GWF_OK;
	hipError_t e = hipSuccess;
	// pass 1: the reads in input order, in batches of equal workspaces sized
	// for the batch's longest read that fit the workspace region
	for (i32 pos = 0; pos < n && e == hipSuccess;) {
		i32 end = pos, maxql = 0;
		while (end < n) {
			const i32 q = std::max(maxql, qlh[end]);
			if (end > pos && gwf_ws_bytes(gwf_caps_for(q, 1, g.mixed != 0)) * (size_t)(end - pos + 1) > ws_bytes) break;
			maxql = q, ++end;
		}
		const Caps c = gwf_caps_for(maxql, 1, g.mixed != 0);
		if (gwf_ws_bytes(c) > ws_bytes) {
			gwf_set_error("gwf_ed_run_noalloc: read %d (length %d) needs a %zu-byte workspace; the scratch holds %zu",
			              pos, qlh[pos], gwf_ws_bytes(c), ws_bytes);
			return GWF_ERR_CAPACITY;
		}
		e = gwf_launch_plain_uniform(g, rv, nullptr, pos, end - pos, c.x, ws, ed, status, nullptr, nullptr);
		st.n_launch += 1;
		pos = end;
	}
	// re-runs of the reads that outgrew their workspace, 4x larger each time
	i32 h[2] = {0, 0};
	if (e == hipSuccess) e = collect(status, dr->r.ql, nullptr, 0, n, list[0], hdr, h);
	int cur = 0, scale = 1;
	while (e == hipSuccess && h[0] > 0) {
		st.n_retry += h[0];
		scale *= 4;
		const Caps c = gwf_caps_for(h[1], scale, g.mixed != 0);
		const size_t wb = gwf_ws_bytes(c);
		if (wb > ws_bytes || (c.x >= (1 << 25) && scale > (1 << 12))) {
			gwf_set_error("gwf_ed_run_noalloc: %d reads need workspaces of %zu bytes; the scratch holds %zu", h[0], wb,
			              ws_bytes);
			return GWF_ERR_CAPACITY;
		}
		const i32 per = (i32)std::min<size_t>(ws_bytes / wb, (size_t)h[0]);
		for (i32 o = 0; o < h[0] && e == hipSuccess; o += per) {
			e = gwf_launch_plain_uniform(g, rv, list[cur] + o, 0, std::min(per, h[0] - o), c.x, ws, ed, status,
			                             nullptr, nullptr);
			st.n_launch += 1;
		}
		const i32 cnt = h[0];
		if (e == hipSuccess) e = collect(status, dr->r.ql, list[cur], 0, cnt, list[cur ^ 1], hdr, h);
		cur ^= 1;
	}
	// every read must have ended normally; `out` holds the statuses first
	if (e == hipSuccess) e = hipMemcpy(out, status, 4 * (size_t)n, hipMemcpyDeviceToHost);
	for (i32 i = 0; e == hipSuccess && i < n; ++i)
		if (out[i] != GWF_ST_OK) {
			gwf_set_error("gwf_ed_run_noalloc: read %d ended with internal status %d", i, out[i]);
			return GWF_ERR_STATE;
		}
	if (e == hipSuccess) e = hipMemcpy(out, ed, 4 * (size_t)n, hipMemcpyDeviceToHost);
	if (e != hipSuccess) {
		gwf_set_error("gwf_ed_run_noalloc: %s", hipGetErrorString(e));
		rc = GWF_ERR_HIP;
	}
	st.wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
	st.kernel_ms = st.wall_ms;
	if (stats) *stats = st;
	return rc;
}
