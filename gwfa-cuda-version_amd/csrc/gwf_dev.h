// gwf_dev.h — types shared by the HIP kernels and the host runtime.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

typedef uint64_t u64;
typedef uint32_t u32;
typedef int32_t i32;

// A wavefront state (vertex v, diagonal d, offset k) is packed into one u64
// of three fields, from the top: v (bv bits), d + 2^(bd-1) (bd bits) and
// 2^bk - 2 - k (bk bits, -1 <= k <= 2^bk - 2), so ascending keys sort by
// vertex, then diagonal, then k descending; the first key of every (v, d)
// group is the furthest-reaching state.  The top bv + bd bits ("diag id")
// identify (v, d); consecutive diagonals of one vertex have consecutive diag
// ids, which is what the forbidden-interval history stores.  The kernels are
// built for GWF_N_LAYOUTS splits; a run uses the first that fits its graph and
// reads (gwf_pick_layout): v must stay below 2^bv - 1 (all ones is the empty
// key's), -2^(bd-1) < d < 2^(bd-1) and -1 <= k <= 2^bk - 2; d = i - k runs
// from -(vertex length) to read length + 1.
struct GwfKeyLayout {
	i32 bv, bd, bk;
};
#define GWF_N_LAYOUTS 4
__host__ __device__ constexpr GwfKeyLayout gwf_key_layout(int l) {
	return l == 1 ? GwfKeyLayout{28, 18, 18}    // large graphs
	     : l == 2 ? GwfKeyLayout{20, 22, 22}    // long reads and segments
	     : l == 3 ? GwfKeyLayout{32, 16, 16}    // any 32-bit vertex id
	              : GwfKeyLayout{24, 20, 20};
}
// Largest vertex count, vertex length and read length layout l represents.
inline uint64_t gwf_layout_max_vtx(int l) { return (1ull << gwf_key_layout(l).bv) - 2; }
inline u32 gwf_layout_max_vlen(int l) {
	const GwfKeyLayout k = gwf_key_layout(l);
	const u32 a = (1u << (k.bd - 1)) - 1, b = (1u << k.bk) - 2;
	return a < b ? a : b;
}
inline i32 gwf_layout_max_ql(int l) { return (1 << (gwf_key_layout(l).bd - 1)) - 2; }
// The first layout that fits, or -1.
inline int gwf_pick_layout(uint64_t n_vtx, u32 max_vlen, i32 max_ql) {
	for (int l = 0; l < GWF_N_LAYOUTS; ++l)
		if (n_vtx <= gwf_layout_max_vtx(l) && max_vlen <= gwf_layout_max_vlen(l) && max_ql <= gwf_layout_max_ql(l))
			return l;
	return -1;
}
// Bounds over all layouts (what an upload accepts before a run picks one).
#define GWF_MAX_VLEN ((1 << 21) - 1)
#define GWF_MAX_QLEN ((1 << 21) - 2)

// Device graph (packed CSR).  Two 16-byte records per vertex: vinfo[2 v]:
// x = byte offset of v's sequence in `seq`, y = length, z = first out-arc,
// w = out-degree; vinfo[2 v + 1]: the first two out-arcs' (x, y) as in arc[]
// (zero where v has fewer), so the closure crosses most vertices without a
// second dependent load.  arc[j] carries what crossing the arc needs in one
// 16-byte load: x = target w, y = overlap o | GWF_ARC_SINGLE_IN when this arc
// is w's only in-arc | (base w[o] << 24), z = w's sequence offset, w = w's
// length.  (o < 2^21, below the flag bit.)
#define GWF_ARC_SINGLE_IN (1u << 23)
#define GWF_ARC_OL_MASK 0x7FFFFFu
struct DGraph {
	const uint4 *vinfo;
	const uint8_t *seq;
	const uint4 *arc;
	u32 n_vtx;
	u32 total_len;  // sum of vertex lengths
	u32 mixed;      // 1: some vertex has in-arcs with different overlaps (per-read visited set)
	u32 layout;     // key layout of the kernels a launch runs (gwf_key_layout)
};

struct DReads {
	const uint8_t *seq;   // packed reads, padded past the end
	const int64_t *off;   // byte offset of read i
	const i32 *ql;
	const i32 *v0;
	i32 n;
	i32 v1;               // end vertex of every read (-1: end anywhere)
};

// Element capacities of one read's working arrays (in HBM or in LDS).
struct Caps {
	i32 m;   // wavefront P and successor/merge buffer Q
	i32 x;   // closure successors (a power of two when LDS-resident)
	i32 f;   // closure frontier, each of two buffers
	i32 t;   // exits recorded in one step
	i32 h;   // forbidden-interval history, each of two buffers
	i32 v;   // per-read visited set, slots (a power of two; 0: none, the per-step LDS set is used)
};

// Slots of the per-read visited set of a workspace whose other capacities
// derive from a (Caps {2a, a, a, a, a}): graphs with mixed in-overlaps only.
__host__ __device__ inline i32 gwf_vis_slots(i32 a, bool mixed) {
	if (!mixed) return 0;
	i32 v = 1024;
	while (v < 4 * a && v < (1 << 27)) v <<= 1;
	return v;
}

// Per-read counters written by the kernels: score steps, wavefront states
// processed, characters compared, device wall-clock ticks at start and end,
// and the step at which the read left LDS for HBM (-1: never).
#define GWF_NCOUNT 6

// Per-read result status.
#define GWF_ST_OK        0
#define GWF_ST_OVERFLOW  1   // an HBM workspace capacity was exceeded; re-run larger
#define GWF_ST_LIMIT     2   // score passed ql + 1 (cannot happen for valid input)

// Launch parameters of the persistent kernel for one bin of reads.
struct PersistCfg {
	i32 nw;          // waves per read (1, 2 or 4)
	i32 wgs;         // resident workgroups (grid size)
	Caps lds;        // LDS-resident capacities
	Caps hbm;        // HBM spill capacities
	size_t lds_bytes;  // dynamic LDS: the LDS-resident arrays, grown to the occupancy's share (HBM-mode staging)
	i32 mw;          // LDS sort: at most mw runs of 64 are merged by one rank pass
};

size_t gwf_ws_bytes(const Caps &c);
size_t gwf_lds_bytes(const Caps &c);

// Plain kernel: one workgroup (one wave) per read_idx[0..n).  Read j's
// workspace is scratch + ws_off[j] (u64 elements) with the capacities of
// ws_a[j] (Caps {2a, a, a, a, a}): every read gets a workspace of its own size.
hipError_t gwf_launch_plain(const DGraph &g, const DReads &r, const i32 *read_idx, i32 n, const u64 *ws_off,
                            const i32 *ws_a, u64 *scratch, i32 *out_ed, i32 *out_status, int64_t *out_counts,
                            hipStream_t st);
// The same kernel with equal workspaces: block j aligns read read_idx[j]
// (read_idx == nullptr: read base + j) in scratch + j * gwf_ws_bytes(c) / 8,
// c = Caps {2a, a, a, a, a}.  out_counts may be null.
hipError_t gwf_launch_plain_uniform(const DGraph &g, const DReads &r, const i32 *read_idx, i32 base, i32 n, i32 a,
                                    u64 *scratch, i32 *out_ed, i32 *out_status, int64_t *out_counts, hipStream_t st);

// Persistent kernel: cfg.wgs workgroups pull the reads of bin `bin`,
// order[bin_off[bin] .. bin_off[bin + 1]) (device arrays; order == nullptr:
// the identity), through the device counter heads[bin] (zero at launch)
// until the queue is empty.  beat (may be null): one GwfBeat per workgroup
// in host-coherent memory.
struct GwfBeat;
hipError_t gwf_launch_persistent(const DGraph &g, const DReads &r, const i32 *order, const i32 *bin_off, i32 bin,
                                 i32 *heads, const PersistCfg &cfg, u64 *scratch, size_t ws_bytes, i32 *out_ed,
                                 i32 *out_status, int64_t *out_counts, GwfBeat *beat, hipStream_t st);

// ------------------------------------------------------------ device-side binning
// Tuning bins: a read goes to the first bin whose est_max is at least its
// predicted wavefront size (~3.2 x predicted ED + 16; the predicted ED is
// the read's length times the error rate its probe prefix showed).
#define GWF_MAX_BINS 8
#define GWF_COST_BUCKETS 128  // GWF_MAX_BINS x this <= 1024 (one scan workgroup)
struct BinBounds {
	i32 n_bins;
	i32 est_max[GWF_MAX_BINS];
};
// Order reads for the binned persistent launches, entirely on the device (so
// the sequence can be captured in a graph): from the probe's EDs, every
// read's bin and cost (ql + 4 ED^2) are computed; a counting sort over
// (bin, descending log-cost bucket) fills order[] with each bin's reads,
// longest first, and bin_off[0..n_bins] with the bins' boundaries.  hist
// holds n_bins * GWF_COST_BUCKETS counters and key n entries of scratch;
// heads[0..GWF_MAX_BINS) is cleared for the launches that follow.
hipError_t gwf_launch_binning(const i32 *probe_ed, const i32 *probe_status, const i32 *ql, i32 n, const BinBounds &bb,
                              i32 *key, i32 *hist, i32 *order, i32 *bin_off, i32 *heads, hipStream_t st);

// Largest dynamic LDS a workgroup may use on this device.
size_t gwf_max_lds_bytes();

// Resident persistent workgroups per CU for nw waves and lds_bytes of LDS.
int gwf_persist_occupancy(int nw, size_t lds_bytes);
size_t gwf_persist_grow_lds(int nw, size_t lds_bytes, int per_cu, size_t margin);

// ------------------------------------------------------------ heartbeat
// Progress record of one workgroup, written by the kernel into host-coherent
// memory every GWF_BEAT_EVERY score steps (and when it picks up a read), so a
// host thread can tell a slow read from a stuck one while the kernel runs.
#define GWF_BEAT_EVERY 16
struct GwfBeat {
	u64 beat;     // increases with every record
	i32 read;     // read index (batch) or request sequence number (serving); -1 idle
	i32 step;     // its current score step
	i32 n_a;      // its current wavefront size
	i32 in_hbm;   // 1 once its state left LDS
	u64 pad;
};

// ------------------------------------------------------------ serving
// A request published to the device: the host writes the entry and the
// read's bytes (host-coherent memory), then advances SrvCtl::word.  Entry
// seq lives at ring[seq & mask] until a workgroup claims it; the host keeps
// fewer than mask + 1 requests in flight, so it never overwrites an
// unclaimed entry.  The result goes to res[res_slot], a slot the host owns
// until it has read the result.
struct SrvEntry {
	u64 q_off;     // offset of the read's bytes in the staging arena (16-aligned,
	               // zero-filled to the next 16 bytes and one more 16-byte word)
	u64 d_off;     // offset of the request's region in the device pool: the
	               // read's copy (kSeqPad past it), then its HBM workspace
	u64 ws_off;    // offset of the HBM workspace inside the region
	i32 ql, v0;
	Caps hc;       // HBM workspace capacities
	u32 res_slot;
	i32 v1;        // end vertex (< 0: anywhere)
};
struct SrvResult {
	i32 ed, status;
	int64_t steps, states, chars, t0, t1, spill_step;
	int64_t t_claim;  // device clock when a workgroup claimed the request
	i32 wg;           // that workgroup
	i32 polls;        // empty polls it made before the claim
	u64 done;         // request sequence + 1 once the fields above are final
};
// word = tail (requests published) | stop << 63
struct SrvCtl {
	u64 word;
	u64 pad[7];
};
#define GWF_SRV_STOP (1ull << 63)

struct SrvArgs {
	i32 mw;                 // LDS sort variant (see PersistCfg)
	DGraph g;
	const SrvEntry *ring;   // host-coherent, mask + 1 entries
	SrvResult *res;         // host-coherent, mask + 1 results
	const SrvCtl *ctl;      // host-coherent
	u64 *claimed;           // device: next claim ticket (= requests claimed while no ticket is held)
	const uint8_t *stage;   // host-coherent staging arena
	uint8_t *pool;          // device pool
	GwfBeat *beat;          // host-coherent, one per workgroup (may be null)
	u64 mask;
	Caps lc;                // LDS capacities
	i32 idle_polls;         // a workgroup exits after this many empty polls
};

// Device server: cfg.wgs workgroups claim published requests until the stop
// bit is set (and no request is left) or they have polled idle_polls times.
hipError_t gwf_launch_serve(const SrvArgs &a, const PersistCfg &cfg, hipStream_t st);

// Resident server workgroups per CU for nw waves and lds_bytes of LDS.
int gwf_serve_occupancy(int nw, size_t lds_bytes);

// p[0..n) = v (a kernel, so it can be part of a captured graph).
hipError_t gwf_launch_fill(i32 *p, i32 n, i32 v, hipStream_t st);
// Set the dynamic-LDS limits of the persistent and server kernels; call once
// per device before the first launch (and before any graph capture).
hipError_t gwf_kernels_init();
