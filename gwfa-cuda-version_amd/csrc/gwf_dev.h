// gwf_dev.h — types shared by the HIP kernels and the host runtime.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

typedef uint64_t u64;
typedef uint32_t u32;
typedef int32_t i32;

// A wavefront state (vertex v, diagonal d, offset k) is packed into one u64:
//   bits 63..40  v                     (n_vtx  < 2^24)
//   bits 39..20  d + GWF_D_BIAS        (|d|    < 2^19)
//   bits 19..0   GWF_KINV_MAX - k      (-1 <= k <= GWF_KINV_MAX)
// so ascending keys sort by vertex, then diagonal, then k descending; the
// first key of every (v, d) group is the furthest-reaching state.  The top 44
// bits ("diag id") identify (v, d); consecutive diagonals of one vertex have
// consecutive diag ids, which is what the forbidden-interval history stores.
#define GWF_D_BIAS   (1 << 19)
#define GWF_KINV_MAX ((1 << 20) - 2)
#define GWF_MAX_VTX  (1u << 24)
#define GWF_MAX_VLEN ((1 << 19) - 1)
#define GWF_MAX_QLEN ((1 << 19) - 2)

// Device graph (packed CSR).  vinfo[v]: x = byte offset of v's sequence in
// `seq`, y = length, z = first out-arc, w = out-degree.  arc[j] carries what
// crossing the arc needs in one 16-byte load: x = target w, y = overlap o |
// (base w[o] << 24), z = w's sequence offset, w = w's length.
struct DGraph {
	const uint4 *vinfo;
	const uint8_t *seq;
	const uint4 *arc;
	u32 n_vtx;
};

struct DReads {
	const uint8_t *seq;   // packed reads, padded past the end
	const int64_t *off;   // byte offset of read i
	const i32 *ql;
	const i32 *v0;
	i32 n;
};

// Element capacities of one read's working arrays (in HBM or in LDS).
struct Caps {
	i32 m;   // wavefront P and successor/merge buffer Q
	i32 x;   // closure successors (a power of two when LDS-resident)
	i32 f;   // closure frontier, each of two buffers
	i32 t;   // exits recorded in one step
	i32 h;   // forbidden-interval history, each of two buffers
};

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
	size_t lds_bytes;
};

size_t gwf_ws_bytes(const Caps &c);
size_t gwf_lds_bytes(const Caps &c);

hipError_t gwf_launch_plain(const DGraph &g, const DReads &r, const i32 *read_idx, i32 n, const Caps &caps,
                            u64 *scratch, size_t ws_bytes, i32 *out_ed, i32 *out_status, int64_t *out_counts,
                            hipStream_t st);

// Persistent kernel: cfg.wgs workgroups pull reads order[head..n) through the
// device counter *head until the queue is empty.
hipError_t gwf_launch_persistent(const DGraph &g, const DReads &r, const i32 *order, i32 n, i32 *head,
                                 const PersistCfg &cfg, u64 *scratch, size_t ws_bytes, i32 *out_ed,
                                 i32 *out_status, int64_t *out_counts, hipStream_t st);

// Largest dynamic LDS a workgroup may use on this device.
size_t gwf_max_lds_bytes();

// Resident persistent workgroups per CU for nw waves and lds_bytes of LDS.
int gwf_persist_occupancy(int nw, size_t lds_bytes);
