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

// Per-vertex record of the device graph: x = byte offset of the sequence in
// the packed sequence buffer, y = length, z = first out-arc, w = out-degree.
struct DGraph {
	const uint4 *vinfo;
	const uint8_t *seq;
	const u32 *arc_w;
	const i32 *arc_o;
	u32 n_vtx;
};

struct DReads {
	const uint8_t *seq;   // packed reads, 16-byte padded
	const int64_t *off;   // byte offset of read i
	const i32 *ql;
	const i32 *v0;
	i32 n;
};

// Element capacities of one read's workspace.
struct Caps {
	i32 a;   // wavefront (A and the compacted output)
	i32 s;   // in-vertex successors (<= 3 * a)
	i32 x;   // closure successors
	i32 f;   // closure frontier (each of two buffers)
	i32 t;   // exits recorded this step
	i32 h;   // forbidden-interval history (each of two buffers)
};

// Per-read counters written by the kernels: score steps, wavefront states
// processed, characters compared, device wall-clock ticks at start and end.
#define GWF_NCOUNT 5

// Per-read result status.
#define GWF_ST_OK        0
#define GWF_ST_OVERFLOW  1   // a workspace capacity was exceeded; re-run larger
#define GWF_ST_LIMIT     2   // score passed ql + 1 (cannot happen for valid input)

// Kernel launch entry points (defined in gwf_kernels.hip).
hipError_t gwf_launch_plain(const DGraph &g, const DReads &r, const i32 *read_idx, i32 n,
                            const Caps &caps, u64 *scratch, size_t ws_bytes, i32 *out_ed,
                            i32 *out_status, int64_t *out_counts, hipStream_t st);
size_t gwf_ws_bytes(const Caps &c);
