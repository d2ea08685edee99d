// gwf_policy.h — host-side sizing policy shared by the batch runtime and the
// server: workspace capacities per read length, LDS capacities, error text.
#pragma once
#include <vector>

#include "gwf_dev.h"
#include "gwfa.h"

// Bytes readable past the end of a packed sequence or a read's device copy
// (extension loads run up to 512 + 16 bytes past the last compared byte).
constexpr size_t kSeqPad = 2048;

int64_t gwf_env_int(const char *name, int64_t dflt);
// HBM workspace capacities of a read of length ql, times `scale` after
// overflows; `mixed`: the graph has vertices with in-arcs of different
// overlaps (the workspace then holds the read's visited set too).
Caps gwf_caps_for(i32 ql, int scale, bool mixed);
// LDS capacities for a wavefront of up to ~m states.
Caps gwf_lds_caps(i32 m);
// Set this thread's gwf_last_error() text.
void gwf_set_error(const char *fmt, ...);

// A number unique to every upload in the process (never 0, never reused), so
// that state derived from a handle (a captured graph's device pointers) can
// tell a handle from a later one that the allocator placed at the same host
// address.
uint64_t gwf_next_generation();

// Device graph handle (gwf_dgraph_upload).
struct gwf_dgraph_s {
	DGraph g;
	void *mem = nullptr;
	u32 n_vtx = 0;
	u32 max_vlen = 0;
	uint64_t gen = 0;
};

// Device read-set handle (gwf_dreads_upload); ql and v0 are host copies.
struct gwf_dreads_s {
	DReads r;
	const i32 *ql_probe = nullptr;  // device: min(ql, probe length) per read
	void *mem = nullptr;
	uint64_t gen = 0;
	std::vector<i32> ql, v0;
};

// The device graph for a run of dr on dg: dg's, with the first state layout
// that holds dg's vertex count and vertex lengths and dr's read lengths
// (gwf_pick_layout); GWF_ERR_ARG (with `what` in the error text) if none does.
int gwf_run_graph(const gwf_dgraph_t *dg, const gwf_dreads_t *dr, const char *what, DGraph *out);
