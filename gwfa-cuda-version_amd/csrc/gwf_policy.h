// gwf_policy.h — host-side sizing policy shared by the batch runtime and the
// server: workspace capacities per read length, LDS capacities, error text.
#pragma once
#include "gwf_dev.h"
#include "gwfa.h"

// Bytes readable past the end of a packed sequence or a read's device copy
// (extension loads run up to 512 + 16 bytes past the last compared byte).
constexpr size_t kSeqPad = 2048;

int64_t gwf_env_int(const char *name, int64_t dflt);
// HBM workspace capacities of a read of length ql, times `scale` after overflows.
Caps gwf_caps_for(i32 ql, int scale);
// LDS capacities for a wavefront of up to ~m states.
Caps gwf_lds_caps(i32 m);
// Set this thread's gwf_last_error() text.
void gwf_set_error(const char *fmt, ...);

// Device graph handle (gwf_dgraph_upload).
struct gwf_dgraph_s {
	DGraph g;
	void *mem = nullptr;
	u32 n_vtx = 0;
};
