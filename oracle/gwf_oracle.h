/*
 * gwf_oracle.h — TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of the reference's GWFA edit distance (gwf-ed.cu:80-1337)
 * and of its graph index (main.cu:242-277), plus an independent brute-force
 * dynamic-programming checker.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * or the timed CPU baseline; the product never links or calls it.
 */
#ifndef GWF_ORACLE_H
#define GWF_ORACLE_H

#include <stdint.h>
#include "../include/gwfa.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
	int64_t steps;     /* score iterations run (ED + 1 when an end is found) */
	int64_t states;    /* W: sum over iterations of the wavefront size n_a */
	int64_t chars;     /* C: graph/query character pairs compared by extension */
	int64_t max_na;    /* largest wavefront seen */
	int64_t a_events;  /* states processed by the A-queue closure */
	int64_t crossings; /* arc traversals in the closure */
	int64_t vis_lossy; /* visited-set claims refused to an arc of another overlap (not in the reference) */
} oracle_stats_t;

/* gwf_ed with exact semantics (signed max_k, growable buffers, no drops).
 * Returns the ED or -1 when no end is reachable (v1 >= 0).  st may be NULL;
 * when given it is accumulated into, not cleared. */
int32_t oracle_gwf_ed(const gwf_graph_t *g, int32_t ql, const char *q, int32_t v0, int32_t v1,
                      uint32_t max_lag, oracle_stats_t *st);

/* Brute-force DP edit distance over graph positions (any graph, cycles too);
 * same start/end semantics as oracle_gwf_ed.  O(total_len * ql). */
int32_t oracle_dp_ed(const gwf_graph_t *g, int32_t ql, const char *q, int32_t v0, int32_t v1);

/* Restatement of gwf_ed_index (main.cu:246-270) on a caller-owned graph:
 * sorts g->arc by `a` and fills aux_out[n_vtx] (caller-allocated, zeroed here). */
void oracle_ed_index(gwf_graph_t *g, uint64_t *aux_out);

/* Multi-threaded batch runner (the timed CPU baseline): reads i in [0,n) are
 * q = seqs + off[i], length ql[i], start v0[i] (v0 may be NULL).  out[i] gets
 * the ED.  n_threads <= 0 means all online cores.  Returns the threads used. */
int oracle_batch(const gwf_graph_t *g, int32_t n, const char *seqs, const int64_t *off,
                 const int32_t *ql, const int32_t *v0, int32_t v1, int32_t *out,
                 int n_threads, oracle_stats_t *st);

#ifdef __cplusplus
}
#endif

#endif
