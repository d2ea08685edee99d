/*
 * gwfa.h — drop-in public header of the MI355X GWFA graph edit-distance library
 * (libgwfa_hip.so).
 *
 * The types and the first three entry points keep the names, field order and
 * memory layout of the reference header (gwfa.h:6-95), so C/C++ host code
 * written against the reference compiles and links unchanged.  Differences:
 *   - the header is plain C99/C++11: no __device__, __align__ or u_int64_t;
 *   - gwf_ed() is host-callable and runs on the GPU (the reference declares
 *     it __device__, gwfa.h:95); `buf` and `diag_shared` may be NULL;
 *   - batch, device-graph, device-read-set and serving entry points are added
 *     (they have no counterpart in the reference, whose batching lives inline
 *     in main.cu:454-735).
 * All arithmetic is exact integer.  Every entry point that can fail returns a
 * GWF_* status code; nothing is dropped silently.
 */
#ifndef GWFA_H
#define GWFA_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
#define GWF_RESTRICT __restrict
#else
#define GWF_RESTRICT restrict
#endif
#define GWF_ALIGN16 __attribute__((aligned(16)))

/* ---- types: identical layout to the reference (gwfa.h:6-80) ------------- */

typedef struct {          /* gwfa.h:6-9   16 B */
	uint64_t a;           /* source vertex << 32 | target vertex */
	int32_t o;            /* overlap: first offset used in the target */
} gwf_arc_t;

typedef struct {          /* gwfa.h:11-18 48 B */
	uint32_t n_vtx, n_arc;
	uint32_t *len;        /* vertex lengths */
	uint32_t *src;        /* oriented GFA segment id of each vertex */
	char **seq;           /* NUL-terminated vertex sequences */
	gwf_arc_t *arc;       /* arcs, sorted by `a` after gwf_ed_index() */
	uint64_t *aux;        /* CSR index: aux[v] = first_arc << 32 | n_arcs */
} gwf_graph_t;

typedef struct {          /* gwfa.h:20-25 */
	uint64_t *GWF_RESTRICT vd;
	int32_t *GWF_RESTRICT k;
	uint32_t *GWF_RESTRICT xo;
} gwf_diag_soa_t;

typedef struct {          /* gwfa.h:27-31 16 B */
	uint64_t vd;          /* vertex << 32 | (diagonal + 0x40000000) */
	int32_t k;            /* last matched offset in the vertex (-1: none) */
	uint32_t xo;          /* anti-diagonal << 1 | out-of-order flag */
} gwf_diag_t;

typedef struct {          /* gwfa.h:33-36 16 B */
	uint64_t vd0;
	uint64_t vd1;
} gwf_intv_t;

typedef struct GWF_ALIGN16 { /* gwfa.h:38-41 */
	uint64_t *GWF_RESTRICT vd0;
	uint64_t *GWF_RESTRICT vd1;
} gwf_intv_soa_t;

typedef struct {          /* gwfa.h:44-52 28 B */
	int32_t n_intv;
	int32_t n_tmp;
	int32_t n_swap;
	int32_t n_ooo;
	int32_t n_A;
	int32_t n_B;
	int32_t n_a;
} counters;

typedef struct {          /* gwfa.h:54-80 192 B; accepted and ignored by gwf_ed() */
	gwf_intv_soa_t intv;
	int32_t max_intv;
	gwf_intv_soa_t tmp;
	int32_t max_tmp;
	gwf_intv_soa_t swap;
	int32_t max_swap;
	gwf_diag_t *ooo;
	int32_t max_ooo;
	gwf_diag_soa_t a_buf;
	int32_t max_a;
	gwf_diag_t *B;
	int32_t max_B;
	gwf_diag_t *GWF_RESTRICT A;
	int32_t max_A;
	int32_t head_A;
	int32_t tail_A;
	uint64_t *GWF_RESTRICT ha;
	uint32_t ha_mask;
} gwf_edbuf_t;

/* ---- status codes -------------------------------------------------------- */

/* -1 is never an error code: it is gwf_ed()'s "no end reachable" result. */
#define GWF_OK             0
#define GWF_ERR_ARG       -2   /* invalid argument (NULL pointer, bad vertex id, ...) */
#define GWF_ERR_HIP       -3   /* HIP runtime error; see gwf_last_error() */
#define GWF_ERR_NOMEM     -4   /* host or device allocation failed */
#define GWF_ERR_NODEV     -5   /* no HIP device visible */
#define GWF_ERR_CAPACITY  -6   /* a read needs more wavefront state than the hard cap */
#define GWF_ERR_STATE     -7   /* call not valid in the current state (e.g. server stopped) */
#define GWF_ERR_BUDGET    -8   /* a request can never fit the serving state budget */

/* ED result value of a read whose alignment has no end (only when v1 >= 0). */
#define GWF_ED_NONE       -1

#ifdef __cplusplus
extern "C" {
#endif

/* ---- reference entry points (gwfa.h:87-88, 95) --------------------------- */

/* Build the CSR arc index: sorts g->arc by `a` and allocates g->aux
 * (main.cu:242-270).  Host only; not thread-safe on the same graph. */
void gwf_ed_index(gwf_graph_t *g);

/* Free g->aux (main.cu:272-277) and drop any device copy of g cached by
 * gwf_ed(). */
void gwf_cleanup(gwf_graph_t *g);

/* Edit distance of query q[0..ql) against g, starting at offset 0 of vertex
 * v0 and ending anywhere (v1 < 0) or at the last base of vertex v1 (v1 >= 0).
 * Same arguments and result as the reference device routine (gwfa.h:95,
 * gwf-ed.cu:1292-1337): the ED, or -1 when v1 >= 0 is unreachable.
 * max_lag is accepted and ignored, as in the reference (gwf-ed.cu:1286-1287).
 * buf and diag_shared may be NULL (they are ignored: state is library-owned).
 * Host-callable; runs on the current HIP device.  The device copy of g is
 * cached per graph pointer until gwf_cleanup(g).  Returns a negative GWF_ERR_*
 * code (< -1) on failure. */
int32_t gwf_ed(const gwf_graph_t *g, int32_t ql, const char *q, int32_t v0, int32_t v1,
               uint32_t max_lag, gwf_edbuf_t *buf, uint64_t *diag_shared);

/* ---- additions: device-resident graph and reads, batch alignment ---------- */
/* These have no counterpart in the reference API; they replace its inline
 * batch runtime (main.cu:417-452 graph upload, main.cu:454-735 stream loop). */

typedef struct gwf_dgraph_s gwf_dgraph_t;   /* packed CSR graph resident in HBM */
typedef struct gwf_dreads_s gwf_dreads_t;   /* packed read set resident in HBM */

/* Kernel selection for gwf_ed_batch / gwf_ed_run. */
#define GWF_KERNEL_PLAIN       0  /* one 64-lane wavefront per read, grid = reads */
#define GWF_KERNEL_PERSISTENT  1  /* persistent workgroups, cost-ordered read queue */

typedef struct {
	double   kernel_ms;        /* device time of the alignment launches (HIP events) */
	double   wall_ms;          /* host wall time of the call */
	int64_t  n_launch;         /* alignment kernel launches */
	int64_t  n_retry;          /* reads re-run with a larger state capacity */
	int64_t  sum_steps;        /* sum over reads of score steps (ED + 1) */
	int64_t  sum_states;       /* sum over reads and steps of wavefront size */
	int64_t  sum_chars;        /* characters compared by extension */
	int64_t  n_spill;          /* reads whose state left LDS for HBM */
	int64_t  sum_hbm_steps;    /* of sum_steps, those run with the state in HBM */
} gwf_run_stats_t;

/* Upload an indexed graph (gwf_ed_index() done) to the current device. */
int gwf_dgraph_upload(const gwf_graph_t *g, gwf_dgraph_t **out);
void gwf_dgraph_free(gwf_dgraph_t *dg);

/* Upload n reads: seqs + off[i] holds read i of length ql[i]; v0[i] is its
 * start vertex (NULL: all 0). */
int gwf_dreads_upload(int32_t n, const char *seqs, const int64_t *off, const int32_t *ql,
                      const int32_t *v0, gwf_dreads_t **out);
void gwf_dreads_free(gwf_dreads_t *dr);

/* A number unique to this upload within the process (never 0 and never
 * reused, unlike the handle's address after a free).  Captured launches are
 * replayed only for the same pair of generations. */
uint64_t gwf_dgraph_generation(const gwf_dgraph_t *dg);
uint64_t gwf_dreads_generation(const gwf_dreads_t *dr);

/* Align every read of dr against dg; out[i] (host memory) receives read i's
 * ED.  stats may be NULL.  Reads whose state outgrows the working capacity
 * are re-run with a larger one; GWF_ERR_CAPACITY only past the hard cap. */
int gwf_ed_run(gwf_dgraph_t *dg, gwf_dreads_t *dr, int32_t v1, uint32_t max_lag,
               int32_t kernel, int32_t *out, gwf_run_stats_t *stats);

/* gwf_ed_run plus per-read device timestamps: t_start_ns[i] and t_end_ns[i]
 * (either may be NULL) receive the time read i started and wrote its ED, in
 * ns since the first read of the run started (device wall clock).  A read
 * re-run after a workspace overflow reports its final run. */
int gwf_ed_run_timed(gwf_dgraph_t *dg, gwf_dreads_t *dr, int32_t v1, uint32_t max_lag,
                     int32_t kernel, int32_t *out, gwf_run_stats_t *stats,
                     int64_t *t_start_ns, int64_t *t_end_ns);

/* Convenience: upload reads from host pointers, run, free.  q[i] points at
 * read i (ql[i] bytes); v0 may be NULL. */
int gwf_ed_batch(gwf_dgraph_t *dg, int32_t n, const char *const *q, const int32_t *ql,
                 const int32_t *v0, int32_t v1, uint32_t max_lag, int32_t kernel,
                 int32_t *out, gwf_run_stats_t *stats);

/* ---- addition: an entry point that allocates nothing --------------------- */
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory where they align (tight serving loops, harnesses that forbid
 * hipMalloc): gwf_ed_run on the plain kernel with every device buffer the call
 * needs (EDs, statuses, re-run lists, every read's workspace) taken from one
 * static device array of GWF_NOALLOC_SCRATCH_BYTES per device, held until the
 * process ends.  dg and dr are uploaded beforehand (set-up, which allocates).
 * Reads go in input order in batches of equal workspaces; a read that
 * outgrows its workspace is re-run with a 4x larger one.  Returns
 * GWF_ERR_CAPACITY for more than GWF_NOALLOC_MAX_READS reads or a read whose
 * workspace cannot fit the array.  Synchronous; out is host memory.  Calls
 * in one process are serialised (they share the array, which keeps the last
 * call's contents).  stats: kernel_ms is the call's wall time, the state and
 * character counters are 0. */
#define GWF_NOALLOC_SCRATCH_BYTES (128ull << 20)
#define GWF_NOALLOC_MAX_READS (1 << 20)
int gwf_ed_run_noalloc(gwf_dgraph_t *dg, gwf_dreads_t *dr, int32_t v1, uint32_t max_lag, int32_t *out,
                       gwf_run_stats_t *stats);

/* Text of the last error seen by this thread ("" if none). */
const char *gwf_last_error(void);

/* ---- additions: length/error bins and their tuning ------------------------ */
/* GWF_KERNEL_PERSISTENT runs entirely on the device: every read's first 256
 * bases are aligned (a probe), the read's predicted wavefront size
 * (~3.2 x predicted ED + 16, predicted ED = ql x probe ED / probe length)
 * selects a bin, and each bin's reads are aligned longest-first by one
 * persistent launch configured from the table below.  Steady-state runs
 * (same graph, reads and table) replay a captured hipGraph (GWF_GRAPH=0
 * disables capture).  No counterpart in the reference. */

#define GWF_TUNE_MAX_BINS 8
typedef struct {
	int32_t est_max;     /* reads with predicted wavefront size <= est_max (first match) */
	int32_t nw;          /* 64-lane waves per read: 1, 2 or 4 */
	int32_t lds_states;  /* LDS wavefront capacity per read (larger spills to HBM) */
	int32_t merge_runs;  /* sort variant: <= merge_runs sorted runs of 64 are merged in one
	                        rank pass, more by pairwise merging (0: always pairwise) */
} gwf_tune_bin_t;
typedef struct {
	int32_t n_bins;
	gwf_tune_bin_t bin[GWF_TUNE_MAX_BINS];
} gwf_tune_t;

int gwf_tune_get(gwf_tune_t *out);
/* Replace the process-wide table (NULL: the built-in one).  est_max must
 * increase from bin to bin; the last bin takes every larger read. */
int gwf_tune_set(const gwf_tune_t *t);
/* Text file: lines "bin <est_max> <nw> <lds_states> <merge_runs>".  The
 * table in GWF_TUNE_FILE, if set, is loaded at first use. */
int gwf_tune_save(const char *path);
int gwf_tune_load(const char *path);

/* The autotuner: times complete binned runs of dr (the workload, or a
 * sample large enough that no single read dominates) under candidate
 * tables: first every configuration (waves per read x LDS capacity) as a
 * single bin; then the built-in bin boundaries with that configuration and,
 * costliest bin first, each other configuration for that bin alone (kept
 * only if the binned table beats the single bin again on a second run);
 * then the sort variant.  A change is kept when it is >= 2% faster.  At most
 * max_evals runs (0: 64).  Installs the best table and copies it to out (may
 * be NULL); report (may be NULL) receives one line per run. */
int gwf_autotune(gwf_dgraph_t *dg, gwf_dreads_t *dr, int32_t max_evals, gwf_tune_t *out,
                 char *report, int64_t report_len);

/* ---- additions: serving (continuous batching) ----------------------------- */
/* No counterpart in the reference, whose driver aligns one fixed batch
 * (main.cu:454-750).  A server keeps the device graph resident and runs a
 * device loop that claims requests as they are published: a request that
 * arrives while others are aligning is admitted as soon as a ring slot and
 * state budget are free, and its alignment overlaps theirs on the GPU.  Each
 * request reserves its predicted state footprint (its read's device copy plus
 * the HBM workspace its length calls for) from a fixed device budget until
 * its result is returned.  A request whose workspace overflows is re-admitted
 * with a 4x larger one; results are bit-identical to gwf_ed_run's. */

typedef struct gwf_serve_s gwf_serve_t;

typedef struct {
	int64_t state_budget_bytes; /* device bytes for admitted requests (0: 1 GiB) */
	int64_t staging_bytes;      /* pinned host bytes for admitted reads (0: 64 MiB) */
	int32_t ring_slots;         /* most requests on the device at once, power of two (0: 4096) */
	int32_t nw;                 /* waves per read: 1, 2 or 4 (0: 4) */
	int32_t lds_states;         /* LDS wavefront capacity per read (0: 1024) */
	int32_t wgs;                /* resident workgroups (0: occupancy x CUs) */
	int32_t park_ms;            /* idle time after which the device loop is parked (0: 20) */
	int32_t watchdog_ms;        /* heartbeat silence that counts a busy workgroup as stalled (0: 10000) */
} gwf_serve_opts_t;

typedef struct {
	int64_t id;            /* as returned by gwf_serve_submit */
	int32_t ed;            /* edit distance (GWF_ED_NONE: no end) */
	int32_t status;        /* GWF_OK, GWF_ERR_BUDGET or GWF_ERR_CAPACITY */
	int64_t t_arrival_ns;  /* gwf_serve_clock_ns() at which the request became visible */
	int64_t t_admit_ns;    /* published to the device (last attempt) */
	int64_t t_done_ns;     /* result seen by the server thread */
	int64_t service_ns;    /* device time from claim to result (last attempt) */
	int64_t steps;         /* score steps (ED + 1) */
	int64_t states;        /* wavefront states processed */
	int32_t attempts;      /* 1 + re-admissions after a workspace overflow */
	int32_t spilled;       /* 1 if the state left LDS for HBM */
} gwf_serve_result_t;

typedef struct {
	int64_t submitted, completed, queued, inflight;
	int64_t readmitted;    /* re-admissions after workspace overflow */
	int64_t rejected;      /* requests that can never fit the budget */
	int64_t launches;      /* device-loop launches (relaunched after parking) */
	int64_t budget_bytes, budget_used, budget_peak;
	int64_t inflight_peak;
	int64_t stalled;       /* workgroups whose heartbeat went silent past watchdog_ms */
	int64_t bases_done;
} gwf_serve_stats_t;

/* Steady clock (ns) used for every t_* field and for release times. */
int64_t gwf_serve_clock_ns(void);

/* Start a server on the current device for dg (which must outlive it). */
int gwf_serve_start(gwf_dgraph_t *dg, const gwf_serve_opts_t *opts, gwf_serve_t **out);

/* Queue n requests: read i is seqs + off[i], ql[i] bytes, start vertex v0[i]
 * (v0 NULL: 0).  t_release_ns[i] (NULL: now) is the gwf_serve_clock_ns() time
 * from which request i is visible to admission (a load generator's arrival
 * time).  ids[i] (may be NULL) receives each request's id.  Copies the bytes. */
int gwf_serve_submit(gwf_serve_t *s, int32_t n, const char *seqs, const int64_t *off, const int32_t *ql,
                     const int32_t *v0, const int64_t *t_release_ns, int64_t *ids);

/* gwf_serve_submit with an end vertex per request, as gwf_ed's v1 argument
 * (gwfa.h:95 of the reference): v1[i] >= 0 ends request i at the last base
 * of vertex v1[i] (gwf-ed.cu:1231-1254), v1[i] < 0 anywhere; v1 NULL: every
 * request ends anywhere.  A request with no reachable end gets GWF_ED_NONE. */
int gwf_serve_submit_v1(gwf_serve_t *s, int32_t n, const char *seqs, const int64_t *off, const int32_t *ql,
                        const int32_t *v0, const int32_t *v1, const int64_t *t_release_ns, int64_t *ids);

/* Move up to max finished results into out, waiting up to wait_ms for the
 * first one.  Returns the number moved (>= 0) or a GWF_ERR_* code. */
int gwf_serve_poll(gwf_serve_t *s, gwf_serve_result_t *out, int32_t max, int32_t wait_ms);

/* Wait until every submitted request has a result (timeout_ms < 0: no limit);
 * GWF_ERR_STATE on timeout. */
int gwf_serve_drain(gwf_serve_t *s, int32_t timeout_ms);

int gwf_serve_stats(gwf_serve_t *s, gwf_serve_stats_t *out);

/* Drop requests not yet admitted, wait for those on the device, stop the
 * device loop and free the server. */
int gwf_serve_stop(gwf_serve_t *s);

#ifdef __cplusplus
}
#endif

#endif /* GWFA_H */
