/*
 * gwfa_host.h — host-side input and synthetic-data library (libgwfa_host.so).
 *
 * Pure C++ (no HIP): the GFA and FASTA/FASTQ readers that feed the alignment
 * library, the graph printer of the reference CLI, and a seeded generator of
 * the benchmark graphs and reads.  Memory returned through out-pointers is
 * released with gwf_free() (graphs) or gwf_host_free() (arrays).
 */
#ifndef GWFA_HOST_H
#define GWFA_HOST_H

#include <stdint.h>
#include "gwfa.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Read a GFA (S/L lines; overlaps as `*`, CIGAR or `ov:ow`; +/- strands) or a
 * FASTA file (one segment per record) and extract the subgraph reachable from
 * the forward strand of segment `start` (NULL: the first segment).  Vertex 0
 * is the start; vertices follow a topological order when the subgraph is a
 * DAG; `-` strands are reverse-complemented.  This is gfa_read + gwf_gfa2gwf
 * (gfa-io.c:354, main.cu:294-345).  The graph is not indexed: call
 * gwf_ed_index() before aligning.  Returns GWF_OK or a GWF_ERR_* code. */
int gwf_gfa_load(const char *path, const char *start, gwf_graph_t **out);

/* Free a graph from gwf_gfa_load / gwf_synth_graph (main.cu:347-352); does
 * not free g->aux (gwf_cleanup does). */
void gwf_free(gwf_graph_t *g);

/* Build a graph from arrays: vertex v has len[v] bytes starting at seq +
 * sum(len[0..v)); arc i goes src[i] -> dst[i] with overlap ovl[i] (NULL: 0).
 * Not indexed. */
int gwf_graph_build(uint32_t n_vtx, const char *seq, const uint32_t *len, uint32_t n_arc,
                    const uint32_t *src, const uint32_t *dst, const int32_t *ovl, gwf_graph_t **out);

/* Print the graph as the reference's `gwf-test -p` (main.cu:354-359). */
int gwf_graph_print_file(const gwf_graph_t *g, const char *path);

/* Write the graph as GFA 1 (S and L lines, forward strands, `<o>M` links). */
int gwf_graph_write_gfa(const gwf_graph_t *g, const char *path);

/* Read all records of a FASTA/FASTQ file, gzip or plain, with kseq
 * semantics (kseq.h:192-232): multi-line sequences, '\r' stripped, bytes
 * kept as-is.  seqs holds the reads back to back, read i at off[i], length
 * ql[i], each followed by a NUL. */
int gwf_fasta_load(const char *path, int32_t *n, char **seqs, int64_t **off, int32_t **ql);

void gwf_host_free(void *p);

/* ---- synthetic graphs ---------------------------------------------------- */

typedef struct {
	int64_t  backbone_len;   /* uniform-random ACGT backbone length */
	int32_t  seg_min;        /* segment lengths between bubbles, uniform [seg_min, seg_max] */
	int32_t  seg_max;
	int32_t  bubble_every;   /* mean bubble spacing (bp), uniform in [0.5x, 1.5x] */
	double   p_snp;          /* fraction of simple bubbles that are SNPs (rest: indels) */
	int32_t  indel_max;      /* indel allele length, uniform [1, indel_max] */
	double   p_multi;        /* fraction of bubbles with 3-4 alleles */
	double   p_nested;       /* probability that an allele >= 8 bp holds a nested SNP bubble */
	double   p_tandem;       /* fraction of bubbles whose alleles are tandem repeats */
	int32_t  tr_motif_min, tr_motif_max;   /* tandem-repeat motif length range */
	int32_t  tr_copy_min, tr_copy_max;     /* tandem-repeat copy-number range */
	uint64_t seed;
} gwf_synth_graph_opts_t;

/* Build a bubble graph; vertex 0 is the source and vertices are in
 * topological order.  Not indexed. */
int gwf_synth_graph(const gwf_synth_graph_opts_t *o, gwf_graph_t **out);

#define GWF_SYNTH_LEN_FIXED      0  /* every read len_a bases */
#define GWF_SYNTH_LEN_LOGNORMAL  1  /* median len_a, sigma len_sigma, clipped to [len_min, len_max] */

#define GWF_SYNTH_ERR_CYCLE      0  /* read i gets err_levels[i % n_err_levels] */
#define GWF_SYNTH_ERR_UNIFORM    1  /* uniform on [err_lo, err_hi] */

#define GWF_SYNTH_START_SOURCE   0  /* every read starts at offset 0 of vertex 0 */
#define GWF_SYNTH_START_RANDOM   1  /* offset 0 of a random vertex on a random path; v0[i] set */

typedef struct {
	int32_t  n_reads;
	int32_t  len_mode;
	int32_t  len_a;
	double   len_sigma;
	int32_t  len_min, len_max;
	int32_t  err_mode;
	int32_t  n_err_levels;
	double   err_levels[8];
	double   err_lo, err_hi;
	int32_t  start_mode;
	uint64_t seed;
} gwf_synth_reads_opts_t;

/* Sample reads along random paths with sub/ins/del errors in equal parts.
 * Outputs as gwf_fasta_load plus v0[i] (start vertex) and err[i] (the error
 * rate used).  g need not be indexed. */
int gwf_synth_reads(const gwf_graph_t *g, const gwf_synth_reads_opts_t *o, int32_t *n,
                    char **seqs, int64_t **off, int32_t **ql, int32_t **v0, float **err);

#ifdef __cplusplus
}
#endif

#endif
