/* ref_gfa_dump.c — test infrastructure (never shipped, never on the GPU path).
 *
 * Reads a GFA with the reference's own GFA layer (gfa_read, gfa_name2id and
 * gfa_sub_from from gfa-io.c, gfa-base.c and gfa-sub.c, compiled from the
 * reference sources by `make ref`) and prints the graph the reference's
 * gwf_gfa2gwf (main.cu:294-345) builds from it, in the format of the
 * reference's `gwf-test -p` (main.cu:354-359):
 *   S <i> <seq of vertex i> LN:i:<len>     vertex i = i-th vertex of the subgraph
 *                                          reachable from segment <start>, forward
 *                                          strand; '-' strands reverse-complemented
 *   L <i> + <j> + <overlap>M               one per arc, in the subgraph's order
 * Usage: gfa_dump <graph.gfa> [start-segment-name]
 * The loader under test (gwf_gfa_load) must produce the same vertices and arcs. */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include "gfa.h"
#include "gfa-priv.h"

int main(int argc, char *argv[])
{
	gfa_t *gfa;
	gfa_sub_t *sub;
	uint32_t v0 = 0, i, j;
	int32_t l;
	if (argc < 2) { fprintf(stderr, "usage: gfa_dump <graph.gfa> [start]\n"); return 1; }
	gfa = gfa_read(argv[1]);
	if (gfa == 0) { fprintf(stderr, "gfa_dump: cannot read %s\n", argv[1]); return 1; }
	if (argc > 2) {
		int32_t sid = gfa_name2id(gfa, argv[2]);
		if (sid >= 0) v0 = (uint32_t)sid << 1;
		else fprintf(stderr, "gfa_dump: no segment '%s', using the first\n", argv[2]);
	}
	sub = gfa_sub_from(gfa, v0, 1 << 30);
	for (i = 0; i < (uint32_t)sub->n_v; ++i) {
		uint32_t v = sub->v[i].v;
		const gfa_seg_t *s = &gfa->seg[v >> 1];
		printf("S\t%u\t", i);
		for (j = 0; j < (uint32_t)s->len; ++j)
			putchar(v & 1 ? gfa_comp_table[(uint8_t)s->seq[s->len - 1 - j]] : s->seq[j]);
		printf("\tLN:i:%d\n", s->len);
	}
	for (i = 0; i < (uint32_t)sub->n_v; ++i)
		for (l = 0; l < sub->v[i].n; ++l) {
			uint64_t a = sub->a[sub->v[i].off + l];
			printf("L\t%u\t+\t%u\t+\t%dM\n", i, (uint32_t)(a >> 32), gfa->arc[(uint32_t)a].ow);
		}
	free(sub->a);  /* gfa_sub_destroy is declared but not defined in gfa-sub.c */
	free(sub->v);
	free(sub);
	gfa_destroy(gfa);
	return 0;
}
