// gwf-test — command-line driver with the reference's interface
// (main.cu:361-750): align every read of a FASTA/FASTQ(.gz) file against a
// GFA graph and print one edit distance per line.
//
//   gwf-test [options] <target.gfa> <query.fa>
//     -s STR   starting segment name [first]
//     -l INT   max lag (accepted and ignored, as in the reference)
//     -p       print the graph in the forward strand and exit
//     -k STR   kernel: persistent (default) or plain
//     -T FILE  load a bin tuning table (gwf_tune_save format)
//
// stdout, as the reference prints it: "\n righe: N\n" once the reads are
// read, the N edit distances one per line, then
// "\nTempo GPU streams: <seconds> secondi\n" (reads upload + alignment +
// copy-back, graph load and index excluded).
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <getopt.h>
#include <vector>

#include "gwfa.h"
#include "gwfa_host.h"

static void usage() {
	fprintf(stderr, "Usage: gwf-test [options] <target.gfa|fa> <query.fa>\n");
	fprintf(stderr, "Options:\n");
	fprintf(stderr, "  -l INT    max lag behind the furthest wavefront; 0 to disable [0]\n");
	fprintf(stderr, "  -s STR    starting segment name [first]\n");
	fprintf(stderr, "  -p        output GFA in the forward strand\n");
	fprintf(stderr, "  -k STR    kernel: persistent or plain [persistent]\n");
	fprintf(stderr, "  -T FILE   bin tuning table [built-in]\n");
}

int main(int argc, char *argv[]) {
	int c, print_graph = 0, kernel = GWF_KERNEL_PERSISTENT;
	const char *sname = nullptr, *tune = nullptr;
	uint32_t max_lag = 0;
	while ((c = getopt(argc, argv, "ptl:s:k:T:")) >= 0) {
		if (c == 'p') print_graph = 1;
		else if (c == 'l') max_lag = (uint32_t)atoi(optarg);
		else if (c == 's') sname = optarg;
		else if (c == 'T') tune = optarg;
		else if (c == 'k') {
			if (!strcmp(optarg, "plain")) kernel = GWF_KERNEL_PLAIN;
			else if (!strcmp(optarg, "persistent")) kernel = GWF_KERNEL_PERSISTENT;
			else { usage(); return 1; }
		} else if (c != 't') { usage(); return 1; }
	}
	if ((!print_graph && argc - optind < 2) || (print_graph && argc == optind)) {
		usage();
		return 1;
	}
	gwf_graph_t *g = nullptr;
	int rc = gwf_gfa_load(argv[optind], sname, &g);
	if (rc == GWF_ERR_ARG && sname) {  // as the reference: report and keep the first segment
		fprintf(stderr, "ERROR: failed to find segment '%s'\n", sname);
		rc = gwf_gfa_load(argv[optind], nullptr, &g);
	}
	if (rc != GWF_OK) {
		fprintf(stderr, "ERROR: failed to read graph %s (%d)\n", argv[optind], rc);
		return 1;
	}
	if (print_graph) {
		gwf_graph_print_file(g, "-");
		gwf_free(g);
		return 0;
	}
	gwf_ed_index(g);
	if (tune && gwf_tune_load(tune) != GWF_OK) {
		fprintf(stderr, "ERROR: %s\n", gwf_last_error());
		return 1;
	}
	gwf_dgraph_t *dg = nullptr;
	if ((rc = gwf_dgraph_upload(g, &dg)) != GWF_OK) {
		fprintf(stderr, "ERROR: %s\n", gwf_last_error());
		return 1;
	}
	int32_t n = 0;
	char *seqs = nullptr;
	int64_t *off = nullptr;
	int32_t *ql = nullptr;
	if (gwf_fasta_load(argv[optind + 1], &n, &seqs, &off, &ql) != GWF_OK) {
		fprintf(stderr, "ERROR: failed to read %s\n", argv[optind + 1]);
		return 1;
	}
	printf("\n righe: %d\n", n);
	std::vector<int32_t> ed((size_t)(n > 0 ? n : 1));
	const auto t0 = std::chrono::steady_clock::now();
	gwf_dreads_t *dr = nullptr;
	rc = gwf_dreads_upload(n, seqs, off, ql, nullptr, &dr);
	if (rc == GWF_OK) rc = gwf_ed_run(dg, dr, -1, max_lag, kernel, ed.data(), nullptr);
	const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
	if (rc != GWF_OK) {
		fprintf(stderr, "ERROR: %s\n", gwf_last_error());
		return 1;
	}
	for (int32_t i = 0; i < n; ++i) printf("%d\n", ed[i]);
	printf("\nTempo GPU streams: %.6f secondi\n", secs);
	gwf_dreads_free(dr);
	gwf_dgraph_free(dg);
	gwf_host_free(seqs);
	gwf_host_free(off);
	gwf_host_free(ql);
	gwf_cleanup(g);
	gwf_free(g);
	return 0;
}
