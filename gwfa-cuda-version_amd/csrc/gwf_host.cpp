// gwf_host.cpp — libgwfa_host.so: GFA/FASTA readers, graph printers and the
// seeded synthetic graph/read generator.  Host C++ only (no HIP).
//
// The GFA reader restates the reference's GFA layer step by step, so that the
// gwf graph (vertex order, lengths, sequences and arcs) is the one the
// reference builds: S/L/FASTA line parsing (gfa_parse_S, gfa_parse_L,
// gfa-io.c:115-336, 354-427), gfa_finalize (gfa-base.c:455-466: segments
// without a sequence dropped, arcs sorted, semi-specified overlaps filled,
// missing complement arcs added, overlaps clamped and arc lengths set, deleted
// arcs removed), the subgraph extraction gfa_sub_from (gfa-sub.c:41-272: a
// Kahn-style order driven by an ordered set keyed on (forced flag,
// unresolved in-arcs, distance, vertex)) and the conversion gwf_gfa2gwf
// (main.cu:294-345: '-' strands reverse-complemented, arc overlap = the
// target-side overlap).  Multi-arcs are kept, as there.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <cmath>
#include <string>
#include <vector>
#include <unordered_map>
#include <algorithm>
#include <deque>
#include <memory>
#include <set>
#include <zlib.h>
#include "../../include/gwfa.h"
#include "../../include/gwfa_host.h"

namespace {

// ---------------------------------------------------------------- RNG
struct Rng {  // xoshiro256** seeded by splitmix64
	uint64_t s[4];
	explicit Rng(uint64_t seed) {
		uint64_t z = seed;
		for (int i = 0; i < 4; ++i) {
			z += 0x9e3779b97f4a7c15ULL;
			uint64_t x = z;
			x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
			x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
			s[i] = x ^ (x >> 31);
		}
	}
	static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
	uint64_t next() {
		const uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
		s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3]; s[2] ^= t; s[3] = rotl(s[3], 45);
		return r;
	}
	double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
	uint64_t below(uint64_t n) { return n ? next() % n : 0; }
	int64_t range(int64_t lo, int64_t hi) { return lo + (int64_t)below((uint64_t)(hi - lo + 1)); }
	double normal() {  // Box-Muller
		double u1 = uniform(), u2 = uniform();
		if (u1 < 1e-300) u1 = 1e-300;
		return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
	}
};

const char kBases[4] = {'A', 'C', 'G', 'T'};

char other_base(Rng &r, char c) {
	char x;
	do x = kBases[r.below(4)]; while (x == c);
	return x;
}

std::string random_seq(Rng &r, int64_t n) {
	std::string s((size_t)n, 'A');
	for (auto &c : s) c = kBases[r.below(4)];
	return s;
}

// ------------------------------------------------------- graph assembly
struct Builder {
	std::vector<std::string> seq;
	std::vector<uint32_t> src;
	std::vector<std::pair<uint32_t, uint32_t>> arcs;  // (from, to)
	std::vector<int32_t> ovl;
	uint32_t add(std::string s, uint32_t srcid = 0) {
		seq.push_back(std::move(s));
		src.push_back(srcid);
		return (uint32_t)seq.size() - 1;
	}
	void link(uint32_t u, uint32_t w, int32_t o = 0) { arcs.push_back({u, w}); ovl.push_back(o); }
	void link_all(const std::vector<uint32_t> &from, uint32_t w) {
		for (uint32_t u : from) link(u, w);
	}
	gwf_graph_t *finish() const {
		gwf_graph_t *g = (gwf_graph_t *)calloc(1, sizeof(gwf_graph_t));
		if (!g) return nullptr;
		g->n_vtx = (uint32_t)seq.size();
		g->n_arc = (uint32_t)arcs.size();
		g->len = (uint32_t *)malloc(sizeof(uint32_t) * (g->n_vtx + 1));
		g->src = (uint32_t *)malloc(sizeof(uint32_t) * (g->n_vtx + 1));
		g->seq = (char **)malloc(sizeof(char *) * (g->n_vtx + 1));
		g->arc = (gwf_arc_t *)malloc(sizeof(gwf_arc_t) * (g->n_arc + 1));
		for (uint32_t v = 0; v < g->n_vtx; ++v) {
			g->len[v] = (uint32_t)seq[v].size();
			g->src[v] = src[v];
			g->seq[v] = (char *)malloc(seq[v].size() + 1);
			memcpy(g->seq[v], seq[v].data(), seq[v].size());
			g->seq[v][seq[v].size()] = 0;
		}
		for (uint32_t i = 0; i < g->n_arc; ++i) {
			g->arc[i].a = (uint64_t)arcs[i].first << 32 | arcs[i].second;
			g->arc[i].o = ovl[i];
		}
		g->aux = nullptr;
		return g;
	}
};

// ------------------------------------------------------- synthetic graph
struct Synth {
	const gwf_synth_graph_opts_t &o;
	Rng r;
	Builder b;
	std::string bb;
	Synth(const gwf_synth_graph_opts_t &opt) : o(opt), r(opt.seed) {}

	// one allele as a chain (possibly with a nested SNP bubble); returns tails
	std::vector<uint32_t> allele(const std::string &s, const std::vector<uint32_t> &from) {
		if (s.empty()) return from;
		if (s.size() >= 8 && o.p_nested > 0 && r.uniform() < o.p_nested) {
			size_t a = 1 + (size_t)r.below(s.size() - 2);
			uint32_t p = b.add(s.substr(0, a));
			b.link_all(from, p);
			uint32_t x = b.add(std::string(1, s[a]));
			uint32_t y = b.add(std::string(1, other_base(r, s[a])));
			b.link(p, x), b.link(p, y);
			uint32_t z = b.add(s.substr(a + 1));
			b.link(x, z), b.link(y, z);
			return {z};
		}
		uint32_t v = b.add(s);
		b.link_all(from, v);
		return {v};
	}

	// a bubble at backbone position pos; returns bases of backbone consumed
	int64_t bubble(int64_t pos, std::vector<uint32_t> &tails) {
		std::vector<std::string> al;
		int64_t used = 0;
		const int64_t left = (int64_t)bb.size() - pos;
		if (o.p_tandem > 0 && r.uniform() < o.p_tandem) {
			std::string motif = random_seq(r, r.range(o.tr_motif_min, o.tr_motif_max));
			int64_t c0 = r.range(o.tr_copy_min, o.tr_copy_max), c1;
			do c1 = r.range(o.tr_copy_min, o.tr_copy_max); while (c1 == c0 && o.tr_copy_max > o.tr_copy_min);
			std::string a0, a1;
			for (int64_t i = 0; i < c0; ++i) a0 += motif;
			for (int64_t i = 0; i < c1; ++i) a1 += motif;
			al = {a0, a1};
		} else {
			const bool multi = o.p_multi > 0 && r.uniform() < o.p_multi;
			const bool snp = r.uniform() < o.p_snp;
			if (snp) {
				al.push_back(std::string(1, bb[pos]));
				used = 1;
				int na = multi ? (int)r.range(3, 4) : 2;
				std::string seen(1, bb[pos]);
				while ((int)al.size() < na) {
					char c = kBases[r.below(4)];
					if (seen.find(c) != std::string::npos) continue;
					seen += c;
					al.push_back(std::string(1, c));
				}
			} else {
				int64_t m = r.range(1, std::max(1, o.indel_max));
				if (m > left - 1) m = std::max<int64_t>(0, left - 1);
				bool del = r.uniform() < 0.5 && m > 0;
				if (del) {
					al.push_back(bb.substr(pos, m));
					used = m;
					al.push_back(std::string());
				} else {
					al.push_back(std::string());
					al.push_back(random_seq(r, std::max<int64_t>(1, m)));
				}
				if (multi) {
					int na = (int)r.range(3, 4);
					while ((int)al.size() < na)
						al.push_back(random_seq(r, r.range(1, std::max(1, o.indel_max))));
				}
			}
		}
		std::vector<uint32_t> nt;
		bool pass = false;
		for (auto &s : al) {
			if (s.empty()) {
				if (pass) continue;  // at most one pass-through
				pass = true;
			}
			auto t = allele(s, tails);
			for (uint32_t x : t)
				if (std::find(nt.begin(), nt.end(), x) == nt.end()) nt.push_back(x);
		}
		tails = nt;
		return used;
	}

	gwf_graph_t *run() {
		bb = random_seq(r, o.backbone_len);
		std::vector<uint32_t> tails;
		int64_t pos = 0;
		const int64_t B = std::max(1, o.bubble_every);
		auto spacing = [&]() { return std::max<int64_t>(1, r.range(B / 2, B + B / 2)); };
		int64_t next_bubble = o.bubble_every > 0 ? spacing() : INT64_MAX;
		bool last_was_seg = false;
		const int64_t L = (int64_t)bb.size();
		while (pos < L) {
			if (pos < next_bubble || !last_was_seg || L - pos < 2) {
				int64_t len = r.range(std::max(1, o.seg_min), std::max(o.seg_min, o.seg_max));
				if (next_bubble > pos) len = std::min(len, next_bubble - pos);
				len = std::min(len, L - pos);
				uint32_t v = b.add(bb.substr(pos, len));
				b.link_all(tails, v);
				tails = {v};
				pos += len;
				last_was_seg = true;
			} else {
				pos += bubble(pos, tails);
				next_bubble = pos + spacing();
				last_was_seg = false;
			}
		}
		if (!last_was_seg) {  // end on a segment so the graph has one sink
			uint32_t v = b.add(random_seq(r, std::max(1, o.seg_min)));
			b.link_all(tails, v);
		}
		return b.finish();
	}
};

// ------------------------------------------------------- read sampling
struct Adj {
	std::vector<uint32_t> off, w;
	std::vector<int32_t> ol;
	explicit Adj(const gwf_graph_t *g) {
		off.assign(g->n_vtx + 1, 0);
		for (uint32_t i = 0; i < g->n_arc; ++i) ++off[(uint32_t)(g->arc[i].a >> 32) + 1];
		for (uint32_t v = 0; v < g->n_vtx; ++v) off[v + 1] += off[v];
		w.resize(g->n_arc), ol.resize(g->n_arc);
		std::vector<uint32_t> fill(off.begin(), off.end() - 1);
		for (uint32_t i = 0; i < g->n_arc; ++i) {
			uint32_t u = (uint32_t)(g->arc[i].a >> 32);
			w[fill[u]] = (uint32_t)g->arc[i].a, ol[fill[u]++] = g->arc[i].o;
		}
	}
};

bool sample_read(const gwf_graph_t *g, const Adj &adj, Rng &r, uint32_t v, int32_t L, double e,
                 std::string &out) {
	out.clear();
	int64_t p = 0;
	bool alive = g->len[v] > 0;
	auto advance = [&]() {
		if (++p >= (int64_t)g->len[v]) {
			uint32_t n = adj.off[v + 1] - adj.off[v];
			if (n == 0) { alive = false; return; }
			uint32_t j = adj.off[v] + (uint32_t)r.below(n);
			v = adj.w[j], p = adj.ol[j];
			if (p >= (int64_t)g->len[v]) alive = false;
		}
	};
	while ((int32_t)out.size() < L) {
		if (!alive) return false;
		const char c = g->seq[v][p];
		const double x = r.uniform();
		if (x < e / 3) { out += other_base(r, c); advance(); }
		else if (x < 2 * e / 3) out += kBases[r.below(4)];
		else if (x < e) advance();
		else { out += c; advance(); }
	}
	return true;
}

// ------------------------------------------------------- GFA parsing
// A segment and an arc as the reference's GFA layer holds them (gfa.h:33-67).
constexpr int32_t kNoOvl = INT32_MAX;  // overlap not given
struct GfaSeg {
	std::string name, seq;
	int64_t len = 0;
	bool has_seq = false, del = false;
};
struct GfaArc {
	uint32_t v, w;        // oriented vertices: segment << 1 | strand
	uint32_t lv = 0;      // len(v) - ov, set once overlaps are final
	int32_t ov, ow;       // overlap on v's side and on w's side
	uint32_t link = 0;    // pairs an arc with its complement
	bool del = false, comp = false;
	uint64_t key() const { return (uint64_t)v << 32 | lv; }
};

// The overlap field of an L line (gfa_parse_L, gfa-io.c:222-300): "*" is 0;
// "ov:ow" or ":ow"; a CIGAR sums M/D/N into ov and M/I/S into ow; anything
// else (or no field) gives 0.
void parse_overlap(const char *q, int32_t &ov, int32_t &ow) {
	ov = ow = 0;
	if (*q == '*' || *q == 0) return;
	if (*q == ':') {
		ov = kNoOvl;
		ow = isdigit((unsigned char)q[1]) ? (int32_t)strtol(q + 1, nullptr, 10) : kNoOvl;
		return;
	}
	if (!isdigit((unsigned char)*q)) return;
	char *r;
	const long x = strtol(q, &r, 10);
	if (isupper((unsigned char)*r)) {
		const char *p = q;
		do {
			char *e;
			const long l = strtol(p, &e, 10);
			p = e;
			if (*p == 'M' || *p == 'D' || *p == 'N') ov += (int32_t)l;
			if (*p == 'M' || *p == 'I' || *p == 'S') ow += (int32_t)l;
			++p;
		} while (isdigit((unsigned char)*p));
	} else if (*r == ':') {
		ov = (int32_t)x;
		ow = isdigit((unsigned char)r[1]) ? (int32_t)strtol(r + 1, nullptr, 10) : kNoOvl;
	}
}

struct LineReader {
	gzFile fp = nullptr;
	std::vector<char> buf;
	explicit LineReader(const char *path) {
		fp = strcmp(path, "-") ? gzopen(path, "r") : gzdopen(0, "r");
		if (fp) gzbuffer(fp, 1 << 20);
		buf.resize(1 << 16);
	}
	~LineReader() { if (fp) gzclose(fp); }
	bool ok() const { return fp != nullptr; }
	bool next(std::string &line) {  // strips "\n" and a trailing "\r"
		line.clear();
		for (;;) {
			if (!gzgets(fp, buf.data(), (int)buf.size())) return !line.empty();
			size_t n = strlen(buf.data());
			line.append(buf.data(), n);
			if (n && buf[n - 1] == '\n') break;
		}
		if (!line.empty() && line.back() == '\n') line.pop_back();
		if (line.size() > 1 && line.back() == '\r') line.pop_back();
		return true;
	}
};

void split_tabs(const std::string &s, std::vector<std::string> &f) {
	f.clear();
	size_t st = 0;
	for (size_t i = 0; i <= s.size(); ++i)
		if (i == s.size() || s[i] == '\t') { f.push_back(s.substr(st, i - st)); st = i + 1; }
}

const unsigned char *comp_table() {
	static unsigned char t[256];
	static bool init = false;
	if (!init) {
		for (int i = 0; i < 256; ++i) t[i] = (unsigned char)i;
		const char *a = "ACGTUMRWSYKVHDBN", *b = "TGCAAKYWSRMBDHVN";
		for (int i = 0; a[i]; ++i) {
			t[(unsigned char)a[i]] = (unsigned char)b[i];
			t[(unsigned char)tolower(a[i])] = (unsigned char)tolower(b[i]);
		}
		init = true;
	}
	return t;
}

}  // namespace

extern "C" {

void gwf_host_free(void *p) { free(p); }

void gwf_free(gwf_graph_t *g) {
	if (!g) return;
	if (g->seq)
		for (uint32_t i = 0; i < g->n_vtx; ++i) free(g->seq[i]);
	free(g->len); free(g->seq); free(g->arc); free(g->src); free(g);
}

int gwf_graph_build(uint32_t n_vtx, const char *seq, const uint32_t *len, uint32_t n_arc, const uint32_t *src,
                    const uint32_t *dst, const int32_t *ovl, gwf_graph_t **out) {
	if (!out || (n_vtx && (!seq || !len)) || (n_arc && (!src || !dst))) return GWF_ERR_ARG;
	Builder b;
	size_t o = 0;
	for (uint32_t v = 0; v < n_vtx; ++v) {
		b.add(std::string(seq + o, len[v]), v << 1);
		o += len[v];
	}
	for (uint32_t i = 0; i < n_arc; ++i) {
		if (src[i] >= n_vtx || dst[i] >= n_vtx) return GWF_ERR_ARG;
		b.link(src[i], dst[i], ovl ? ovl[i] : 0);
	}
	*out = b.finish();
	return *out ? GWF_OK : GWF_ERR_NOMEM;
}

int gwf_graph_print_file(const gwf_graph_t *g, const char *path) {
	if (!g || !path) return GWF_ERR_ARG;
	FILE *fp = strcmp(path, "-") ? fopen(path, "w") : stdout;
	if (!fp) return GWF_ERR_ARG;
	for (uint32_t i = 0; i < g->n_vtx; ++i) fprintf(fp, "S\t%d\t%s\tLN:i:%d\n", i, g->seq[i], g->len[i]);
	for (uint32_t i = 0; i < g->n_arc; ++i)
		fprintf(fp, "L\t%d\t+\t%d\t+\t%dM\n", (uint32_t)(g->arc[i].a >> 32), (uint32_t)g->arc[i].a, g->arc[i].o);
	if (fp != stdout) fclose(fp);
	else fflush(fp);
	return GWF_OK;
}

int gwf_graph_write_gfa(const gwf_graph_t *g, const char *path) {
	if (!g || !path) return GWF_ERR_ARG;
	FILE *fp = fopen(path, "w");
	if (!fp) return GWF_ERR_ARG;
	fprintf(fp, "H\tVN:Z:1.0\n");
	for (uint32_t i = 0; i < g->n_vtx; ++i) fprintf(fp, "S\ts%u\t%s\n", i, g->seq[i]);
	for (uint32_t i = 0; i < g->n_arc; ++i)
		fprintf(fp, "L\ts%u\t+\ts%u\t+\t%dM\n", (uint32_t)(g->arc[i].a >> 32), (uint32_t)g->arc[i].a, g->arc[i].o);
	fclose(fp);
	return GWF_OK;
}

int gwf_fasta_load(const char *path, int32_t *n_out, char **seqs_out, int64_t **off_out, int32_t **ql_out) {
	if (!path || !n_out || !seqs_out || !off_out || !ql_out) return GWF_ERR_ARG;
	gzFile fp = strcmp(path, "-") ? gzopen(path, "r") : gzdopen(0, "r");
	if (!fp) return GWF_ERR_ARG;
	gzbuffer(fp, 1 << 20);
	std::string all;
	std::vector<int64_t> off;
	std::vector<int32_t> ql;
	// kseq_read (kseq.h:192-232): jump to a header, skip name/comment line,
	// append sequence lines until '>', '+' or '@' at a line start.
	int c = gzgetc(fp);
	for (;;) {
		while (c != -1 && c != '>' && c != '@') c = gzgetc(fp);
		if (c == -1) break;
		while ((c = gzgetc(fp)) != -1 && c != '\n') {}  // header line
		const int64_t st = (int64_t)all.size();
		bool fastq = false;
		while ((c = gzgetc(fp)) != -1 && c != '>' && c != '+' && c != '@') {
			if (c == '\n') continue;
			std::string line(1, (char)c);
			while ((c = gzgetc(fp)) != -1 && c != '\n') line += (char)c;
			if (line.size() > 1 && line.back() == '\r') line.pop_back();
			all += line;
			if (c == -1) break;
		}
		const int64_t len = (int64_t)all.size() - st;
		if (c == '+') {
			fastq = true;
			while ((c = gzgetc(fp)) != -1 && c != '\n') {}
			int64_t got = 0;
			while (got < len && (c = gzgetc(fp)) != -1)
				if (c != '\n' && c != '\r') ++got;
			c = gzgetc(fp);
		}
		(void)fastq;
		if (len > INT32_MAX) { gzclose(fp); return GWF_ERR_ARG; }
		off.push_back(st);
		ql.push_back((int32_t)len);
		all += '\0';
	}
	gzclose(fp);
	const size_t n = ql.size();
	char *seqs = (char *)malloc(all.size() + 1);
	int64_t *o = (int64_t *)malloc(sizeof(int64_t) * (n + 1));
	int32_t *l = (int32_t *)malloc(sizeof(int32_t) * (n + 1));
	if (!seqs || !o || !l) { free(seqs); free(o); free(l); return GWF_ERR_NOMEM; }
	memcpy(seqs, all.data(), all.size());
	seqs[all.size()] = 0;
	for (size_t i = 0; i < n; ++i) o[i] = off[i], l[i] = ql[i];
	*n_out = (int32_t)n, *seqs_out = seqs, *off_out = o, *ql_out = l;
	return GWF_OK;
}

namespace {

// The graph as the reference's GFA layer leaves it after gfa_read.
struct GfaGraph {
	std::vector<GfaSeg> seg;
	std::unordered_map<std::string, uint32_t> id;
	std::vector<GfaArc> arc;
	std::vector<uint64_t> idx;  // per oriented vertex: first arc << 32 | count
	uint32_t add_seg(const std::string &nm) {  // gfa_add_seg
		auto it = id.find(nm);
		if (it != id.end()) return it->second;
		seg.emplace_back();
		seg.back().name = nm;
		id.emplace(nm, (uint32_t)seg.size() - 1);
		return (uint32_t)seg.size() - 1;
	}
	uint32_t n(uint32_t v) const { return (uint32_t)idx[v]; }
	uint32_t first(uint32_t v) const { return (uint32_t)(idx[v] >> 32); }
	void sort_index() {  // arcs by (v, lv); ties keep their order
		std::stable_sort(arc.begin(), arc.end(), [](const GfaArc &a, const GfaArc &b) { return a.key() < b.key(); });
		index();
	}
	void index() {
		idx.assign(seg.size() * 2, 0);
		size_t last = 0;
		for (size_t i = 1; i <= arc.size(); ++i)
			if (i == arc.size() || arc[i - 1].v != arc[i].v) {
				idx[arc[i - 1].v] = (uint64_t)last << 32 | (i - last);
				last = i;
			}
	}
	void finalize();
};

void GfaGraph::finalize() {  // gfa_finalize, gfa-base.c:455-466
	for (GfaSeg &s : seg) s.del = s.len == 0;  // gfa_fix_no_seg
	if (seg.empty()) return;
	sort_index();
	const uint32_t nv = (uint32_t)seg.size() * 2;
	// gfa_fix_semi_arc: an overlap given on one side only is taken from the
	// complement arc when there is exactly one consistent one; else dropped
	for (uint32_t v = 0; v < nv; ++v)
		for (uint32_t i = first(v), e = first(v) + n(v); i < e; ++i) {
			GfaArc &a = arc[i];
			if (a.del || (a.ow != kNoOvl && a.ov != kNoOvl)) continue;
			const uint32_t w = a.w ^ 1;
			int c = 0;
			int64_t jv = -1;
			for (uint32_t j = first(w), f = first(w) + n(w); j < f; ++j)
				if (!arc[j].del && arc[j].w == (v ^ 1)) ++c, jv = j;
			bool multi = false;
			if (c == 1) {
				if (a.ov != kNoOvl && arc[jv].ow != kNoOvl && a.ov != arc[jv].ow) multi = true;
				if (a.ow != kNoOvl && arc[jv].ov != kNoOvl && a.ow != arc[jv].ov) multi = true;
			}
			if (c == 1 && !multi) {
				if (arc[jv].ov != kNoOvl) a.ow = arc[jv].ov;
				if (arc[jv].ow != kNoOvl) a.ov = arc[jv].ow;
			} else {
				a.del = true;
			}
		}
	// gfa_fix_symm_add: every arc needs its complement w^1 -> v^1 with the
	// overlaps swapped; missing ones are appended (and not indexed until the
	// final sort, so later searches do not see them)
	for (uint32_t v = 0; v < nv; ++v)
		for (uint32_t i = first(v), e = first(v) + n(v); i < e; ++i) {
			if (arc[i].del || arc[i].comp) continue;
			const uint32_t w = arc[i].w ^ 1;
			bool found = false;
			for (uint32_t j = first(w), f = first(w) + n(w); j < f; ++j) {
				GfaArc &b = arc[j];
				if (b.del || b.comp) continue;
				if (b.w == (v ^ 1) && b.ov == arc[i].ow && b.ow == arc[i].ov) {
					b.comp = true;
					b.link = arc[i].link;
					found = true;
					break;
				}
			}
			if (!found) {
				GfaArc c;
				c.v = arc[i].w ^ 1, c.w = v ^ 1, c.ov = arc[i].ow, c.ow = arc[i].ov;
				c.link = arc[i].link, c.comp = true;
				arc.push_back(c);
			}
		}
	// gfa_fix_arc_len: clamp overlaps to the segment, drop arcs of dropped
	// segments, lv = len(v) - ov
	for (GfaArc &a : arc) {
		const GfaSeg &sv = seg[a.v >> 1];
		if (!sv.del && sv.len < a.ov) a.ov = (int32_t)sv.len;
		if (sv.del || seg[a.w >> 1].del) a.del = true;
		else a.lv = (uint32_t)(sv.len - a.ov);
	}
	// gfa_cleanup: remove deleted arcs, sort, index
	arc.erase(std::remove_if(arc.begin(), arc.end(), [&](const GfaArc &a) {
		return a.del || seg[a.v >> 1].del || seg[a.w >> 1].del;
	}), arc.end());
	sort_index();
}

// gfa_sub_from (gfa-sub.c:41-272) with an unbounded radius: the vertices
// reachable from v0 in the order the reference emits them, and per vertex
// the arcs into that set as (new target index << 32 | arc index), sorted.
struct SubNode {
	uint64_t xnd;     // bit 63: not yet forced; bits 32-62: unresolved in-arcs; bits 0-31: distance
	uint32_t v;
	uint32_t pos = 0; // position in the output order
	bool in_tree = true, forced = false;
};

void extract_sub(const GfaGraph &g, uint32_t v0, std::vector<uint32_t> &order, std::vector<uint64_t> &sub_arc,
                 std::vector<uint32_t> &sub_off) {
	auto tn_n = [](const SubNode *p) { return (uint32_t)(p->xnd << 1 >> 33); };
	std::unordered_map<uint32_t, SubNode *> h;
	std::vector<std::unique_ptr<SubNode>> pool;
	std::set<std::pair<uint64_t, uint32_t>> tree;  // (xnd, v): the order of tn_cmp
	auto gen = [&](uint32_t v, uint32_t d) {
		pool.emplace_back(new SubNode());
		SubNode *p = pool.back().get();
		p->v = v;
		p->xnd = 1ull << 63 | (uint64_t)g.n(v ^ 1) << 32 | d;
		h[v] = p;
		return p;
	};
	auto find = [&](uint32_t v) -> SubNode * {
		auto it = h.find(v);
		return it == h.end() ? nullptr : it->second;
	};
	{
		SubNode *p = gen(v0, 0);
		tree.insert({p->xnd, p->v});
	}
	std::vector<SubNode *> L;
	int32_t orphan_inv = 0;
	while (!tree.empty()) {
		SubNode *q = nullptr;
		SubNode *r = find(tree.begin()->second);
		if (orphan_inv) {  // first the vertices whose complement has already left the set
			for (auto it = tree.begin(); it != tree.end(); ++it) {
				SubNode *x = find(it->second);
				SubNode *c = find(x->v ^ 1);
				if (c && !c->in_tree) {
					--orphan_inv;
					q = x;
					tree.erase(it);
					break;
				}
			}
		} else if (tn_n(r) > 0) {
			// the smallest vertex still has unresolved in-arcs: unless every
			// in-arc comes from a vertex whose complement alone was emitted,
			// clear every vertex's "not forced" bit
			const uint32_t u = r->v ^ 1;
			uint32_t i = 0;
			for (; i < g.n(u); ++i) {
				const GfaArc &a = g.arc[g.first(u) + i];
				SubNode *k1 = find(a.w ^ 1), *k2 = find(a.w);
				if (!((k1 == nullptr && k2 && !k2->in_tree) || (k2 == nullptr && k1 && !k1->in_tree))) break;
			}
			if (i < g.n(u)) {
				std::vector<SubNode *> all;
				for (const auto &e : tree) all.push_back(find(e.second));
				tree.clear();
				for (SubNode *x : all) {
					x->xnd &= ~(1ull << 63);
					tree.insert({x->xnd, x->v});
				}
			}
		}
		if (!q) {
			q = find(tree.begin()->second);
			tree.erase(tree.begin());
		}
		q->forced = tn_n(q) > 0 || (q->xnd >> 63) == 0;
		q->in_tree = false;
		L.push_back(q);
		if (SubNode *c = find(q->v ^ 1)) orphan_inv += c->in_tree;
		const uint32_t d = (uint32_t)q->xnd;
		for (uint32_t i = g.first(q->v), e = g.first(q->v) + g.n(q->v); i < e; ++i) {
			const GfaArc &a = g.arc[i];
			const uint32_t dt = d + (uint32_t)g.seg[a.w >> 1].len;
			if (SubNode *c = find(a.w ^ 1))
				if (!c->in_tree && !c->forced) continue;  // the other strand already left: a bidirected step
			SubNode *p = find(a.w);
			if (!p) {
				p = gen(a.w, dt);
			} else {
				if (!p->in_tree) continue;  // a cycle back to an emitted vertex
				tree.erase({p->xnd, p->v});
				if (dt < (uint32_t)p->xnd) p->xnd = p->xnd >> 32 << 32 | dt;
			}
			if (tn_n(p) > 0) p->xnd -= 1ull << 32;
			tree.insert({p->xnd, p->v});
		}
	}
	order.clear();
	for (size_t j = 0; j < L.size(); ++j) L[j]->pos = (uint32_t)j, order.push_back(L[j]->v);
	sub_arc.clear();
	sub_off.assign(L.size() + 1, 0);
	for (size_t j = 0; j < L.size(); ++j) {
		const size_t o0 = sub_arc.size();
		for (uint32_t i = g.first(L[j]->v), e = g.first(L[j]->v) + g.n(L[j]->v); i < e; ++i)
			if (SubNode *p = find(g.arc[i].w)) sub_arc.push_back((uint64_t)p->pos << 32 | i);
		std::sort(sub_arc.begin() + o0, sub_arc.end());
		sub_off[j + 1] = (uint32_t)sub_arc.size();
	}
}

}  // namespace

int gwf_gfa_load(const char *path, const char *start, gwf_graph_t **out) {
	if (!path || !out) return GWF_ERR_ARG;
	*out = nullptr;
	LineReader lr(path);
	if (!lr.ok()) return GWF_ERR_ARG;
	GfaGraph g;
	std::string line;
	std::vector<std::string> f;
	int64_t fa_cur = -1;  // segment of the FASTA record being read
	std::string fa_seq;
	auto fa_flush = [&]() {
		if (fa_cur >= 0) {
			g.seg[fa_cur].seq = fa_seq;
			g.seg[fa_cur].has_seq = true;
			g.seg[fa_cur].len = (int64_t)fa_seq.size();
		}
		fa_cur = -1;
		fa_seq.clear();
	};
	while (lr.next(line)) {
		if (!line.empty() && line[0] == '>') {  // a FASTA record is a segment named s<index + 1>
			fa_flush();                           // (gfa_parse_fa_hdr, gfa-io.c:326-340)
			fa_cur = g.add_seg("s" + std::to_string(g.seg.size() + 1));
			continue;
		}
		if (fa_cur >= 0) {
			if (!(line.size() >= 3 && line[1] == '\t')) {
				fa_seq += line;
				continue;
			}
			fa_flush();
		}
		if (line.size() < 3 || line[1] != '\t') continue;
		split_tabs(line, f);
		if (line[0] == 'S') {  // gfa_parse_S
			if (f.size() < 3) continue;
			const uint32_t sid = g.add_seg(f[1]);
			GfaSeg &s = g.seg[sid];
			int64_t LN = -1;
			for (size_t i = 3; i < f.size(); ++i)
				if (f[i].compare(0, 5, "LN:i:") == 0) LN = atoll(f[i].c_str() + 5);
			s.has_seq = f[2] != "*";
			s.seq = s.has_seq ? f[2] : std::string();
			s.len = s.has_seq ? (int64_t)s.seq.size() : (LN >= 0 ? LN : 0);
		} else if (line[0] == 'L') {  // gfa_parse_L
			if (f.size() < 5) continue;
			if ((f[2] != "+" && f[2] != "-") || (f[4] != "+" && f[4] != "-")) continue;
			GfaArc a;
			parse_overlap(f.size() > 5 ? f[5].c_str() : "", a.ov, a.ow);
			a.v = g.add_seg(f[1]) << 1 | (f[2] == "-");
			a.w = g.add_seg(f[3]) << 1 | (f[4] == "-");
			a.link = (uint32_t)g.arc.size();
			for (size_t i = 6; i < f.size(); ++i) {  // L1/L2: lengths implied by the link
				if (f[i].compare(0, 5, "L1:i:") == 0 && a.ov != kNoOvl)
					g.seg[a.v >> 1].len = std::max<int64_t>(g.seg[a.v >> 1].len, a.ov + atoll(f[i].c_str() + 5));
				if (f[i].compare(0, 5, "L2:i:") == 0 && a.ow != kNoOvl)
					g.seg[a.w >> 1].len = std::max<int64_t>(g.seg[a.w >> 1].len, a.ow + atoll(f[i].c_str() + 5));
			}
			g.arc.push_back(a);
		}
	}
	fa_flush();
	if (g.seg.empty()) return GWF_ERR_ARG;
	g.finalize();
	uint32_t v0 = 0;  // the first segment, forward strand (main.cu:368)
	if (start) {
		auto it = g.id.find(start);
		if (it == g.id.end()) return GWF_ERR_ARG;
		v0 = it->second << 1;
	}
	std::vector<uint32_t> order, sub_off;
	std::vector<uint64_t> sub_arc;
	extract_sub(g, v0, order, sub_arc, sub_off);
	Builder b;
	const unsigned char *ct = comp_table();
	for (uint32_t v : order) {  // gwf_gfa2gwf
		const GfaSeg &sg = g.seg[v >> 1];
		// a segment given as "*" has no bases; N stands for them (the
		// reference would dereference a null sequence here)
		std::string s = sg.has_seq ? sg.seq : std::string((size_t)sg.len, 'N');
		s.resize((size_t)sg.len, 'N');
		if (v & 1) {
			std::string rc(s.size(), 'N');
			for (size_t j = 0; j < s.size(); ++j) rc[j] = (char)ct[(unsigned char)s[s.size() - 1 - j]];
			b.add(rc, v);
		} else b.add(s, v);
	}
	for (size_t i = 0; i < order.size(); ++i)
		for (uint32_t k = sub_off[i]; k < sub_off[i + 1]; ++k)
			b.link((uint32_t)i, (uint32_t)(sub_arc[k] >> 32), g.arc[(uint32_t)sub_arc[k]].ow);
	*out = b.finish();
	return *out ? GWF_OK : GWF_ERR_NOMEM;
}

int gwf_synth_graph(const gwf_synth_graph_opts_t *o, gwf_graph_t **out) {
	if (!o || !out || o->backbone_len <= 0 || o->seg_min <= 0 || o->seg_max < o->seg_min) return GWF_ERR_ARG;
	if (o->p_tandem > 0 && (o->tr_motif_min < 1 || o->tr_motif_max < o->tr_motif_min ||
	                        o->tr_copy_min < 1 || o->tr_copy_max < o->tr_copy_min))
		return GWF_ERR_ARG;
	Synth s(*o);
	*out = s.run();
	return *out ? GWF_OK : GWF_ERR_NOMEM;
}

int gwf_synth_reads(const gwf_graph_t *g, const gwf_synth_reads_opts_t *o, int32_t *n_out, char **seqs_out,
                    int64_t **off_out, int32_t **ql_out, int32_t **v0_out, float **err_out) {
	if (!g || !o || !n_out || !seqs_out || !off_out || !ql_out || o->n_reads < 0 || g->n_vtx == 0)
		return GWF_ERR_ARG;
	if (o->err_mode == GWF_SYNTH_ERR_CYCLE && (o->n_err_levels < 1 || o->n_err_levels > 8)) return GWF_ERR_ARG;
	Rng r(o->seed);
	Adj adj(g);
	const int32_t n = o->n_reads;
	std::vector<int32_t> lens(n), v0s(n);
	std::vector<float> errs(n);
	std::string all, rd;
	std::vector<int64_t> off(n);
	for (int32_t i = 0; i < n; ++i) {
		int32_t L = o->len_a;
		if (o->len_mode == GWF_SYNTH_LEN_LOGNORMAL) {
			double x = std::exp(std::log((double)o->len_a) + o->len_sigma * r.normal());
			L = (int32_t)std::llround(std::min<double>(std::max<double>(x, o->len_min), o->len_max));
		}
		double e = o->err_mode == GWF_SYNTH_ERR_CYCLE ? o->err_levels[i % o->n_err_levels]
		                                              : o->err_lo + (o->err_hi - o->err_lo) * r.uniform();
		bool ok = false;
		uint32_t v0 = 0;
		for (int t = 0; t < 10000 && !ok; ++t) {
			v0 = o->start_mode == GWF_SYNTH_START_RANDOM ? (uint32_t)r.below(g->n_vtx) : 0;
			ok = sample_read(g, adj, r, v0, L, e, rd);
			if (!ok && o->start_mode != GWF_SYNTH_START_RANDOM) break;
		}
		if (!ok) return GWF_ERR_ARG;  // graph too short for the requested read length
		off[i] = (int64_t)all.size();
		all += rd;
		all += '\0';
		lens[i] = L, v0s[i] = (int32_t)v0, errs[i] = (float)e;
	}
	char *seqs = (char *)malloc(all.size() + 1);
	int64_t *po = (int64_t *)malloc(sizeof(int64_t) * (n + 1));
	int32_t *pl = (int32_t *)malloc(sizeof(int32_t) * (n + 1));
	int32_t *pv = (int32_t *)malloc(sizeof(int32_t) * (n + 1));
	float *pe = (float *)malloc(sizeof(float) * (n + 1));
	if (!seqs || !po || !pl || !pv || !pe) { free(seqs); free(po); free(pl); free(pv); free(pe); return GWF_ERR_NOMEM; }
	memcpy(seqs, all.data(), all.size());
	seqs[all.size()] = 0;
	for (int32_t i = 0; i < n; ++i) po[i] = off[i], pl[i] = lens[i], pv[i] = v0s[i], pe[i] = errs[i];
	*n_out = n, *seqs_out = seqs, *off_out = po, *ql_out = pl;
	if (v0_out) *v0_out = pv; else free(pv);
	if (err_out) *err_out = pe; else free(pe);
	return GWF_OK;
}

}  // extern "C"
