// gwf_host.cpp — libgwfa_host.so: GFA/FASTA readers, graph printers and the
// seeded synthetic graph/read generator.  Host C++ only (no HIP).
//
// GFA semantics follow the reference's GFA layer: S/L line parsing
// (gfa-io.c:115-336), complement arcs (gfa_fix_symm_add, gfa-base.c:303),
// removal of duplicate links, and the conversion to a gwf graph rooted at the
// start segment with reverse-complemented `-` strands and arc overlap = the
// target-side overlap `ow` (main.cu:294-345).  The vertex order is a
// topological order rooted at the start vertex (Kahn's algorithm over the
// reachable subgraph); edit distances do not depend on vertex numbering.
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
struct GfaSeg { std::string name, seq; int64_t ln = -1; bool has_s = false; };
struct GfaArc { uint32_t v, w; int32_t ov, ow; };

bool parse_overlap(const char *q, int32_t &ov, int32_t &ow) {
	if (*q == 0) { ov = ow = 0; return true; }
	if (*q == '*') { ov = ow = 0; return true; }
	if (isdigit((unsigned char)*q)) {
		char *e;
		long x = strtol(q, &e, 10);
		if (*e == ':') { ov = (int32_t)x; ow = isdigit((unsigned char)e[1]) ? (int32_t)strtol(e + 1, 0, 10) : 0; return true; }
		if (isupper((unsigned char)*e)) {  // CIGAR
			ov = ow = 0;
			const char *p = q;
			while (isdigit((unsigned char)*p)) {
				long l = strtol(p, (char **)&p, 10);
				if (*p == 'M' || *p == 'D' || *p == 'N') ov += (int32_t)l;
				if (*p == 'M' || *p == 'I' || *p == 'S') ow += (int32_t)l;
				++p;
			}
			return true;
		}
		if (*e == 0) { ov = ow = (int32_t)x; return true; }
	}
	return false;
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

int gwf_gfa_load(const char *path, const char *start, gwf_graph_t **out) {
	if (!path || !out) return GWF_ERR_ARG;
	*out = nullptr;
	LineReader lr(path);
	if (!lr.ok()) return GWF_ERR_ARG;
	std::vector<GfaSeg> segs;
	std::unordered_map<std::string, uint32_t> id;
	std::vector<GfaArc> arcs;
	auto seg_id = [&](const std::string &nm) {
		auto it = id.find(nm);
		if (it != id.end()) return it->second;
		uint32_t i = (uint32_t)segs.size();
		segs.push_back(GfaSeg());
		segs.back().name = nm;
		id.emplace(nm, i);
		return i;
	};
	std::string line;
	std::vector<std::string> f;
	int64_t fa_cur = -1;
	while (lr.next(line)) {
		if (!line.empty() && line[0] == '>') {  // FASTA record = one segment
			size_t e = line.find_first_of(" \t", 1);
			fa_cur = seg_id(line.substr(1, e == std::string::npos ? std::string::npos : e - 1));
			segs[fa_cur].has_s = true;
			segs[fa_cur].seq.clear();
			continue;
		}
		if (fa_cur >= 0 && !(line.size() >= 3 && line[1] == '\t')) {
			segs[fa_cur].seq += line;
			continue;
		}
		fa_cur = -1;
		if (line.size() < 3 || line[1] != '\t') continue;
		if (line[0] == 'S') {
			split_tabs(line, f);
			if (f.size() < 3) return GWF_ERR_ARG;
			uint32_t s = seg_id(f[1]);
			segs[s].has_s = true;
			segs[s].seq = f[2] == "*" ? std::string() : f[2];
			for (size_t i = 3; i < f.size(); ++i)
				if (f[i].compare(0, 5, "LN:i:") == 0) segs[s].ln = atoll(f[i].c_str() + 5);
			if (f[2] == "*" && segs[s].ln > 0) segs[s].seq.assign((size_t)segs[s].ln, 'N');
		} else if (line[0] == 'L') {
			split_tabs(line, f);
			if (f.size() < 5) return GWF_ERR_ARG;
			if ((f[2] != "+" && f[2] != "-") || (f[4] != "+" && f[4] != "-")) return GWF_ERR_ARG;
			GfaArc a;
			a.v = seg_id(f[1]) << 1 | (f[2] == "-");
			a.w = seg_id(f[3]) << 1 | (f[4] == "-");
			if (!parse_overlap(f.size() > 5 ? f[5].c_str() : "", a.ov, a.ow)) return GWF_ERR_ARG;
			arcs.push_back(a);
		}
	}
	if (segs.empty()) return GWF_ERR_ARG;
	// complement arcs (gfa_fix_symm_add) and duplicate removal
	const size_t na0 = arcs.size();
	for (size_t i = 0; i < na0; ++i) {
		GfaArc c;
		c.v = arcs[i].w ^ 1, c.w = arcs[i].v ^ 1, c.ov = arcs[i].ow, c.ow = arcs[i].ov;
		arcs.push_back(c);
	}
	std::stable_sort(arcs.begin(), arcs.end(), [](const GfaArc &x, const GfaArc &y) {
		return x.v != y.v ? x.v < y.v : x.w < y.w;
	});
	arcs.erase(std::unique(arcs.begin(), arcs.end(), [](const GfaArc &x, const GfaArc &y) {
		return x.v == y.v && x.w == y.w;
	}), arcs.end());
	// drop links touching segments that never had an S line
	arcs.erase(std::remove_if(arcs.begin(), arcs.end(), [&](const GfaArc &x) {
		return !segs[x.v >> 1].has_s || !segs[x.w >> 1].has_s;
	}), arcs.end());
	const uint32_t nov = (uint32_t)segs.size() * 2;
	std::vector<uint32_t> aoff(nov + 1, 0);
	for (auto &a : arcs) ++aoff[a.v + 1];
	for (uint32_t v = 0; v < nov; ++v) aoff[v + 1] += aoff[v];  // arcs are sorted by v
	uint32_t v0;
	if (start) {
		auto it = id.find(start);
		if (it == id.end() || !segs[it->second].has_s) return GWF_ERR_ARG;
		v0 = it->second << 1;
	} else {
		uint32_t s = 0;
		while (s < segs.size() && !segs[s].has_s) ++s;
		if (s == segs.size()) return GWF_ERR_ARG;
		v0 = s << 1;
	}
	// reachable oriented vertices from v0, in discovery order
	std::vector<int64_t> idx(nov, -1);
	std::vector<uint32_t> reach;
	std::deque<uint32_t> bfs{v0};
	idx[v0] = 0, reach.push_back(v0);
	while (!bfs.empty()) {
		uint32_t v = bfs.front();
		bfs.pop_front();
		for (uint32_t j = aoff[v]; j < aoff[v + 1]; ++j)
			if (idx[arcs[j].w] < 0) {
				idx[arcs[j].w] = (int64_t)reach.size();
				reach.push_back(arcs[j].w);
				bfs.push_back(arcs[j].w);
			}
	}
	// Kahn order over the reachable subgraph, v0 forced first; on a cycle, the
	// earliest-discovered remaining vertex is taken next.
	const size_t nr = reach.size();
	std::vector<uint32_t> indeg(nr, 0), order;
	std::vector<char> done(nr, 0);
	for (uint32_t v : reach)
		for (uint32_t j = aoff[v]; j < aoff[v + 1]; ++j) ++indeg[idx[arcs[j].w]];
	std::deque<uint32_t> ready;
	size_t scan = 0;
	auto emit = [&](uint32_t ri) {
		done[ri] = 1;
		order.push_back(ri);
		uint32_t v = reach[ri];
		for (uint32_t j = aoff[v]; j < aoff[v + 1]; ++j) {
			uint32_t wi = (uint32_t)idx[arcs[j].w];
			if (!done[wi] && --indeg[wi] == 0) ready.push_back(wi);
		}
	};
	emit(0);
	while (order.size() < nr) {
		if (ready.empty()) {
			while (done[scan]) ++scan;
			emit((uint32_t)scan);
			continue;
		}
		uint32_t ri = ready.front();
		ready.pop_front();
		if (!done[ri]) emit(ri);
	}
	std::vector<uint32_t> newid(nr);
	for (size_t i = 0; i < nr; ++i) newid[order[i]] = (uint32_t)i;
	Builder b;
	const unsigned char *ct = comp_table();
	for (size_t i = 0; i < nr; ++i) {
		uint32_t v = reach[order[i]];
		const std::string &s = segs[v >> 1].seq;
		if (v & 1) {
			std::string rc(s.size(), 'N');
			for (size_t j = 0; j < s.size(); ++j) rc[j] = (char)ct[(unsigned char)s[s.size() - 1 - j]];
			b.add(rc, v);
		} else b.add(s, v);
	}
	for (size_t i = 0; i < nr; ++i) {
		uint32_t v = reach[order[i]];
		for (uint32_t j = aoff[v]; j < aoff[v + 1]; ++j)
			b.link((uint32_t)i, newid[idx[arcs[j].w]], arcs[j].ow);
	}
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
