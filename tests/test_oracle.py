"""Pin the oracle (the CPU restatement of gwf_ed, gwf-ed.cu:80-1337) before
trusting it: against the independent brute-force DP on random graphs,
against closed forms, invariants and the reference's fixture inputs."""
import os
import random

import numpy as np
import pytest

from conftest import GOLDEN, read_ints


def rand_graph(gw, rng, nv, dag=True, ovl=False, maxlen=12):
    seqs = []
    for _ in range(nv):
        L = rng.randint(1, 3 if rng.random() < 0.5 else maxlen)
        alph = "ACGT"[: rng.randint(2, 4)]
        seqs.append("".join(rng.choice(alph) for _ in range(L)))
    fixed_o = [((w * 7 + 3) % len(seqs[w])) if ovl else 0 for w in range(nv)]
    arcs = []
    for v in range(nv):
        for _ in range(rng.randint(0, 2)):
            if dag:
                if v + 1 >= nv:
                    break
                w = rng.randint(v + 1, nv - 1)
            else:
                w = rng.randrange(nv)
            arcs.append((v, w, fixed_o[w]))
    g = gw.Graph.from_lists(seqs, arcs)
    g.index()
    return g, seqs


def rand_read(rng, n):
    return "".join(rng.choice("ACGT") for _ in range(n))


@pytest.mark.parametrize("dag,ovl", [(True, False), (False, False), (True, True), (False, True)])
def test_oracle_equals_dp_random(gw, orc, dag, ovl):
    """Exact ED for v1 < 0 when all in-arcs of a vertex share one overlap."""
    rng = random.Random(hash((dag, ovl)) & 0xffff)
    for it in range(600):
        g, seqs = rand_graph(gw, rng, rng.randint(1, 12), dag, ovl)
        q = rand_read(rng, rng.randint(0, 30))
        assert orc.ed(g, q) == orc.dp(g, q), (it, seqs, g.arcs(), q)


def test_oracle_equals_dp_larger(gw, orc):
    rng = random.Random(7)
    for it in range(40):
        g, seqs = rand_graph(gw, rng, rng.randint(20, 60), True, False, maxlen=40)
        q = rand_read(rng, rng.randint(50, 300))
        assert orc.ed(g, q) == orc.dp(g, q), it


def test_dp_catches_planted_mistakes(gw, orc):
    """The DP pins plausible mistakes: an oracle that charged 2 per
    substitution, or ignored the free end, would disagree on these."""
    g = gw.Graph.from_lists(["ACGTACGTACGT"], [])
    g.index()
    assert orc.dp(g, "ACGAACGT") == 1 == orc.ed(g, "ACGAACGT")      # one substitution
    assert orc.dp(g, "ACGT") == 0 == orc.ed(g, "ACGT")              # free end in the graph
    assert orc.dp(g, "ACGTTACGT") == 1 == orc.ed(g, "ACGTTACGT")    # one insertion
    assert orc.dp(g, "ACGACGT") == 1 == orc.ed(g, "ACGACGT")        # one deletion
    assert orc.dp(g, "TACGT") == 1 == orc.ed(g, "TACGT")            # leading insertion


def test_v1_target_vertex_is_upper_bound(gw, orc):
    """For v1 >= 0 the reference algorithm skips the insertion of a vertex-end
    state when every out-arc matches (gwf-ed.cu:1225); its result is then an
    upper bound of the exact end-at-v1 distance, never below it."""
    rng = random.Random(11)
    for _ in range(300):
        g, _ = rand_graph(gw, rng, rng.randint(1, 10))
        q = rand_read(rng, rng.randint(0, 20))
        v1 = rng.randrange(g.n_vtx)
        a, b = orc.ed(g, q, v1=v1), orc.dp(g, q, v1=v1)
        assert (a == -1) == (b == -1)
        assert a >= b


def sellers(t, q):
    """Semi-global edit distance: q fully aligned to a prefix of t."""
    prev = np.arange(len(q) + 1)
    best = prev[-1]
    for c in t:
        cur = np.empty_like(prev)
        cur[0] = prev[0] + 1
        for j in range(1, len(q) + 1):
            cur[j] = min(prev[j] + 1, cur[j - 1] + 1, prev[j - 1] + (q[j - 1] != c))
        prev = cur
        best = min(best, prev[-1])
    return int(best)


def test_closed_forms(gw, orc):
    rng = random.Random(3)
    t = rand_read(rng, 200)
    g = gw.Graph.from_lists([t], [])
    g.index()
    assert orc.ed(g, t[:137]) == 0                     # exact prefix
    assert orc.ed(g, "") == 0                          # empty read
    for _ in range(30):                                # linear graph = Sellers DP
        q = rand_read(rng, rng.randint(1, 60))
        assert orc.ed(g, q) == sellers(t, q)
    q = list(t[:150])                                  # k separated substitutions
    for p in (10, 40, 70, 100, 130):
        q[p] = "A" if q[p] != "A" else "C"
    assert orc.ed(g, "".join(q)) == 5
    for _ in range(20):                                # ED <= ql
        q = rand_read(rng, rng.randint(0, 50))
        assert orc.ed(g, q) <= len(q)


def test_all_insertion_end(gw, orc):
    """ED == ql through an end with offset -1 (the reference loop would spin)."""
    g = gw.Graph.from_lists(["A", "AAAA"], [(0, 1)])
    g.index()
    assert orc.ed(g, "TT") == 2
    assert orc.ed(g, "TTTTTTTT") == 8


def test_short_start_vertex(gw, orc):
    """A start segment shorter than 32 bp with an exact-prefix read gives 0
    (the reference's unsigned max_k would skip that extension)."""
    g = gw.Graph.from_lists(["ACGTACGTAC", "GGTTAACC"], [(0, 1)])
    g.index()
    assert orc.ed(g, "ACGTACGTACGGTTAA") == 0


def test_invariants(gw, orc):
    rng = random.Random(5)
    for _ in range(60):
        g, seqs = rand_graph(gw, rng, rng.randint(2, 10))
        arcs = g.arcs()
        q = rand_read(rng, rng.randint(0, 25))
        ed = orc.ed(g, q)
        # relabel vertices (start stays 0)
        perm = [0] + rng.sample(range(1, len(seqs)), len(seqs) - 1)
        inv = {old: new for new, old in enumerate(perm)}
        g2 = gw.Graph.from_lists([seqs[o] for o in perm], [(inv[a], inv[b], o) for a, b, o in arcs])
        g2.index()
        assert orc.ed(g2, q) == ed
        # start index given through v0 instead of being vertex 0
        rot = list(range(1, len(seqs))) + [0]
        inv2 = {old: new for new, old in enumerate(rot)}
        g3 = gw.Graph.from_lists([seqs[o] for o in rot], [(inv2[a], inv2[b], o) for a, b, o in arcs])
        g3.index()
        assert orc.ed(g3, q, v0=inv2[0]) == ed
        # unreachable vertices change nothing
        g4 = gw.Graph.from_lists(seqs + ["ACGTTT", "GG"], arcs + [(len(seqs), len(seqs) + 1, 0)])
        g4.index()
        assert orc.ed(g4, q) == ed
        # split a vertex into a chain of two
        v = rng.randrange(len(seqs))
        if len(seqs[v]) >= 2:
            cut = rng.randint(1, len(seqs[v]) - 1)
            s5 = list(seqs) + [seqs[v][cut:]]
            s5[v] = seqs[v][:cut]
            nw = len(seqs)
            a5 = [(nw if a == v else a, b, o) for a, b, o in arcs] + [(v, nw, 0)]
            g5 = gw.Graph.from_lists(s5, a5)
            g5.index()
            assert orc.ed(g5, q) == ed


def test_golden_fixture_1000_reads(gw, orc):
    g = gw.Graph.from_gfa(os.path.join(GOLDEN, "graph_1000.gfa")).index()
    r = gw.Reads.from_fasta(os.path.join(GOLDEN, "reads_1000.fa.gz"))
    want = read_ints(os.path.join(GOLDEN, "reads_1000.ed"))
    got, _, _ = orc.batch(g, r)
    assert got.tolist() == want
    for i in range(0, 1000, 97):            # re-check a sample against the DP
        assert orc.dp(g, r.read(i)) == want[i]


def test_golden_fixture_grafo(gw, orc):
    g = gw.Graph.from_gfa(os.path.join(GOLDEN, "grafo_test.gfa")).index()
    r = gw.Reads.from_fasta(os.path.join(GOLDEN, "grafo_reads.fa"))
    want = read_ints(os.path.join(GOLDEN, "grafo_reads.ed"))
    got, _, _ = orc.batch(g, r)
    assert got.tolist() == want


def test_index_matches_reference_restatement(gw, orc):
    """gwf_ed_index (product) vs the oracle's restatement of main.cu:246-270."""
    import ctypes as C
    rng = random.Random(9)
    for _ in range(50):
        g, seqs = rand_graph(gw, rng, rng.randint(1, 30), dag=False)
        arcs = g.arcs()
        h = gw.Graph.from_lists(seqs, arcs)
        aux = np.zeros(max(h.n_vtx, 1), dtype=np.uint64)
        orc.lib().oracle_ed_index(C.cast(h.ptr, C.c_void_p), aux.ctypes.data)
        assert g.aux().tolist() == aux[:h.n_vtx].tolist()
        assert [a[:2] for a in g.arcs()] == [a[:2] for a in h.arcs()]


def test_batch_runner_threads(gw, orc):
    g, r = gw.make_workload(1, 200)
    g.index()
    a, st1, _ = orc.batch(g, r, threads=1)
    b, st8, t = orc.batch(g, r, threads=4)
    assert a.tolist() == b.tolist()
    assert st1["states"] == st8["states"] and st1["steps"] == int(a.sum()) + len(a)


def test_v1_skip_insertion_case(gw, orc):
    """A case where the reference algorithm's end-at-v1 result is not the
    exact distance: the read needs an insertion after the last base of v1
    (vertex 13), but every arc out of v1 matches the next query base, so the
    reference adds no insertion there (gwf-ed.cu:1225-1226) and ends later.
    Recording the exits of vertices entered in the same step
    (gwf-ed.cu:1177-1183) also matters here: without them the result is 5."""
    seqs = ['G', 'ACCCACCCACCAAACAAACCCCCA', 'A', 'A', 'CCA', 'A', 'AGGCGACCAAGA', 'AAA', 'CC',
            'ACCACCCCCCGCGAGG', 'GGGAGACAC', 'ACAAGGGCCGCAAGCGCGC', 'AA', 'C', 'CCA']
    arcs = [(0, 2, 0), (0, 4, 1), (1, 4, 1), (1, 12, 1), (2, 9, 2), (3, 6, 9), (3, 13, 0), (5, 7, 1), (7, 12, 1),
            (9, 13, 0), (10, 12, 1), (10, 13, 0), (13, 14, 2)]
    g = gw.Graph.from_lists(seqs, arcs)
    g.index()
    q = "GACACCCCCCGCGAGGCA"
    assert orc.dp(g, q, v1=13) == 1
    assert orc.ed(g, q, v1=13) == 8
    assert orc.ed(g, q[:-1], v1=13) == 0      # without the trailing base the end is exact


def test_mixed_in_overlaps_make_the_result_order_dependent(gw, orc):
    """Why the GPU follows the reference's closure order on a graph where one
    vertex has in-arcs with different overlaps: the visited set keys a vertex
    entry by (target, query position) (gwf-ed.cu:1196), so two arcs entering
    one vertex at one query position with different overlaps give two
    different states under one key, and the reference keeps whichever its FIFO
    closure reaches first.  Relabelling the vertices (start kept at 0) then
    changes its answer; with one overlap per target it never does
    (test_invariants)."""
    seqs = ["CAA", "ACC", "GTGG"]
    arcs = [(0, 1, 0), (1, 1, 2), (1, 2, 3), (1, 1, 1), (2, 0, 2), (2, 2, 1), (2, 1, 2)]
    q = "GCGCTCGGAAGTATTATT"
    g = gw.Graph.from_lists(seqs, arcs)
    g.index()
    perm = [0, 2, 1]
    inv = {old: new for new, old in enumerate(perm)}
    g2 = gw.Graph.from_lists([seqs[o] for o in perm], [(inv[a], inv[b], o) for a, b, o in arcs])
    g2.index()
    assert (orc.ed(g, q), orc.ed(g2, q), orc.dp(g, q)) == (12, 11, 11)


# ---- mixed in-overlaps.  The visited set refuses a second claim on a key
# (w, i+1) (gwf-ed.cu:1196-1212).  A refusal to an arc of the same overlap
# drops a copy of a state already reached, which changes nothing; only a
# refusal to an arc of another overlap (counted by the oracle as vis_lossy)
# can drop a state the exact distance needs.  So: no such refusal -> the
# exact DP; otherwise never below it.  Which claim wins is fixed by the
# reference's order: arcs of one vertex in CSR order (the stable sort of
# gwf_ed_index keeps input order between parallel arcs), the A-queue first in,
# first out, and the key claimed before the first-base test.

def _mixed_case(gw, rng, dag):
    nv = rng.randint(1, 12)
    seqs = ["".join(rng.choice("ACGT"[:rng.randint(2, 4)]) for _ in range(rng.randint(1, 3 if rng.random() < 0.4 else 10)))
            for _ in range(nv)]
    arcs = set()
    for v in range(nv):
        for _ in range(rng.randint(0, 3)):
            if dag:
                if v + 1 >= nv:
                    break
                w = rng.randint(v + 1, nv - 1)
            else:
                w = rng.randrange(nv)
            arcs.add((v, w, rng.randrange(len(seqs[w]))))
    g = gw.Graph.from_lists(seqs, sorted(arcs))
    g.index()
    return g, rand_read(rng, rng.randint(0, 25))


@pytest.mark.parametrize("dag", [True, False])
def test_mixed_in_overlaps_exact_without_lossy_claims(gw, orc, dag):
    rng = random.Random(41 + dag)
    lossy = above = 0
    for it in range(2000):
        g, q = _mixed_case(gw, rng, dag)
        st = {}
        e, d = orc.ed(g, q, stats=st), orc.dp(g, q)
        if st["vis_lossy"]:
            lossy += 1
            above += e > d
            assert e >= d, (it, g.arcs(), q)
        else:
            assert e == d, (it, g.arcs(), q)
    assert lossy > 300 and above > 20   # both kinds of run occur


def _claim_order_cases(gw):
    """(graph, query, the reference's answer) derived by hand from
    gwf-ed.cu:1186-1212.  1: one vertex "AC" with two arcs into "GT", of
    overlaps 0 and 1; q = "ACT".  At the end of "AC" (query position 1) the
    first arc in CSR order claims (w, 2): with overlap 0 its first base G
    mismatches T, so only a mismatch/deletion state enters w and the claim of
    the overlap-1 arc (which would match T and end at 0) is refused: answer 1.
    With the arcs the other way round the overlap-1 arc claims first: 0.
    2: "A" -> "C" (vertex 1), "A" -> "C" (vertex 2), both -> "GT" with overlaps
    o1 and o2; q = "ACT".  The two "C" states enter the A-queue in CSR order
    (1 before 2) and, first in first out, vertex 1 claims (3, 2) first: 1 when
    o1 = 0, 0 when o1 = 1.  The exact distance is 0 in all four."""
    out = []
    for arcs, want in (([(0, 1, 0), (0, 1, 1)], 1), ([(0, 1, 1), (0, 1, 0)], 0)):
        out.append((gw.Graph.from_lists(["AC", "GT"], arcs), "ACT", want))
    for o1, o2, want in ((0, 1, 1), (1, 0, 0)):
        out.append((gw.Graph.from_lists(["A", "C", "C", "GT"], [(0, 1, 0), (0, 2, 0), (1, 3, o1), (2, 3, o2)]),
                    "ACT", want))
    for g, _, _ in out:
        g.index()
    return out


def test_mixed_claim_order_hand_cases(gw, orc):
    for g, q, want in _claim_order_cases(gw):
        assert orc.ed(g, q) == want
        assert orc.dp(g, q) == 0


# Plausible mistakes in the oracle's closure crossing (gwf_oracle.c, the
# restatement of gwf-ed.cu:1186-1212), each planted in a copy compiled here;
# the checks above must reject every one.
MIXED_PLANTS = [
    ("visited key without the query position",
     "or_set_put(b, (uint64_t)w << 32 | (uint32_t)(ii + 1), ol)", "or_set_put(b, (uint64_t)w << 32, ol)"),
    ("entry at the vertex start instead of at the overlap",
     "or_A_push(b, or_gen_vd(w, ii + 1 - ol), ol,", "or_A_push(b, or_gen_vd(w, ii + 1), 0,"),
    ("a key claimed only by a matching arc",
     "int absent = or_set_put(b, (uint64_t)w << 32 | (uint32_t)(ii + 1), ol);",
     "int absent = q[ii + 1] == g->seq[w][ol] ? or_set_put(b, (uint64_t)w << 32 | (uint32_t)(ii + 1), ol) : 1;"),
    ("the A-queue last in, first out",
     "odiag_t t = b->A[b->head_A];\n\t++b->head_A, --b->n_A;",
     "odiag_t t = b->A[b->head_A + b->n_A - 1];\n\t--b->n_A;"),
]


def _compile_plant(orc, tmp_path, name, old, new):
    import ctypes as C
    import subprocess
    src = open(os.path.join(os.path.dirname(orc.__file__), "gwf_oracle.c")).read()
    assert src.count(old) == 1, name
    path = tmp_path / f"{name}.c"
    path.write_text(src.replace(old, new))
    so = tmp_path / f"{name}.so"
    subprocess.run(["gcc", "-O1", "-fPIC", "-shared", "-w", f"-I{os.path.dirname(orc.__file__)}", "-o", str(so),
                    str(path), "-lpthread"], check=True)
    lib = C.CDLL(str(so))
    lib.oracle_gwf_ed.argtypes = [C.c_void_p, C.c_int32, C.c_char_p, C.c_int32, C.c_int32, C.c_uint32, C.c_void_p]
    lib.oracle_gwf_ed.restype = C.c_int32
    return lib


def test_mixed_pin_rejects_planted_mistakes(gw, orc, tmp_path):
    import ctypes as C
    rng = random.Random(43)
    cases = [_mixed_case(gw, rng, dag) for dag in (True, False) for _ in range(600)]
    hand = _claim_order_cases(gw)
    for i, (what, old, new) in enumerate(MIXED_PLANTS):
        lib = _compile_plant(orc, tmp_path, f"mixed{i}", old, new)
        wrong = 0
        for g, q in cases:
            st = orc.OracleStats()
            qb = q.encode()
            e = lib.oracle_gwf_ed(C.cast(g.ptr, C.c_void_p), len(qb), qb, 0, -1, 0, C.byref(st))
            d = orc.dp(g, q)
            wrong += (e != d) if st.vis_lossy == 0 else (e < d)
        for g, q, want in hand:
            qb = q.encode()
            wrong += lib.oracle_gwf_ed(C.cast(g.ptr, C.c_void_p), len(qb), qb, 0, -1, 0, None) != want
        assert wrong > 0, f"planted mistake not caught: {what}"


# ---- v1 >= 0: the domain where the reference algorithm's end-at-v1 result is
# exact.  The vertex-end insertion is skipped when every out-arc matches the
# next query base (gwf-ed.cu:1225).  At any vertex other than v1 that only
# removes paths dominated by the matching entries, but at v1 it removes the
# path that inserts the rest of the query after v1's last base.  So whenever
# v1 has no out-arc, or its out-arcs do not all start with the same base, the
# skip cannot fire at v1 and the result must be the exact end-at-v1 DP.

def _v1_domain_case(gw, rng, dag):
    while True:
        g, seqs = rand_graph(gw, rng, rng.randint(1, 12), dag)
        outs = {}
        for a, b, o in g.arcs():
            outs.setdefault(a, []).append(seqs[b][o])
        ok = [v for v in range(g.n_vtx) if v not in outs or len(set(outs[v])) >= 2]
        if ok:
            return g, seqs, rng.choice(ok), rand_read(rng, rng.randint(0, 25))


@pytest.mark.parametrize("dag", [True, False])
def test_v1_exact_where_the_skip_cannot_fire_at_v1(gw, orc, dag):
    rng = random.Random(23 + dag)
    reach = 0
    for it in range(2500):
        g, seqs, v1, q = _v1_domain_case(gw, rng, dag)
        want = orc.dp(g, q, v1=v1)
        assert orc.ed(g, q, v1=v1) == want, (it, seqs, g.arcs(), q, v1)
        reach += want >= 0
    assert reach > 800        # a good share of the cases have an end


def levenshtein(a, b):
    prev = list(range(len(b) + 1))
    for i, ca in enumerate(a, 1):
        cur = [i] + [0] * len(b)
        for j, cb in enumerate(b, 1):
            cur[j] = min(prev[j] + 1, cur[j - 1] + 1, prev[j - 1] + (ca != cb))
        prev = cur
    return prev[-1]


def test_v1_closed_forms(gw, orc):
    rng = random.Random(29)
    # a read spelling a path exactly, ending at the last base of its last vertex -> 0
    for _ in range(200):
        g, seqs = rand_graph(gw, rng, rng.randint(2, 10), dag=True, ovl=rng.random() < 0.5)
        outs = {}
        for a, b, o in g.arcs():
            outs.setdefault(a, []).append((b, o))
        v, q = 0, seqs[0]
        for _ in range(rng.randint(0, 6)):
            if v not in outs:
                break
            v, o = rng.choice(outs[v])
            q += seqs[v][o:]
        assert orc.ed(g, q, v1=v) == 0 == orc.dp(g, q, v1=v)
    # a chain ending at its last vertex (a sink): the global edit distance to the
    # whole chain's sequence; at an earlier vertex never below that vertex's prefix
    for _ in range(60):
        seqs = [rand_read(rng, rng.randint(1, 8)) for _ in range(rng.randint(1, 6))]
        g = gw.Graph.from_lists(seqs, [(i, i + 1) for i in range(len(seqs) - 1)])
        g.index()
        q = rand_read(rng, rng.randint(0, 30))
        last = len(seqs) - 1
        assert orc.ed(g, q, v1=last) == levenshtein(q, "".join(seqs))
        j = rng.randrange(len(seqs))
        assert orc.ed(g, q, v1=j) >= levenshtein(q, "".join(seqs[:j + 1]))


# Plausible mistakes in the oracle's end-vertex branches (gwf_oracle.c, the
# restatement of gwf-ed.cu:1231-1254), each planted in a copy compiled here;
# the exact-domain check above must reject every one.
V1_PLANTS = [
    ("no deletion after a query end mid-vertex (gwf-ed.cu:1237-1241)",
     "\t\t\tor_B_push(b, v, d - 1, k + 1, x0 + 1, ooo);\n\t\t} else if ((int32_t)v != v1) {",
     "\t\t\t;\n\t\t} else if ((int32_t)v != v1) {"),
    ("that deletion on the wrong diagonal",
     "\t\t\tor_B_push(b, v, d - 1, k + 1, x0 + 1, ooo);\n\t\t} else if ((int32_t)v != v1) {",
     "\t\t\tor_B_push(b, v, d, k + 1, x0 + 1, ooo);\n\t\t} else if ((int32_t)v != v1) {"),
    ("no deletion into the next vertices at a vertex end (gwf-ed.cu:1244-1254)",
     "\t\t\t\tor_B_push(b, w, ii - ol, ol, x0 + 1, 1);\n\t\t\t}\n\t\t}\n\t}",
     "\t\t\t\t(void)w, (void)ol;\n\t\t\t}\n\t\t}\n\t}"),
    ("those deletions through the first out-arc only",
     "nv = (int32_t)g->aux[v], j;\n\t\t\tfor (j = 0; j < nv; ++j) {",
     "nv = (int32_t)g->aux[v], j;\n\t\t\tfor (j = 0; j < (nv < 1 ? nv : 1); ++j) {"),
    ("an end anywhere in v1 instead of at its last base (gwf-ed.cu:1231)",
     "((int32_t)v == v1 && k + 1 == vl)", "((int32_t)v == v1)"),
]


def test_v1_pin_rejects_planted_mistakes(gw, orc, tmp_path):
    import ctypes as C
    rng = random.Random(31)
    cases = [_v1_domain_case(gw, rng, dag) for dag in (True, False) for _ in range(700)]
    for i, (what, old, new) in enumerate(V1_PLANTS):
        lib = _compile_plant(orc, tmp_path, f"v1_{i}", old, new)
        wrong = 0
        for g, _, v1, q in cases:
            qb = q.encode()
            wrong += lib.oracle_gwf_ed(C.cast(g.ptr, C.c_void_p), len(qb), qb, 0, v1, 0, None) != orc.dp(g, q, v1=v1)
        assert wrong > 0, f"planted mistake not caught: {what}"
