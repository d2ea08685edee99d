"""GPU parity: every ED from the HIP path (through the C-ABI) must equal the
oracle's (the CPU restatement of gwf_ed) read by read, bit-exact."""
import os
import random

import numpy as np
import pytest

from conftest import GOLDEN, read_ints

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _gpu_eds(gw, graph, reads, kernel=0):
    dg = gw.DeviceGraph(graph)
    dr = gw.DeviceReads(reads)
    try:
        return gw.ed_run(dg, dr, kernel=kernel)
    finally:
        dr.free()
        dg.free()


KERNELS = [0, 1]   # plain, persistent


@pytest.mark.parametrize("kernel", KERNELS)
def test_fixture_1000_reads(gpu, kernel):
    _say(f"case: exposed_testing/1000_reads (golden EDs), kernel {kernel}")
    g = gpu.Graph.from_gfa(os.path.join(GOLDEN, "graph_1000.gfa"))
    r = gpu.Reads.from_fasta(os.path.join(GOLDEN, "reads_1000.fa.gz"))
    want = np.array(read_ints(os.path.join(GOLDEN, "reads_1000.ed")))
    got, st = _gpu_eds(gpu, g, r, kernel)
    bad = np.nonzero(got != want)[0]
    assert bad.size == 0, f"{bad.size} mismatches, first {bad[:5]}: gpu {got[bad[:5]]} want {want[bad[:5]]}"
    assert st["sum_steps"] == int(want.sum()) + len(want)


@pytest.mark.parametrize("kernel", KERNELS)
def test_fixture_grafo_test(gpu, kernel):
    _say(f"case: exposed_testing/generated/grafo_test.gfa + sampled reads (golden EDs), kernel {kernel}")
    g = gpu.Graph.from_gfa(os.path.join(GOLDEN, "grafo_test.gfa"))
    r = gpu.Reads.from_fasta(os.path.join(GOLDEN, "grafo_reads.fa"))
    want = np.array(read_ints(os.path.join(GOLDEN, "grafo_reads.ed")))
    got, _ = _gpu_eds(gpu, g, r, kernel)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("c,n", [(1, 1000), (2, 240), (3, 300), (4, 12)])
def test_configs_vs_oracle(gpu, orc, c, n, kernel):
    _say(f"case: config {c}, first {n} reads, kernel {kernel}")
    g, r = gpu.make_workload(c, n)
    g.index()
    want, ost, _ = orc.batch(g, r)
    got, st = _gpu_eds(gpu, g, r, kernel)
    bad = np.nonzero(got != want)[0]
    assert bad.size == 0, f"config {c}: {bad.size} mismatches, first {bad[:5]}: gpu {got[bad[:5]]} oracle {want[bad[:5]]}"
    assert st["sum_steps"] == ost["steps"]


def _random_graph(gw, rng, nv, dag=True, ovl=False):
    seqs = []
    for _ in range(nv):
        L = rng.randint(1, 3 if rng.random() < 0.5 else 30)
        alph = "ACGT"[: rng.randint(2, 4)]
        seqs.append("".join(rng.choice(alph) for _ in range(L)))
    fixed_o = [((w * 7 + 3) % len(seqs[w])) if ovl else 0 for w in range(nv)]
    arcs = set()
    for v in range(nv):
        for _ in range(rng.randint(0, 2)):
            if dag:
                if v + 1 >= nv:
                    break
                w = rng.randint(v + 1, nv - 1)
            else:
                w = rng.randrange(nv)
            arcs.add((v, w))
    return gw.Graph.from_lists(seqs, [(v, w, fixed_o[w]) for v, w in sorted(arcs)]), seqs


def _sample(rng, graph_seqs, arcs_of, ql, err):
    v, p, out = 0, 0, []
    while len(out) < ql:
        x = rng.random()
        if x < err / 3:
            out.append(rng.choice("ACGT"))
            continue
        c = graph_seqs[v][p]
        if x < 2 * err / 3:
            out.append(rng.choice("ACGT"))
        elif x >= err:
            out.append(c)
        p += 1
        if p >= len(graph_seqs[v]):
            nxt = arcs_of.get(v, [])
            if not nxt:
                break
            v, p = rng.choice(nxt)
    return "".join(out)


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("seed,dag,ovl", [(1, True, False), (2, False, False), (3, True, True), (4, False, True)])
def test_random_graphs_vs_oracle(gpu, orc, seed, dag, ovl, kernel):
    _say(f"case: random graphs seed={seed} dag={dag} ovl={ovl} kernel {kernel}")
    rng = random.Random(seed)
    for it in range(25):
        g, seqs = _random_graph(gpu, rng, rng.randint(1, 40), dag, ovl)
        arcs_of = {}
        for v, w, o in g.arcs():
            arcs_of.setdefault(v, []).append((w, o))
        reads = [_sample(rng, seqs, arcs_of, rng.randint(0, 300), rng.choice([0.0, 0.02, 0.1, 0.3]))
                 for _ in range(20)]
        reads += ["", "ACGTTGCA" * 3, "".join(rng.choice("ACGT") for _ in range(rng.randint(1, 200)))]
        r = gpu.Reads.from_list(reads)
        g.index()
        want, _, _ = orc.batch(g, r)
        got, _ = _gpu_eds(gpu, g, r, kernel)
        assert np.array_equal(got, want), f"iteration {it}: gpu {got.tolist()} oracle {want.tolist()}"


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("seed,dag,ovl", [(11, True, False), (12, False, True)])
def test_high_degree_vertices_vs_oracle(gpu, orc, seed, dag, ovl, kernel):
    """Vertices with up to 11 out-arcs (the closure takes a state's arcs in
    groups of four, so several groups per state), parallel alleles of one or
    a few bases, against the oracle."""
    _say(f"case: high-degree graphs seed={seed} dag={dag} ovl={ovl} kernel {kernel}")
    rng = random.Random(seed)
    for it in range(12):
        nv = rng.randint(6, 40)
        seqs = ["".join(rng.choice("ACGT") for _ in range(rng.randint(1, 4 if rng.random() < 0.6 else 25)))
                for _ in range(nv)]
        fixed_o = [((w * 5 + 1) % len(seqs[w])) if ovl else 0 for w in range(nv)]
        arcs = set()
        for v in range(nv):
            for _ in range(rng.randint(0, 11)):
                if dag:
                    if v + 1 >= nv:
                        break
                    w = rng.randint(v + 1, nv - 1)
                else:
                    w = rng.randrange(nv)
                arcs.add((v, w))
        g = gpu.Graph.from_lists(seqs, [(v, w, fixed_o[w]) for v, w in sorted(arcs)])
        arcs_of = {}
        for v, w, o in g.arcs():
            arcs_of.setdefault(v, []).append((w, o))
        assert max(len(a) for a in arcs_of.values()) > 4
        reads = [_sample(rng, seqs, arcs_of, rng.randint(1, 250), rng.choice([0.0, 0.02, 0.1, 0.3]))
                 for _ in range(24)]
        r = gpu.Reads.from_list(reads)
        g.index()
        want, _, _ = orc.batch(g, r)
        got, _ = _gpu_eds(gpu, g, r, kernel)
        assert np.array_equal(got, want), f"iteration {it}: gpu {got.tolist()} oracle {want.tolist()}"


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("seed,cyclic", [(31, False), (32, True)])
def test_single_in_arc_chains_vs_oracle(gpu, orc, seed, cyclic, kernel):
    """Diamonds whose two branches carry the same sequence (both reach the
    merge vertex at one query position), followed by chains of vertices with
    a single in-arc, which the closure crosses without visited-set claims;
    optionally a back arc from the last chain vertex to the first diamond.
    Repeated frontier states then run down the chains twice: the EDs must
    still equal the oracle's."""
    _say(f"case: single-in-arc chains seed={seed} cyclic={cyclic} kernel {kernel}")
    rng = random.Random(seed)
    for it in range(10):
        seqs, arcs = [], []
        prev = None
        for _ in range(rng.randint(2, 6)):
            split = len(seqs)
            seqs.append("".join(rng.choice("ACGT") for _ in range(rng.randint(1, 6))))
            br = "".join(rng.choice("ACGT") for _ in range(rng.randint(1, 5)))
            seqs += [br, br if rng.random() < 0.8 else br[::-1]]
            merge = len(seqs)
            seqs.append("".join(rng.choice("ACGT") for _ in range(rng.randint(1, 6))))
            arcs += [(split, split + 1), (split, split + 2), (split + 1, merge), (split + 2, merge)]
            last = merge
            for _ in range(rng.randint(1, 4)):  # single-in-arc chain
                seqs.append("".join(rng.choice("ACGT") for _ in range(rng.randint(1, 8))))
                arcs.append((last, len(seqs) - 1))
                last = len(seqs) - 1
            if prev is not None:
                arcs.append((prev, split))
            prev = last
        if cyclic:
            arcs.append((prev, 0))
        g = gpu.Graph.from_lists(seqs, [(v, w, 0) for v, w in arcs])
        arcs_of = {}
        for v, w, o in g.arcs():
            arcs_of.setdefault(v, []).append((w, o))
        reads = [_sample(rng, seqs, arcs_of, rng.randint(1, 200), rng.choice([0.0, 0.05, 0.2])) for _ in range(24)]
        r = gpu.Reads.from_list(reads)
        g.index()
        want, _, _ = orc.batch(g, r)
        got, _ = _gpu_eds(gpu, g, r, kernel)
        assert np.array_equal(got, want), f"iteration {it}: gpu {got.tolist()} oracle {want.tolist()}"


@pytest.mark.parametrize("kernel", KERNELS)
def test_edge_cases(gpu, orc, kernel):
    _say(f"case: edge cases, kernel {kernel}")
    # short start vertex (< 32 bp) read as an exact prefix: the reference's
    # unsigned max_k skips this extension; exact semantics give 0
    g = gpu.Graph.from_lists(["ACGTACGTAC", "GGTTAACC", "TTTT"], [(0, 1), (1, 2)])
    cases = [("ACGTACGTACGG", 0), ("", 0), ("CCCCCCCCCCCC", None), ("ACGTACGTACGGTTAACCTTTTAAAA", None),
             ("T" * 50, None)]
    r = gpu.Reads.from_list([c for c, _ in cases])
    g.index()
    want, _, _ = orc.batch(g, r)
    got, _ = _gpu_eds(gpu, g, r, kernel)
    assert np.array_equal(got, want)
    for (q, exp), x in zip(cases, got):
        if exp is not None:
            assert x == exp, q
    # ED == ql through an all-insertion end (the reference's score loop would spin)
    g2 = gpu.Graph.from_lists(["A", "AAAA"], [(0, 1)])
    r2 = gpu.Reads.from_list(["TT", "TTTTTTTT"])
    got2, _ = _gpu_eds(gpu, g2, r2, kernel)
    assert got2.tolist() == [2, 8]


def test_per_read_start_vertex(gpu, orc):
    _say("case: per-read v0")
    g, r = gpu.make_workload(2, 60)
    g.index()
    assert (r.v0 > 0).any()
    want, _, _ = orc.batch(g, r)
    got, _ = _gpu_eds(gpu, g, r)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("kernel", KERNELS)
def test_workspace_overflow_retry(gpu, orc, monkeypatch, kernel):
    _say(f"case: forced tiny workspace -> overflow -> re-run, kernel {kernel}")
    monkeypatch.setenv("GWF_DEBUG_CAP_A", "16")
    monkeypatch.setenv("GWF_PERSIST_M", "64")
    g, r = gpu.make_workload(2, 30)
    g.index()
    want, _, _ = orc.batch(g, r)
    got, st = _gpu_eds(gpu, g, r, kernel)
    assert st["n_retry"] > 0
    assert np.array_equal(got, want)


def test_lds_spill_to_hbm(gpu, orc, monkeypatch):
    _say("case: persistent kernel, small LDS capacity -> wavefronts spill to HBM mid-read")
    monkeypatch.setenv("GWF_PERSIST_M", "128")
    monkeypatch.setenv("GWF_PERSIST_NW", "1")
    g, r = gpu.make_workload(2, 60)
    g.index()
    want, _, _ = orc.batch(g, r)
    got, st = _gpu_eds(gpu, g, r, 1)
    assert st["n_spill"] > 0 and st["n_retry"] == 0
    assert np.array_equal(got, want)
    # steps run in HBM: some, fewer than all (each spilled read starts in LDS)
    assert 0 < st["sum_hbm_steps"] < st["sum_steps"]


@pytest.mark.parametrize("nw,margin", [(1, 1 << 20), (2, 1 << 20), (2, 0)])
def test_hbm_staging_fallbacks(gpu, orc, monkeypatch, nw, margin):
    """Spilled reads with an LDS scratch too small for the closure
    successors, the history merge, Q or the history (sorted and merged in
    HBM, kept states compacted through the workspace), or as large as the
    occupancy allows (margin 0: the arena may keep a workgroup per CU out,
    which only delays it)."""
    _say(f"case: persistent kernel nw={nw} lds m=64 arena margin={margin}, config 4 reads")
    monkeypatch.setenv("GWF_PERSIST_NW", str(nw))
    monkeypatch.setenv("GWF_PERSIST_M", "64")
    monkeypatch.setenv("GWF_LDS_MARGIN", str(margin))
    g, r = gpu.make_workload(4, 8)
    g.index()
    want, _, _ = orc.batch(g, r)
    got, st = _gpu_eds(gpu, g, r, 1)
    assert st["n_spill"] == len(want)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("nw,m", [(1, 256), (2, 512), (4, 2048)])
def test_persistent_configurations(gpu, orc, monkeypatch, nw, m):
    _say(f"case: persistent kernel nw={nw} lds m={m}")
    monkeypatch.setenv("GWF_PERSIST_NW", str(nw))
    monkeypatch.setenv("GWF_PERSIST_M", str(m))
    g, r = gpu.make_workload(3, 200)
    g.index()
    want, _, _ = orc.batch(g, r)
    got, st = _gpu_eds(gpu, g, r, 1)
    assert np.array_equal(got, want)


def test_gwf_ed_single_call(gpu, orc):
    _say("case: gwf_ed host entry point (reference signature)")
    g, r = gpu.make_workload(1, 20)
    g.index()
    for i in range(len(r)):
        q = r.read(i)
        assert gpu.gwf_ed(g, q) == orc.ed(g, q)


def _mixed_graph(gw, rng, nv, dag):
    """A random graph whose arcs carry their own overlaps, so that vertices get
    in-arcs with different overlaps."""
    seqs = ["".join(rng.choice("ACGT"[:rng.randint(2, 4)]) for _ in range(rng.randint(1, 3 if rng.random() < 0.4 else 20)))
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
    return gw.Graph.from_lists(seqs, sorted(arcs)), seqs


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("seed,dag", [(41, True), (42, False)])
def test_mixed_in_overlaps_vs_oracle(gpu, orc, seed, dag, kernel):
    """Vertices with in-arcs of different overlaps: the reference's answer
    depends on the order in which its FIFO closure claims visited keys; the
    GPU follows that order (and one visited set per read), so every ED equals
    the oracle's, with v1 < 0 and v1 >= 0."""
    _say(f"case: mixed in-overlaps, random graphs seed={seed} dag={dag} kernel {kernel}")
    rng = random.Random(seed)
    n_mixed = 0
    for it in range(20):
        g, seqs = _mixed_graph(gpu, rng, rng.randint(2, 30), dag)
        arcs_of = {}
        for v, w, o in g.arcs():
            arcs_of.setdefault(v, []).append((w, o))
        n_mixed += len({(w, o) for _, w, o in g.arcs()}) > len({w for _, w, _ in g.arcs()})
        reads = [_sample(rng, seqs, arcs_of, rng.randint(0, 200), rng.choice([0.0, 0.05, 0.2])) for _ in range(15)]
        reads += ["".join(rng.choice("ACGT") for _ in range(rng.randint(1, 60)))]
        r = gpu.Reads.from_list(reads)
        g.index()
        for v1 in (-1, rng.randrange(g.n_vtx)):
            want, _, _ = orc.batch(g, r, v1=v1)
            dg, dr = gpu.DeviceGraph(g), gpu.DeviceReads(r)
            try:
                got, _ = gpu.ed_run(dg, dr, kernel=kernel, v1=v1)
            finally:
                dr.free()
                dg.free()
            assert np.array_equal(got, want), f"iteration {it} v1={v1}: gpu {got.tolist()} oracle {want.tolist()}"
    assert n_mixed >= 10


def test_mixed_in_overlaps_order_dependent_case(gpu, orc):
    """The graph of test_oracle's order-dependence case, in both labellings:
    the reference answers 12 and 11, and so must the GPU."""
    _say("case: mixed in-overlaps, the order-dependent case in two labellings")
    seqs = ["CAA", "ACC", "GTGG"]
    arcs = [(0, 1, 0), (1, 1, 2), (1, 2, 3), (1, 1, 1), (2, 0, 2), (2, 2, 1), (2, 1, 2)]
    q = "GCGCTCGGAAGTATTATT"
    perm = [0, 2, 1]
    inv = {old: new for new, old in enumerate(perm)}
    for s_, a_ in ((seqs, arcs), ([seqs[o] for o in perm], [(inv[a], inv[b], o) for a, b, o in arcs])):
        g = gpu.Graph.from_lists(s_, a_)
        g.index()
        want = orc.ed(g, q)
        for kernel in KERNELS:
            dg, dr = gpu.DeviceGraph(g), gpu.DeviceReads(gpu.Reads.from_list([q]))
            try:
                got, _ = gpu.ed_run(dg, dr, kernel=kernel)
            finally:
                dr.free()
                dg.free()
            assert got.tolist() == [want]
        assert gpu.gwf_ed(g, q) == want


@pytest.mark.parametrize("env", [{"GWF_PERSIST_M": "64"}, {"GWF_DEBUG_CAP_A": "48"}])
def test_mixed_in_overlaps_spill_and_rerun(gpu, orc, monkeypatch, env):
    """Mixed in-overlaps through an LDS-to-HBM spill (the spilled step repeats
    its visited-set claims with the same stamps) and through workspace
    overflow re-runs (a full visited set re-runs the read with a larger one)."""
    _say(f"case: mixed in-overlaps, {env}")
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    rng = random.Random(43)
    for it in range(8):
        g, seqs = _mixed_graph(gpu, rng, rng.randint(10, 40), it % 2 == 0)
        arcs_of = {}
        for v, w, o in g.arcs():
            arcs_of.setdefault(v, []).append((w, o))
        reads = [_sample(rng, seqs, arcs_of, rng.randint(50, 300), 0.15) for _ in range(12)]
        r = gpu.Reads.from_list(reads)
        g.index()
        want, _, _ = orc.batch(g, r)
        dg, dr = gpu.DeviceGraph(g), gpu.DeviceReads(r)
        try:
            for kernel in KERNELS:
                got, st = gpu.ed_run(dg, dr, kernel=kernel)
                assert np.array_equal(got, want), f"iteration {it} kernel {kernel}"
        finally:
            dr.free()
            dg.free()


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("seed,dag,ovl", [(21, True, False), (22, False, False), (23, True, True), (24, False, True)])
def test_end_vertex_v1_vs_oracle(gpu, orc, seed, dag, ovl, kernel):
    """v1 >= 0 (end at the last base of v1, gwf-ed.cu:1231-1254): the GPU must
    reproduce the reference algorithm's result, which is not always the exact
    distance (the oracle is the reference restated, not the DP)."""
    _say(f"case: end vertex v1, random graphs seed={seed} dag={dag} ovl={ovl} kernel {kernel}")
    rng = random.Random(seed)
    for it in range(12):
        g, seqs = _random_graph(gpu, rng, rng.randint(2, 30), dag, ovl)
        arcs_of = {}
        for v, w, o in g.arcs():
            arcs_of.setdefault(v, []).append((w, o))
        reads = [_sample(rng, seqs, arcs_of, rng.randint(0, 150), rng.choice([0.0, 0.05, 0.2]))
                 for _ in range(12)]
        r = gpu.Reads.from_list(reads)
        g.index()
        for v1 in sorted({0, g.n_vtx - 1, rng.randrange(g.n_vtx)}):
            want, _, _ = orc.batch(g, r, v1=v1)
            dg, dr = gpu.DeviceGraph(g), gpu.DeviceReads(r)
            try:
                got, _ = gpu.ed_run(dg, dr, kernel=kernel, v1=v1)
            finally:
                dr.free()
                dg.free()
            assert np.array_equal(got, want), f"iteration {it} v1={v1}: gpu {got.tolist()} oracle {want.tolist()}"


def test_end_vertex_v1_config_graph(gpu, orc):
    _say("case: end vertex v1 on the small benchmark graph")
    g, r = gpu.make_workload(1, 40)
    g.index()
    dg, dr = gpu.DeviceGraph(g), gpu.DeviceReads(r)
    try:
        for v1 in (5, g.n_vtx // 3):
            want, _, _ = orc.batch(g, r, v1=v1)
            for kernel in KERNELS:
                got, _ = gpu.ed_run(dg, dr, kernel=kernel, v1=v1)
                assert np.array_equal(got, want), (v1, kernel)
        q = r.read(0)
        assert gpu.gwf_ed(g, q, v1=7) == orc.ed(g, q, v1=7)
    finally:
        dr.free()
        dg.free()


@pytest.mark.parametrize("kernel", KERNELS)
def test_end_vertex_v1_exact_domain_vs_dp(gpu, orc, kernel):
    """Where v1 has no out-arc, or out-arcs that do not all start with one base,
    the reference algorithm's end-at-v1 result is the exact distance (pinned by
    tests/test_oracle.py), so the GPU must give the end-at-v1 DP itself."""
    _say(f"case: end vertex v1 in the exact domain, GPU vs the DP, kernel {kernel}")
    from test_oracle import _v1_domain_case
    rng = random.Random(37)
    for it in range(30):
        g, seqs, v1, _ = _v1_domain_case(gpu, rng, dag=it % 2 == 0)
        arcs_of = {}
        for v, w, o in g.arcs():
            arcs_of.setdefault(v, []).append((w, o))
        reads = [_sample(rng, seqs, arcs_of, rng.randint(0, 60), rng.choice([0.0, 0.1, 0.3])) for _ in range(10)]
        want = [orc.dp(g, q, v1=v1) for q in reads]
        dg, dr = gpu.DeviceGraph(g), gpu.DeviceReads(gpu.Reads.from_list(reads))
        try:
            got, _ = gpu.ed_run(dg, dr, kernel=kernel, v1=v1)
        finally:
            dr.free()
            dg.free()
        assert got.tolist() == want, f"iteration {it} v1={v1}: gpu {got.tolist()} dp {want}"


def test_mixed_in_overlaps_order_sensitive_cases(gpu, orc):
    """Cases where the reference's answer changes when the vertices are
    relabelled (start kept at 0), i.e. where the order in which its closure
    claims visited keys decides the ED: the GPU must follow the order in both
    labellings."""
    rng = random.Random(44)
    found, tries = 0, 0
    while found < 25 and tries < 40000:
        tries += 1
        g, seqs = _mixed_graph(gpu, rng, rng.randint(3, 9), rng.random() < 0.5)
        arcs = g.arcs()
        g.index()
        q = "".join(rng.choice("ACGT") for _ in range(rng.randint(5, 25)))
        perm = [0] + rng.sample(range(1, len(seqs)), len(seqs) - 1)
        inv = {old: new for new, old in enumerate(perm)}
        g2 = gpu.Graph.from_lists([seqs[o] for o in perm], [(inv[a], inv[b], o) for a, b, o in arcs])
        g2.index()
        a, b = orc.ed(g, q), orc.ed(g2, q)
        if a == b:
            continue
        found += 1
        _say(f"case: order-sensitive mixed-overlap graph {found} (try {tries}): oracle {a} / {b}")
        for gg, want in ((g, a), (g2, b)):
            for kernel in KERNELS:
                dg, dr = gpu.DeviceGraph(gg), gpu.DeviceReads(gpu.Reads.from_list([q]))
                try:
                    got, _ = gpu.ed_run(dg, dr, kernel=kernel)
                finally:
                    dr.free()
                    dg.free()
                assert got.tolist() == [want], (seqs, arcs, q, kernel)
    assert found == 25


def test_mixed_claim_order_hand_cases(gpu, orc):
    """The hand-derived claim-order cases of test_oracle (which arc wins a
    visited key: CSR order within a vertex, first in first out across the
    closure queue, the key claimed before the first-base test)."""
    _say("case: mixed in-overlaps, hand-derived claim-order cases")
    from test_oracle import _claim_order_cases
    for g, q, want in _claim_order_cases(gpu):
        for kernel in KERNELS:
            dg, dr = gpu.DeviceGraph(g), gpu.DeviceReads(gpu.Reads.from_list([q]))
            try:
                got, _ = gpu.ed_run(dg, dr, kernel=kernel)
            finally:
                dr.free()
                dg.free()
            assert got.tolist() == [want], (g.arcs(), kernel)
        assert gpu.gwf_ed(g, q) == want
