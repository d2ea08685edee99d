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


def test_fixture_1000_reads(gpu):
    _say("case: exposed_testing/1000_reads (golden EDs)")
    g = gpu.Graph.from_gfa(os.path.join(GOLDEN, "graph_1000.gfa"))
    r = gpu.Reads.from_fasta(os.path.join(GOLDEN, "reads_1000.fa.gz"))
    want = np.array(read_ints(os.path.join(GOLDEN, "reads_1000.ed")))
    got, st = _gpu_eds(gpu, g, r)
    bad = np.nonzero(got != want)[0]
    assert bad.size == 0, f"{bad.size} mismatches, first {bad[:5]}: gpu {got[bad[:5]]} want {want[bad[:5]]}"
    assert st["sum_steps"] == int(want.sum()) + len(want)


def test_fixture_grafo_test(gpu):
    _say("case: exposed_testing/generated/grafo_test.gfa + sampled reads (golden EDs)")
    g = gpu.Graph.from_gfa(os.path.join(GOLDEN, "grafo_test.gfa"))
    r = gpu.Reads.from_fasta(os.path.join(GOLDEN, "grafo_reads.fa"))
    want = np.array(read_ints(os.path.join(GOLDEN, "grafo_reads.ed")))
    got, _ = _gpu_eds(gpu, g, r)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("c,n", [(1, 1000), (2, 240), (3, 300), (4, 12)])
def test_configs_vs_oracle(gpu, orc, c, n):
    _say(f"case: config {c}, first {n} reads")
    g, r = gpu.make_workload(c, n)
    g.index()
    want, ost, _ = orc.batch(g, r)
    got, st = _gpu_eds(gpu, g, r)
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


@pytest.mark.parametrize("seed,dag,ovl", [(1, True, False), (2, False, False), (3, True, True), (4, False, True)])
def test_random_graphs_vs_oracle(gpu, orc, seed, dag, ovl):
    _say(f"case: random graphs seed={seed} dag={dag} ovl={ovl}")
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
        got, _ = _gpu_eds(gpu, g, r)
        assert np.array_equal(got, want), f"iteration {it}: gpu {got.tolist()} oracle {want.tolist()}"


def test_edge_cases(gpu, orc):
    _say("case: edge cases")
    # short start vertex (< 32 bp) read as an exact prefix: the reference's
    # unsigned max_k skips this extension; exact semantics give 0
    g = gpu.Graph.from_lists(["ACGTACGTAC", "GGTTAACC", "TTTT"], [(0, 1), (1, 2)])
    cases = [("ACGTACGTACGG", 0), ("", 0), ("CCCCCCCCCCCC", None), ("ACGTACGTACGGTTAACCTTTTAAAA", None),
             ("T" * 50, None)]
    r = gpu.Reads.from_list([c for c, _ in cases])
    g.index()
    want, _, _ = orc.batch(g, r)
    got, _ = _gpu_eds(gpu, g, r)
    assert np.array_equal(got, want)
    for (q, exp), x in zip(cases, got):
        if exp is not None:
            assert x == exp, q
    # ED == ql through an all-insertion end (the reference's score loop would spin)
    g2 = gpu.Graph.from_lists(["A", "AAAA"], [(0, 1)])
    r2 = gpu.Reads.from_list(["TT", "TTTTTTTT"])
    got2, _ = _gpu_eds(gpu, g2, r2)
    assert got2.tolist() == [2, 8]


def test_per_read_start_vertex(gpu, orc):
    _say("case: per-read v0")
    g, r = gpu.make_workload(2, 60)
    g.index()
    assert (r.v0 > 0).any()
    want, _, _ = orc.batch(g, r)
    got, _ = _gpu_eds(gpu, g, r)
    assert np.array_equal(got, want)


def test_workspace_overflow_retry(gpu, orc, monkeypatch):
    _say("case: forced tiny workspace -> overflow -> re-run")
    monkeypatch.setenv("GWF_DEBUG_CAP_A", "16")
    g, r = gpu.make_workload(2, 30)
    g.index()
    want, _, _ = orc.batch(g, r)
    got, st = _gpu_eds(gpu, g, r)
    assert st["n_retry"] > 0
    assert np.array_equal(got, want)


def test_gwf_ed_single_call(gpu, orc):
    _say("case: gwf_ed host entry point (reference signature)")
    g, r = gpu.make_workload(1, 20)
    g.index()
    for i in range(len(r)):
        q = r.read(i)
        assert gpu.gwf_ed(g, q) == orc.ed(g, q)


def test_mixed_overlaps_rejected(gpu):
    _say("case: mixed in-overlaps rejected")
    g = gpu.Graph.from_lists(["ACGT", "CCGT", "GATTACA"], [(0, 2, 0), (1, 2, 2), (0, 1, 0)])
    with pytest.raises(gpu.GwfError):
        gpu.DeviceGraph(g)
