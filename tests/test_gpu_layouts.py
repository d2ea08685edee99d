"""GPU parity across the state-key layouts.

A wavefront state (vertex, diagonal, offset) is one 64-bit key; the kernels
are built for four splits of it (vertex/diagonal/offset bits 24/20/20,
28/18/18, 20/22/22 and 32/16/16).  A run takes the first split that holds
the graph's vertex count and vertex lengths and the reads' lengths, so graphs
beyond 2^24 vertices and reads or segments beyond 2^19 bases run (the
reference packs a diagonal as a 32-bit vertex id over diagonal + 2^26:
gwfa.h:27-31, gwf-ed.cu:80-83).  The
environment variable GWF_KEY_LAYOUT forces one split, which is how every
split is checked here on ordinary workloads.  Every ED must equal the
oracle's, bit-exact."""
import random

import numpy as np
import pytest

from test_gpu_parity import KERNELS, _gpu_eds, _mixed_graph, _random_graph, _sample, _say

pytestmark = pytest.mark.gpu

LAYOUTS = [0, 1, 2, 3]


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("layout", LAYOUTS)
def test_forced_layout_configs_vs_oracle(gpu, orc, monkeypatch, layout, kernel):
    _say(f"case: key layout {layout}, config 2 (240 reads) and config 4 (12 reads), kernel {kernel}")
    monkeypatch.setenv("GWF_KEY_LAYOUT", str(layout))
    for c, n in ((2, 240), (4, 12)):
        g, r = gpu.make_workload(c, n)
        g.index()
        want, ost, _ = orc.batch(g, r)
        got, st = _gpu_eds(gpu, g, r, kernel)
        bad = np.nonzero(got != want)[0]
        assert bad.size == 0, f"layout {layout} config {c}: {bad.size} mismatches, first {bad[:5]}"
        assert st["sum_steps"] == ost["steps"]


@pytest.mark.parametrize("layout", LAYOUTS)
def test_forced_layout_random_graphs_vs_oracle(gpu, orc, monkeypatch, layout):
    """Overlapping arcs, mixed in-overlaps and end vertices under each split."""
    _say(f"case: key layout {layout}, random overlap / mixed graphs with v1")
    monkeypatch.setenv("GWF_KEY_LAYOUT", str(layout))
    rng = random.Random(300 + layout)
    for it in range(10):
        if it % 2:
            g, seqs = _mixed_graph(gpu, rng, rng.randint(2, 30), it % 4 == 1)
        else:
            g, seqs = _random_graph(gpu, rng, rng.randint(2, 30), it % 4 == 0, True)
        arcs_of = {}
        for v, w, o in g.arcs():
            arcs_of.setdefault(v, []).append((w, o))
        reads = [_sample(rng, seqs, arcs_of, rng.randint(0, 200), rng.choice([0.0, 0.05, 0.2])) for _ in range(12)]
        r = gpu.Reads.from_list(reads)
        g.index()
        for v1 in (-1, rng.randrange(g.n_vtx)):
            want, _, _ = orc.batch(g, r, v1=v1)
            for kernel in KERNELS:
                dg, dr = gpu.DeviceGraph(g), gpu.DeviceReads(r)
                try:
                    got, _ = gpu.ed_run(dg, dr, kernel=kernel, v1=v1)
                finally:
                    dr.free()
                    dg.free()
                assert np.array_equal(got, want), \
                    f"layout {layout} iteration {it} v1={v1} kernel {kernel}: gpu {got.tolist()} oracle {want.tolist()}"


def _mutate(rng, s, err):
    """s with substitutions, insertions and deletions at rate err (bytes)."""
    out = bytearray()
    i = 0
    while i < len(s):
        x = rng.random()
        if x < err / 3:
            out.append(rng.choice(b"ACGT"))
            i += 1
        elif x < 2 * err / 3:
            out.append(rng.choice(b"ACGT"))
        elif x < err:
            i += 1
        else:
            out.append(s[i])
            i += 1
    return bytes(out)


def _long_graph(gpu, rng):
    """A chain of five 130 kb segments, a 700 kb segment (longer than the
    2^19 - 1 bases of the default split) and a short tail."""
    seqs = [bytes(rng.choice(b"ACGT") for _ in range(130_000)) for _ in range(5)]
    seqs.append(bytes(rng.choice(b"ACGT") for _ in range(700_000)))
    seqs.append(b"ACGTACGTAC")
    arcs = [(v, v + 1, 0) for v in range(len(seqs) - 1)]
    return gpu.Graph.from_lists(seqs, arcs), seqs


def test_long_reads_and_segments_vs_oracle(gpu, orc, monkeypatch):
    """Reads of 600 kb and 1.1 Mb and a 700 kb segment: beyond the default
    split, they run in the 20/22/22 one; forcing the default is an error."""
    _say("case: long reads (600 kb, 1.1 Mb) and a 700 kb segment")
    rng = random.Random(77)
    g, seqs = _long_graph(gpu, rng)
    path = b"".join(seqs)
    reads = [_mutate(rng, path[:600_000], 0.002), _mutate(rng, path[:1_100_000], 0.001),
             _mutate(rng, seqs[5][:3000], 0.05), _mutate(rng, seqs[0][:20_000], 0.01)]
    v0 = [0, 0, 5, 0]
    r = gpu.Reads.from_list(reads, v0=v0)
    g.index()
    want, _, _ = orc.batch(g, r)
    assert want[0] > 0 and want[1] > 0
    for kernel in KERNELS:
        got, _ = _gpu_eds(gpu, g, r, kernel)
        assert np.array_equal(got, want), f"kernel {kernel}: gpu {got.tolist()} oracle {want.tolist()}"
    assert gpu.gwf_ed(g, reads[0]) == want[0]
    dg = gpu.DeviceGraph(g)
    try:
        with gpu.Server(dg, state_budget_bytes=8 << 30) as srv:
            ids = srv.submit(r)
            res = srv.collect(len(r), timeout_s=300)
        assert res["id"].tolist() == ids.tolist()
        assert res["status"].tolist() == [0] * len(r)
        assert res["ed"].tolist() == want.tolist()
        monkeypatch.setenv("GWF_KEY_LAYOUT", "0")
        dr = gpu.DeviceReads(r)
        try:
            with pytest.raises(gpu.GwfError, match="layout"):
                gpu.ed_run(dg, dr)
        finally:
            dr.free()
    finally:
        dg.free()


def _big_graph(gpu, rng, n_vtx, sub):
    """n_vtx vertices: single-base filler vertices without arcs, then a random
    graph of `sub` vertices on the top ids (its arcs cross 2^24)."""
    lo = n_vtx - sub
    sseq = [bytes(rng.choice(b"ACGT") for _ in range(rng.randint(1, 30))) for _ in range(sub)]
    arcs = set()
    for v in range(sub):
        for _ in range(rng.randint(1, 3)):
            w = rng.randrange(sub)
            arcs.add((v, w, rng.randrange(len(sseq[w])) if rng.random() < 0.3 else 0))
    lens = np.ones(n_vtx, np.uint32)
    lens[lo:] = [len(s) for s in sseq]
    seq = b"A" * lo + b"".join(sseq)
    arcs = sorted(arcs)
    src = np.array([lo + a for a, _, _ in arcs], np.uint32)
    dst = np.array([lo + b for _, b, _ in arcs], np.uint32)
    ovl = np.array([o for _, _, o in arcs], np.int32)
    p = gpu.C.POINTER(gpu.GwfGraph)()
    rc = gpu.host_lib().gwf_graph_build(n_vtx, seq, lens.ctypes.data, len(arcs), src.ctypes.data, dst.ctypes.data,
                                        ovl.ctypes.data, gpu.C.byref(p))
    assert rc == 0
    return gpu.Graph(p), lo, sseq, arcs


def test_vertex_ids_beyond_2_24_vs_oracle(gpu, orc, monkeypatch):
    """A graph of 2^24 + 3000 vertices whose reads start and run on ids on
    both sides of 2^24: the run takes the 28/18/18 split; the default split
    is refused."""
    n_vtx = (1 << 24) + 3000
    _say(f"case: {n_vtx} vertices, reads on ids around 2^24")
    rng = random.Random(2024)
    g, lo, sseq, arcs = _big_graph(gpu, rng, n_vtx, 6000)
    arcs_of = {}
    for a, b, o in arcs:
        arcs_of.setdefault(a, []).append((b, o))
    reads, v0 = [], []
    for _ in range(60):
        s = rng.randrange(len(sseq))
        q = []
        v, p = s, 0
        while len(q) < 400:
            x = rng.random()
            if x < 0.03:
                q.append(rng.choice("ACGT"))
            elif x >= 0.06:
                q.append(chr(sseq[v][p]))
                p += 1
            else:
                p += 1
            if p >= len(sseq[v]):
                v, p = rng.choice(arcs_of[v])
        reads.append("".join(q))
        v0.append(lo + s)
    assert min(v0) < (1 << 24) < max(v0)
    r = gpu.Reads.from_list(reads, v0=v0)
    g.index()
    want, _, _ = orc.batch(g, r)
    for kernel in KERNELS:
        got, _ = _gpu_eds(gpu, g, r, kernel)
        assert np.array_equal(got, want), f"kernel {kernel}: {np.count_nonzero(got != want)} mismatches"
    dg = gpu.DeviceGraph(g)
    try:
        with gpu.Server(dg) as srv:
            ids = srv.submit(r)
            res = srv.collect(len(r), timeout_s=120)
        assert res["id"].tolist() == ids.tolist()
        assert res["ed"].tolist() == want.tolist()
        monkeypatch.setenv("GWF_KEY_LAYOUT", "0")
        dr = gpu.DeviceReads(r)
        try:
            with pytest.raises(gpu.GwfError, match="layout"):
                gpu.ed_run(dg, dr)
        finally:
            dr.free()
    finally:
        dg.free()
    g.free()
