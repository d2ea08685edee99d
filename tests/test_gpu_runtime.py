"""GPU: the batch runtime's caches and budgets.  A captured graph or a cached
device graph must never be reused for a different upload, the persistent
path must keep its workspaces inside the device budget, and the plain path's
multi-stream batches must give the same EDs as one launch.  Every ED is
checked against the oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


ONE_BIN = [dict(est_max=1 << 30, nw=2, lds_states=256, merge_runs=0)]


def _truncated(gw, reads, idx, every, to):
    """Reads idx of `reads`, every `every`-th one cut to `to` bases."""
    seqs = [reads.read(int(i)) for i in idx]
    seqs = [s[:to] if k % every == 0 else s for k, s in enumerate(seqs)]
    return gw.Reads.from_list(seqs, reads.v0[np.asarray(idx)])


@pytest.fixture(scope="module")
def c2(gpu):
    g, r = gpu.make_workload(2, 600)
    g.index()
    return g, r


@pytest.mark.parametrize("table", [None, ONE_BIN], ids=["binned", "direct"])
def test_recycled_read_handles_never_replay(gpu, orc, c2, table):
    """Two read sets with the same count and longest read but different bytes,
    uploaded one after the other into handles that may share an address: each
    run must align its own reads (the replay is keyed by upload generation)."""
    g, r = c2
    a = r.subset(np.arange(0, 300))
    b = _truncated(gpu, r, np.arange(300, 600), 2, 3100)
    assert len(a) == len(b) and a.ql.max() == b.ql.max() and a.bases() != b.bases()
    want_a, _, _ = orc.batch(g, a)
    want_b, _, _ = orc.batch(g, b)
    dg = gpu.DeviceGraph(g)
    try:
        gpu.tune_set(table)
        gens = set()
        for rep in range(2):
            for name, reads, want in (("a", a, want_a), ("b", b, want_b)):
                _say(f"case: recycled handle, table {'one bin' if table else 'built-in'}, reads {name}, pass {rep}")
                dr = gpu.DeviceReads(reads)
                gens.add(dr.generation)
                try:
                    got, _ = gpu.ed_run(dg, dr, kernel=gpu.KERNEL_PERSISTENT)
                finally:
                    dr.free()
                bad = np.nonzero(got != want)[0]
                assert bad.size == 0, f"reads {name} pass {rep}: {bad.size} EDs differ, first {bad[:5]}"
        assert len(gens) == 4 and 0 not in gens
    finally:
        gpu.tune_set(None)
        dg.free()


def test_ed_batch_calls_same_shape(gpu, orc, c2):
    """gwf_ed_batch uploads, runs and frees its reads on every call; two calls
    with the same count and longest read but other bytes must not share a
    captured graph."""
    g, r = c2
    a = r.subset(np.arange(0, 200))
    b = _truncated(gpu, r, np.arange(200, 400), 3, 2500)
    want_a, _, _ = orc.batch(g, a)
    want_b, _, _ = orc.batch(g, b)
    dg = gpu.DeviceGraph(g)
    try:
        for rep in range(2):
            for name, reads, want in (("a", a, want_a), ("b", b, want_b)):
                _say(f"case: gwf_ed_batch persistent, reads {name}, call {rep}")
                got, _ = gpu.ed_batch(dg, reads, kernel=gpu.KERNEL_PERSISTENT)
                assert np.array_equal(got, want), f"reads {name} call {rep}"
    finally:
        dg.free()


def test_device_graph_generations(gpu, c2):
    g, _ = c2
    gens = []
    for _ in range(3):
        dg = gpu.DeviceGraph(g)
        gens.append(dg.generation)
        dg.free()
    assert len(set(gens)) == 3 and min(gens) > 0


def test_gwf_ed_sees_a_reindexed_graph(gpu, orc):
    """gwf_ed caches the device copy per graph pointer; re-indexing the graph
    after its arcs changed must drop that copy."""
    _say("case: gwf_ed, arcs changed and re-indexed at the same graph address")
    lib = gpu.hip_lib()
    g = gpu.Graph.from_lists(["ACGTAC", "TTTTTTTT", "GGGGCCCC"], [(0, 1)])
    g.index()
    q = b"ACGTACGGGGCCCC"
    first = lib.gwf_ed(g.ptr, len(q), q, 0, -1, 0, None, None)
    assert first == orc.ed(g, q) == 8
    g.ptr.contents.arc[0].a = (0 << 32) | 2         # 0 -> 1 becomes 0 -> 2
    lib.gwf_ed_index(g.ptr)
    second = lib.gwf_ed(g.ptr, len(q), q, 0, -1, 0, None, None)
    assert second == orc.ed(g, q) == 0
    assert lib.gwf_ed(g.ptr, len(q), q, 0, 2, 0, None, None) == orc.ed(g, q, v1=2) == 0


def test_long_read_in_short_batch_within_budget(gpu, orc, monkeypatch):
    """One long high-error read among short reads, persistent kernel, with a
    device budget far below the workspaces the longest read would give every
    resident workgroup: the bins' workspaces are sized to the budget and the
    read that outgrows them is re-run by the budgeted plain launches."""
    g, r = gpu.make_workload(3, 2000)
    g.index()
    long_i = int(np.argmax(r.ql))                 # the longest of 2,000 uneven-length reads (~15 kb)
    short = [i for i in range(201) if i != long_i][:200]
    seqs = [r.read(i)[:400] for i in short] + [r.read(long_i)]
    v0 = np.array([int(r.v0[i]) for i in short] + [int(r.v0[long_i])], np.int32)
    reads = gpu.Reads.from_list(seqs, v0)
    assert reads.ql.max() > 8000
    want, _, _ = orc.batch(g, reads)
    monkeypatch.setenv("GWF_BUDGET_BYTES", str(256 << 20))
    dg, dr = gpu.DeviceGraph(g), gpu.DeviceReads(reads)
    try:
        for table in (None, ONE_BIN):
            _say(f"case: long read + 200 short, 256 MiB budget, table {'one bin' if table else 'built-in'}")
            gpu.tune_set(table)
            got, st = gpu.ed_run(dg, dr, kernel=gpu.KERNEL_PERSISTENT)
            bad = np.nonzero(got != want)[0]
            assert bad.size == 0, f"{bad.size} EDs differ, first {bad[:5]}"
            assert st["n_retry"] >= 1      # the long read outgrew the budget-sized workspace
    finally:
        gpu.tune_set(None)
        dr.free()
        dg.free()


def test_plain_batches_on_many_streams(gpu, orc, c2, monkeypatch):
    """A small budget splits the plain launch into many batches spread over
    the streams; the EDs are those of the oracle."""
    _say("case: plain kernel, 48 MiB budget: ~3 MiB arenas, batches of a few reads on 16 streams")
    g, r = c2
    sub = r.subset(np.arange(0, 240))
    want, _, _ = orc.batch(g, sub)
    monkeypatch.setenv("GWF_BUDGET_BYTES", str(48 << 20))
    dg, dr = gpu.DeviceGraph(g), gpu.DeviceReads(sub)
    try:
        got, st = gpu.ed_run(dg, dr, kernel=gpu.KERNEL_PLAIN)
        assert np.array_equal(got, want)
        assert st["n_launch"] >= 16
    finally:
        dr.free()
        dg.free()
