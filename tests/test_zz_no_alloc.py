"""GPU: gwf_ed_run_noalloc (every device buffer in one static device array)
against gwf_ed_run on the same handles, bit-exact, on two calls in a row (the
second smaller), with forced workspace overflows, and its size limit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def c3(gpu):
    g, r = gpu.make_workload(3, 700)
    g.index()
    return g, r


def test_noalloc_matches_ed_run_twice(gpu, orc, c3):
    g, r = c3
    dg = gpu.DeviceGraph(g)
    try:
        for name, idx in (("first call, 700 reads", np.arange(700)), ("second call, 90 reads", np.arange(300, 390))):
            _say(f"case: gwf_ed_run_noalloc {name}")
            sub = r.subset(idx)
            dr = gpu.DeviceReads(sub)
            try:
                want, _ = gpu.ed_run(dg, dr, kernel=gpu.KERNEL_PLAIN)
                got, st = gpu.ed_run_noalloc(dg, dr)
                assert np.array_equal(got, want), f"{name}: {int((got != want).sum())} EDs differ"
                pers, _ = gpu.ed_run(dg, dr, kernel=gpu.KERNEL_PERSISTENT)
                assert np.array_equal(got, pers)
                assert st["n_launch"] >= 1
            finally:
                dr.free()
        oracle_eds, _, _ = orc.batch(g, r.subset(np.arange(300, 390)))
        assert np.array_equal(got, oracle_eds)
    finally:
        dg.free()


def test_noalloc_reruns_overflowed_reads(gpu, c3, monkeypatch):
    _say("case: gwf_ed_run_noalloc with 64-state workspaces: most reads overflow and are re-run")
    g, r = c3
    sub = r.subset(np.arange(200))
    dg, dr = gpu.DeviceGraph(g), gpu.DeviceReads(sub)
    try:
        want, _ = gpu.ed_run(dg, dr, kernel=gpu.KERNEL_PLAIN)
        monkeypatch.setenv("GWF_DEBUG_CAP_A", "64")
        got, st = gpu.ed_run_noalloc(dg, dr)
        assert np.array_equal(got, want)
        assert st["n_retry"] > 0
        monkeypatch.delenv("GWF_DEBUG_CAP_A")
        got2, st2 = gpu.ed_run_noalloc(dg, dr)     # the array still holds the last call's state
        assert np.array_equal(got2, want)
    finally:
        dr.free()
        dg.free()


def test_noalloc_rejects_too_many_reads(gpu):
    _say("case: gwf_ed_run_noalloc with more reads than its limit")
    g = gpu.Graph.from_lists(["ACGT"], [])
    dg = gpu.DeviceGraph(g)
    n = (1 << 20) + 1
    reads = gpu.Reads(np.zeros(n, np.uint8), np.arange(n, dtype=np.int64), np.zeros(n, np.int32))
    dr = gpu.DeviceReads(reads)
    try:
        with pytest.raises(gpu.GwfError, match="GWF_ERR_CAPACITY"):
            gpu.ed_run_noalloc(dg, dr)
    finally:
        dr.free()
        dg.free()
