"""GPU serving (continuous batching) through the C-ABI: every result equals
the oracle's ED, and is the same whether the request ran alone or in a mix
with other traffic admitted while it was on the device."""
import os
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def c3(gpu, orc):
    """The uneven-length workload's graph with 160 of its reads, and the
    oracle's EDs for them."""
    g, r = gpu.make_workload(3, 160)
    g.index()
    want, _, _ = orc.batch(g, r)
    return g, r, want


def _check_results(res, want, ids):
    assert len(res) == len(ids)
    assert np.array_equal(res["id"], np.sort(ids))
    assert (res["status"] == 0).all(), res["status"]
    order = np.argsort(ids)
    got = np.empty(len(ids), np.int32)
    got[order] = res["ed"]
    bad = np.nonzero(got != want)[0]
    assert bad.size == 0, f"{bad.size} mismatches, first {bad[:5]}: got {got[bad[:5]]} want {want[bad[:5]]}"
    assert (res["t_admit_ns"] >= res["t_arrival_ns"]).all()
    assert (res["t_done_ns"] >= res["t_admit_ns"]).all()
    assert (res["steps"] == res["ed"] + 1).all()
    return got


def test_serve_alone_and_in_mix(gpu, c3):
    g, r, want = c3
    dg = gpu.DeviceGraph(g)
    try:
        with gpu.Server(dg) as srv:
            _say("case: serve, 12 requests one at a time (alone)")
            alone_idx = np.arange(0, len(r), len(r) // 12)[:12]
            alone = []
            for i in alone_idx:
                ids = srv.submit(r.subset([i]))
                res = srv.collect(1, timeout_s=60)
                assert res["id"][0] == ids[0] and res["status"][0] == 0
                alone.append(int(res["ed"][0]))
            assert alone == want[alone_idx].tolist()
            _say("case: serve, all requests at once plus a Poisson stream (mix)")
            ids_a = srv.submit(r)
            now = srv.clock_ns()
            rng = np.random.default_rng(7)
            rel = now + np.cumsum(rng.exponential(2e6, len(r))).astype(np.int64)   # ~500 req/s
            ids_b = srv.submit(r, rel)
            res = srv.collect(2 * len(r), timeout_s=120)
            st = srv.stats()
        got_a = _check_results(res[np.isin(res["id"], ids_a)], want, ids_a)
        got_b = _check_results(res[np.isin(res["id"], ids_b)], want, ids_b)
        assert np.array_equal(got_a[alone_idx], np.array(alone))
        assert np.array_equal(got_b, got_a)
        assert st["completed"] == st["submitted"] == 12 + 2 * len(r)
        assert st["inflight_peak"] > 1          # requests overlapped on the device
        assert st["budget_peak"] <= st["budget_bytes"]
        assert st["stalled"] == 0
    finally:
        dg.free()


def test_serve_spill_and_overflow_readmission(gpu, c3, monkeypatch):
    _say("case: serve with tiny LDS and HBM capacities -> spills and re-admissions")
    g, r, want = c3
    monkeypatch.setenv("GWF_DEBUG_CAP_A", "64")
    dg = gpu.DeviceGraph(g)
    try:
        sub = np.arange(40)
        with gpu.Server(dg, lds_states=64, nw=1) as srv:
            ids = srv.submit(r.subset(sub))
            res = srv.collect(len(sub), timeout_s=120)
            st = srv.stats()
        _check_results(res, want[sub], ids)
        assert st["readmitted"] > 0
        assert (res["attempts"] > 1).any()
        assert res["spilled"].any()
    finally:
        dg.free()


def test_serve_budget_admission(gpu, c3):
    _say("case: serve under a small state budget; an oversize request is rejected")
    g, r, want = c3
    dg = gpu.DeviceGraph(g)
    try:
        budget = 24 << 20
        with gpu.Server(dg, state_budget_bytes=budget) as srv:
            ids = srv.submit(r)
            big = gpu.Reads.from_list(["ACGT" * 50000])   # 200 kb: its workspace alone exceeds 24 MiB
            bid = srv.submit(big)
            res = srv.collect(len(r) + 1, timeout_s=120)
            st = srv.stats()
        rb = res[res["id"] == bid[0]]
        assert rb["status"][0] == -8                  # GWF_ERR_BUDGET
        _check_results(res[res["id"] != bid[0]], want, ids)
        assert 0 < st["budget_peak"] <= budget
        assert st["rejected"] == 1
    finally:
        dg.free()


def test_serve_parks_and_relaunches(gpu, c3):
    _say("case: serve parks the device loop when idle and relaunches it")
    g, r, want = c3
    dg = gpu.DeviceGraph(g)
    try:
        with gpu.Server(dg, park_ms=5) as srv:
            ids1 = srv.submit(r.subset(np.arange(10)))
            res1 = srv.collect(10, timeout_s=60)
            time.sleep(0.2)
            ids2 = srv.submit(r.subset(np.arange(10, 20)))
            res2 = srv.collect(10, timeout_s=60)
            st = srv.stats()
        _check_results(res1, want[:10], ids1)
        _check_results(res2, want[10:20], ids2)
        assert st["launches"] >= 2
    finally:
        dg.free()


def test_serve_loop_ends_by_itself(gpu, c3, monkeypatch):
    """Idle workgroups hand their claim tickets back newest first and end; the
    next request relaunches the loop and is claimed at the right sequence."""
    _say("case: serve loop ends after its idle polls (no park) and relaunches")
    monkeypatch.setenv("GWF_SERVE_IDLE_POLLS", "300")
    g, r, want = c3
    dg = gpu.DeviceGraph(g)
    try:
        with gpu.Server(dg, park_ms=600000) as srv:
            ids1 = srv.submit(r.subset(np.arange(40)))
            res1 = srv.collect(40, timeout_s=60)
            time.sleep(3.0)
            ids2 = srv.submit(r.subset(np.arange(40, 60)))
            res2 = srv.collect(20, timeout_s=60)
            time.sleep(3.0)
            ids3 = srv.submit(r.subset(np.arange(60, 61)))
            res3 = srv.collect(1, timeout_s=60)
            st = srv.stats()
        _check_results(res1, want[:40], ids1)
        _check_results(res2, want[40:60], ids2)
        _check_results(res3, want[60:61], ids3)
        assert st["launches"] == 3
    finally:
        dg.free()


def test_serve_end_vertex(gpu, orc):
    """Per-request end vertex (gwf_serve_submit_v1; v1 >= 0 ends at the last
    base of v1, gwf-ed.cu:1231-1254): requests with different end vertices,
    and ones that end anywhere, share one server, and each result equals
    the oracle's for its own v1."""
    _say("case: serve with per-request end vertices, mixed with v1 < 0")
    g, r = gpu.make_workload(1, 40)
    g.index()
    ends = (5, g.n_vtx // 3, -1)
    want = {e: orc.batch(g, r, v1=e)[0] for e in ends}
    dg = gpu.DeviceGraph(g)
    try:
        # default shape, and one wave with the smallest LDS capacity (state spilled to HBM)
        for opts in ({}, dict(nw=1, lds_states=64)):
            with gpu.Server(dg, **opts) as srv:
                ids = {e: srv.submit(r, v1=e) for e in ends}
                res = srv.collect(len(ends) * len(r), timeout_s=120)
            row = {int(i): k for k, i in enumerate(res["id"])}
            for e in ends:
                k = np.array([row[int(i)] for i in ids[e]])
                assert (res["status"][k] == 0).all(), (opts, e, res["status"][k].tolist())
                assert np.array_equal(res["ed"][k], want[e]), (opts, e, res["ed"][k].tolist(), want[e].tolist())
                _say(f"  {opts} v1={e}: {int(res['spilled'][k].sum())} of {len(k)} spilled")
    finally:
        dg.free()
