"""GPU: the device-side binned pipeline, its graph replays and the autotuner.
Every ED must equal the oracle's whatever the bin configuration."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def mix(gpu, orc):
    g, r = gpu.make_workload(3, 400)
    g.index()
    want, _, _ = orc.batch(g, r)
    return g, r, want


def test_graph_replays_match_eager(gpu, mix, monkeypatch):
    _say("case: binned persistent run, captured graph replayed 4x with eager work in between")
    g, r, want = mix
    dg, dr = gpu.DeviceGraph(g), gpu.DeviceReads(r)
    try:
        monkeypatch.setenv("GWF_GRAPH", "0")
        eager, _ = gpu.ed_run(dg, dr, kernel=gpu.KERNEL_PERSISTENT)
        assert np.array_equal(eager, want)
        monkeypatch.setenv("GWF_GRAPH", "1")
        for k in range(4):
            got, st = gpu.ed_run(dg, dr, kernel=gpu.KERNEL_PERSISTENT)
            assert np.array_equal(got, eager), f"replay {k}"
            assert st["sum_steps"] == int(want.sum()) + len(want)
            plain, _ = gpu.ed_run(dg, dr, kernel=gpu.KERNEL_PLAIN)      # eager work between replays
            assert np.array_equal(plain, eager)
    finally:
        dr.free()
        dg.free()


def test_bins_on_one_stream(gpu, mix, monkeypatch):
    _say("case: bins launched one after another on one stream")
    g, r, want = mix
    monkeypatch.setenv("GWF_BIN_STREAMS", "0")
    dg, dr = gpu.DeviceGraph(g), gpu.DeviceReads(r)
    try:
        got, _ = gpu.ed_run(dg, dr, kernel=gpu.KERNEL_PERSISTENT)
        assert np.array_equal(got, want)
    finally:
        dr.free()
        dg.free()


@pytest.mark.parametrize("table", [
    [dict(est_max=1 << 30, nw=1, lds_states=128, merge_runs=0)],
    [dict(est_max=200, nw=1, lds_states=256, merge_runs=32), dict(est_max=1 << 30, nw=2, lds_states=512, merge_runs=2)],
    [dict(est_max=100, nw=4, lds_states=2048, merge_runs=16), dict(est_max=400, nw=2, lds_states=1024, merge_runs=0),
     dict(est_max=900, nw=1, lds_states=512, merge_runs=8), dict(est_max=1 << 30, nw=4, lds_states=1024, merge_runs=4)],
])
def test_bin_tables(gpu, mix, table):
    _say(f"case: tuning table {table}")
    g, r, want = mix
    dg, dr = gpu.DeviceGraph(g), gpu.DeviceReads(r)
    try:
        gpu.tune_set(table)
        got, _ = gpu.ed_run(dg, dr, kernel=gpu.KERNEL_PERSISTENT)
        assert np.array_equal(got, want)
    finally:
        gpu.tune_set(None)
        dr.free()
        dg.free()


def test_autotune(gpu, mix):
    _say("case: autotune on the uneven-length sample, then align with the tuned table")
    g, r, want = mix
    dg, dr = gpu.DeviceGraph(g), gpu.DeviceReads(r)
    try:
        table, report = gpu.autotune(dg, dr, max_evals=12)
        assert "chose" in report and report.count("eval") == 12
        assert gpu.tune_get() == table
        got, _ = gpu.ed_run(dg, dr, kernel=gpu.KERNEL_PERSISTENT)
        assert np.array_equal(got, want)
    finally:
        gpu.tune_set(None)
        dr.free()
        dg.free()


def test_heartbeat_progress_lines(gpu, capfd, monkeypatch):
    """GWF_HEARTBEAT_MS: a batch run prints progress lines from the
    workgroups' records while it runs (main.cu:48-149's progress polling)."""
    _say("case: heartbeat lines during a binned run")
    g, r = gpu.make_workload(4, 300)
    g.index()
    dg, dr = gpu.DeviceGraph(g), gpu.DeviceReads(r)
    try:
        monkeypatch.setenv("GWF_HEARTBEAT_MS", "2")
        got, _ = gpu.ed_run(dg, dr, kernel=gpu.KERNEL_PERSISTENT)
        monkeypatch.delenv("GWF_HEARTBEAT_MS")
        want, _ = gpu.ed_run(dg, dr, kernel=gpu.KERNEL_PERSISTENT)
    finally:
        dr.free()
        dg.free()
    err = capfd.readouterr().err
    lines = [l for l in err.splitlines() if l.startswith("[gwf heartbeat]")]
    assert len(lines) >= 2, err[-2000:]
    assert "busy=" in lines[0] and "max_step=" in lines[0]
    assert np.array_equal(got, want)


@pytest.mark.parametrize("direct", ["1", "0"])
def test_single_bin_direct_or_probed(gpu, mix, monkeypatch, direct):
    """One bin and no more reads than resident workgroups: the run launches
    the bin kernel on the reads in input order (GWF_DIRECT=1: one launch) or
    probes and bins them first (GWF_DIRECT=0: two launches, the probe and the
    bin); both give the oracle's EDs, eagerly and from the captured graph."""
    _say(f"case: single-bin run of {len(mix[1])} reads, GWF_DIRECT={direct}")
    monkeypatch.setenv("GWF_DIRECT", direct)
    g, r, want = mix
    dg, dr = gpu.DeviceGraph(g), gpu.DeviceReads(r)
    try:
        gpu.tune_set([dict(est_max=1 << 30, nw=2, lds_states=128, merge_runs=0)])
        for rep in range(3):  # capture, then replays
            got, st = gpu.ed_run(dg, dr, kernel=gpu.KERNEL_PERSISTENT)
            assert np.array_equal(got, want), rep
            assert st["n_launch"] == (1 if direct == "1" else 2), st["n_launch"]
    finally:
        gpu.tune_set(None)
        dr.free()
        dg.free()
