"""The gwf-test command line (main.cu:361-750 interface): usage, -p graph
print, -s start segment, and (GPU) the ED listing framed as the reference
frames it."""
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, read_ints


def _cli(gw):
    exe = os.path.join(gw.LIB_DIR, "gwf-test")
    if not os.path.exists(exe):
        pytest.skip("gwf-test not built")
    return exe


def test_usage(gw):
    p = subprocess.run([_cli(gw)], capture_output=True, text=True)
    assert p.returncode == 1 and p.stderr.startswith("Usage: gwf-test")


def test_print_graph_matches_loader(gw, tmp_path):
    gfa = os.path.join(GOLDEN, "grafo_test.gfa")
    p = subprocess.run([_cli(gw), "-p", gfa], capture_output=True, text=True, check=True)
    g = gw.Graph.from_gfa(gfa)
    ref = tmp_path / "g.txt"
    g.print_file(str(ref))
    assert p.stdout == ref.read_text()
    assert p.stdout.splitlines()[0].startswith("S\t0\t")


def test_start_segment_and_unknown_name(gw):
    gfa = os.path.join(GOLDEN, "graph_1000.gfa")
    names = [l.split("\t")[1] for l in open(gfa) if l.startswith("S\t")]
    p = subprocess.run([_cli(gw), "-p", "-s", names[1], gfa], capture_output=True, text=True, check=True)
    g = gw.Graph.from_gfa(gfa, start=names[1])
    s0 = g.seq(0).decode() if isinstance(g.seq(0), bytes) else g.seq(0)
    assert p.stdout.splitlines()[0] == f"S\t0\t{s0}\tLN:i:{len(s0)}"
    q = subprocess.run([_cli(gw), "-p", "-s", "no_such_segment", gfa], capture_output=True, text=True, check=True)
    assert "ERROR: failed to find segment 'no_such_segment'" in q.stderr
    first = subprocess.run([_cli(gw), "-p", gfa], capture_output=True, text=True, check=True)
    assert q.stdout == first.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("kernel", ["persistent", "plain"])
def test_cli_eds(gpu, kernel):
    print(f"case: gwf-test CLI on the 1000-read fixture, kernel {kernel}", flush=True)
    exe = _cli(gpu)
    p = subprocess.run([exe, "-k", kernel, os.path.join(GOLDEN, "graph_1000.gfa"),
                        os.path.join(GOLDEN, "reads_1000.fa.gz")], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    lines = p.stdout.split("\n")
    assert lines[0] == "" and lines[1] == " righe: 1000"
    eds = [int(x) for x in lines[2:1002]]
    assert eds == read_ints(os.path.join(GOLDEN, "reads_1000.ed"))
    assert lines[1002] == "" and lines[1003].startswith("Tempo GPU streams: ") and lines[1003].endswith(" secondi")


@pytest.mark.gpu
def test_cli_gfa_with_varying_overlaps(gpu, orc, tmp_path):
    """A GFA whose links carry different CIGAR overlaps into one segment runs
    through gwf-test (the reference aligns any GFA, gwf-ed.cu:1186-1212) with
    the oracle's EDs."""
    print("case: gwf-test CLI, GFA with varying CIGAR overlaps", flush=True)
    gfa = tmp_path / "ovl.gfa"
    gfa.write_text("H\tVN:Z:1.0\n"
                   "S\ts1\tACGTACGGTA\nS\ts2\tTTGCAGGA\nS\ts3\tGGATCCATGCA\nS\ts4\tCATTGACC\n"
                   "L\ts1\t+\ts2\t+\t0M\nL\ts1\t+\ts3\t+\t2M\nL\ts2\t+\ts3\t+\t4M\n"
                   "L\ts3\t+\ts4\t+\t3M\nL\ts2\t+\ts4\t+\t1M\n")
    reads = ["ACGTACGGTATTGCAGGATCCATGCATTGACC", "ACGTACGGTATCCATGCATTGACC", "ACGTTCGGTATTGCAGGATGCATTGAC",
             "ACGTACGGTATTGCAGGACATTGACC", "ACGACGGTACCATGCATGACC", "GGGG"]
    fa = tmp_path / "r.fa"
    fa.write_text("".join(f">r{i}\n{q}\n" for i, q in enumerate(reads)))
    g = gpu.Graph.from_gfa(str(gfa))
    g.index()
    ins = {}
    for _, w, o in g.arcs():
        ins.setdefault(w, set()).add(o)
    assert any(len(o) > 1 for o in ins.values())          # a segment entered with two different overlaps
    want = [orc.ed(g, q) for q in reads]
    p = subprocess.run([_cli(gpu), str(gfa), str(fa)], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    lines = p.stdout.split("\n")
    assert lines[1] == f" righe: {len(reads)}"
    assert [int(x) for x in lines[2:2 + len(reads)]] == want
