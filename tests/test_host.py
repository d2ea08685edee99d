"""Host-side pieces: GFA/FASTA readers, the synthetic generator and the
C-ABI exports (no GPU needed)."""
import gzip
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT


def test_gfa_fixture_1000(gw):
    g = gw.Graph.from_gfa(os.path.join(GOLDEN, "graph_1000.gfa"))
    assert g.n_vtx == 80 and g.n_arc == 99
    assert g.total_len() == 1071
    assert g.seq(0).startswith(b"AATCTGTTCGAGGGCTAGC")     # vertex 0 = first segment (seq1)
    assert all(u < w for u, w, _ in g.arcs())             # topological order
    assert all(o == 0 for _, _, o in g.arcs())            # '*' overlaps read as 0


def test_gfa_fixture_grafo(gw):
    g = gw.Graph.from_gfa(os.path.join(GOLDEN, "grafo_test.gfa"))
    assert g.n_vtx == 820 and g.n_arc == 1108             # P/W lines ignored
    assert all(u < w for u, w, _ in g.arcs())


def test_gfa_strands_overlaps_and_start(gw, tmp_path):
    p = tmp_path / "g.gfa"
    p.write_text("H\tVN:Z:1.0\n"
                 "S\ta\tACGTT\n"
                 "S\tb\tGGCAT\n"
                 "S\tc\t*\tLN:i:3\n"
                 "L\ta\t+\tb\t-\t2M\n"
                 "L\tb\t-\tc\t+\t0M\n")
    g = gw.Graph.from_gfa(str(p))
    assert g.n_vtx == 3
    assert g.seq(0) == b"ACGTT" and g.seq(1) == b"ATGCC"  # b on the - strand is reverse-complemented
    assert g.arcs() == [(0, 1, 2), (1, 2, 0)]
    g2 = gw.Graph.from_gfa(str(p), start="b")
    assert g2.seq(0) == b"GGCAT" and g2.n_vtx == 2       # forward b reaches a- only (complement arcs)
    with pytest.raises(gw.GwfError):
        gw.Graph.from_gfa(str(p), start="nope")


def test_fasta_reader(gw, tmp_path):
    r = gw.Reads.from_fasta(os.path.join(GOLDEN, "reads_1000.fa.gz"))
    assert len(r) == 1000
    assert sorted(set(r.ql.tolist()))[0] == 342
    with gzip.open(os.path.join(GOLDEN, "reads_1000.fa.gz"), "rt") as f:
        seqs = [l.strip() for l in f if not l.startswith(">")]
    assert [r.read(i).decode() for i in range(5)] == seqs[:5]
    p = tmp_path / "x.fq"
    p.write_bytes(b">m1 comment\r\nACGT\r\nAC\r\n\n>m2\nnNac\n@q1\nACGTA\n+\n@@@@I\n@q2\nGG\n+q2\nII\n")
    r2 = gw.Reads.from_fasta(str(p))
    assert [r2.read(i) for i in range(len(r2))] == [b"ACGTAC", b"nNac", b"ACGTA", b"GG"]


def test_synth_deterministic_and_topological(gw):
    o = gw.CONFIGS[1]["graph"]
    g1, g2 = gw.Graph.synth(**o), gw.Graph.synth(**o)
    assert g1.n_vtx == g2.n_vtx and g1.arcs() == g2.arcs()
    assert all(u < w for u, w, _ in g1.arcs())
    assert abs(g1.total_len() - o["backbone_len"]) < 0.1 * o["backbone_len"]
    r1 = gw.Reads.synth(g1, **gw.CONFIGS[1]["reads"])
    r2 = gw.Reads.synth(g2, **gw.CONFIGS[1]["reads"])
    assert np.array_equal(r1.seqs, r2.seqs) and len(r1) == 1000 and set(r1.ql.tolist()) == {1000}


def test_synth_error_free_reads_align_exactly(gw, orc):
    g = gw.Graph.synth(**gw.CONFIGS[4]["graph"] | {"backbone_len": 30_000})
    g.index()
    r = gw.Reads.synth(g, n_reads=40, len_mode=0, len_a=800, err_mode=0, err_levels=[0.0], start_mode=1, seed=5)
    eds, _, _ = orc.batch(g, r)
    assert eds.tolist() == [0] * 40


def _declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return set(re.findall(r"\b(gwf_\w+)\s*\(", txt))


def _exported(lib):
    out = subprocess.run(["nm", "-D", "--defined-only", lib], capture_output=True, text=True, check=True).stdout
    return {l.split()[-1] for l in out.splitlines() if " T " in l}


def test_hip_library_exports_header(gw):
    lib = os.path.join(gw.LIB_DIR, "libgwfa_hip.so")
    missing = _declared("gwfa.h") - _exported(lib)
    assert not missing, missing
    gw.hip_lib()   # loads without a GPU


def test_host_library_exports_header(gw):
    lib = os.path.join(gw.LIB_DIR, "libgwfa_host.so")
    missing = _declared("gwfa_host.h") - _exported(lib)
    assert not missing, missing


def test_reference_type_layout(tmp_path):
    """Types keep the reference's sizes (gwfa.h:6-80) under gcc."""
    src = tmp_path / "l.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "gwfa.h"\n'
                   'int main(void){printf("%zu %zu %zu %zu %zu %zu %zu %zu\\n", sizeof(gwf_arc_t), '
                   'sizeof(gwf_graph_t), sizeof(gwf_diag_t), sizeof(gwf_intv_t), sizeof(gwf_edbuf_t), '
                   'sizeof(counters), offsetof(gwf_graph_t, aux), offsetof(gwf_edbuf_t, ha_mask));return 0;}\n')
    exe = tmp_path / "l"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    assert out == ["16", "48", "16", "16", "192", "28", "40", "184"]


def test_host_entry_points_fail_cleanly_without_device(gw):
    from conftest import gpu_present
    if gpu_present():
        pytest.skip("device present")
    g = gw.Graph.from_lists(["ACGT"], [])
    g.index()
    with pytest.raises(gw.GwfError, match="NODEV"):
        gw.DeviceGraph(g)
    assert gw.gwf_ed.__doc__
    lib = gw.hip_lib()
    assert lib.gwf_ed(g.ptr, 4, b"ACGT", 0, -1, 0, None, None) == -5   # GWF_ERR_NODEV


def test_serving_type_layout(tmp_path, gw):
    """The serving structs have the sizes the ctypes/numpy mirrors assume."""
    import ctypes
    src = tmp_path / "s.c"
    src.write_text('#include <stdio.h>\n#include "gwfa.h"\n'
                   'int main(void){printf("%zu %zu %zu\\n", sizeof(gwf_serve_opts_t), sizeof(gwf_serve_result_t), '
                   'sizeof(gwf_serve_stats_t));return 0;}\n')
    exe = tmp_path / "s"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    assert out == [ctypes.sizeof(gw.ServeOpts), gw.SERVE_RESULT.itemsize, ctypes.sizeof(gw.ServeStats)]


def test_serve_start_fails_cleanly_without_device(gw):
    from conftest import gpu_present
    if gpu_present():
        pytest.skip("device present")
    import ctypes
    lib = gw.hip_lib()
    srv = ctypes.c_void_p()
    assert lib.gwf_serve_start(ctypes.c_void_p(1), None, ctypes.byref(srv)) == -5   # GWF_ERR_NODEV
    assert lib.gwf_serve_start(None, None, ctypes.byref(srv)) == -2                 # GWF_ERR_ARG
    assert lib.gwf_serve_stop(None) == 0


def test_tuning_table_roundtrip(gw, tmp_path):
    """Tuning tables: get/set/save/load through the C-ABI (no device needed)."""
    base = gw.tune_get()
    assert 1 <= len(base) <= 8
    try:
        mine = [dict(est_max=100, nw=1, lds_states=128, merge_runs=0),
                dict(est_max=1000, nw=2, lds_states=512, merge_runs=16),
                dict(est_max=1 << 30, nw=4, lds_states=2048, merge_runs=32)]
        gw.tune_set(mine)
        assert gw.tune_get() == mine
        f = str(tmp_path / "t.txt")
        gw.tune_save(f)
        gw.tune_set(None)
        assert gw.tune_get() == base
        gw.tune_load(f)
        assert gw.tune_get() == mine
        for bad in ([dict(est_max=10, nw=3, lds_states=256, merge_runs=8)],              # nw
                    [dict(est_max=10, nw=1, lds_states=1 << 20, merge_runs=8)],          # LDS too large
                    [dict(est_max=50, nw=1, lds_states=256, merge_runs=8),
                     dict(est_max=40, nw=1, lds_states=256, merge_runs=8)]):             # not increasing
            with pytest.raises(gw.GwfError):
                gw.tune_set(bad)
        (tmp_path / "bad.txt").write_text("bin 10 1 256\n")
        with pytest.raises(gw.GwfError):
            gw.tune_load(str(tmp_path / "bad.txt"))
        assert gw.tune_get() == mine
    finally:
        gw.tune_set(None)
