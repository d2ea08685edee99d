"""GFA loader parity: gwf_gfa_load must build the graph the reference's own
GFA layer builds (gfa_read + gfa_sub_from + gwf_gfa2gwf, gfa-io.c,
gfa-base.c, gfa-sub.c, main.cu:294-345), printed as `gwf-test -p` prints it:
same vertices in the same order, same sequences and lengths, same arcs in
the same order.  The expected prints in tests/golden/gfa_ref were made by
the reference's GFA code itself (tests/golden/make_gfa_goldens.py)."""
import glob
import os
import random
import subprocess
import sys

import pytest

from conftest import GOLDEN, ROOT

REF = os.path.join(GOLDEN, "gfa_ref")
CASES = sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(REF, "*.txt")))


def _gfa_path(name):
    p = os.path.join(REF, name + ".gfa")
    return p if os.path.exists(p) else os.path.join(GOLDEN, name + ".gfa")


def _print(gw, path, start, tmp_path):
    g = gw.Graph.from_gfa(path, start=start)
    out = tmp_path / "p.txt"
    g.print_file(str(out))
    return out.read_text()


def test_golden_set_is_complete():
    assert len(CASES) >= 40
    assert any("@" in c for c in CASES)


@pytest.mark.parametrize("case", CASES)
def test_loader_matches_reference_print(gw, case, tmp_path):
    name, _, start = case.partition("@")
    want = open(os.path.join(REF, case + ".txt")).read()
    assert _print(gw, _gfa_path(name), start or None, tmp_path) == want


def test_cli_print_matches_reference(gw):
    exe = os.path.join(gw.LIB_DIR, "gwf-test")
    if not os.path.exists(exe):
        pytest.skip("gwf-test not built")
    for case in ("edges@c", "cyc_strands@s700_18", "graph_1000@seq40"):
        name, _, start = case.partition("@")
        p = subprocess.run([exe, "-p", "-s", start, _gfa_path(name)], capture_output=True, text=True, check=True)
        assert p.stdout == open(os.path.join(REF, case + ".txt")).read()


def test_live_random_graphs_against_reference_build(gw, tmp_path):
    """Where oracle/_ref/gfa_dump (the reference's GFA code, built by
    `make -C oracle ref`) exists, compare 40 fresh random graphs."""
    dump = os.path.join(ROOT, "oracle", "_ref", "gfa_dump")
    if not os.path.exists(dump):
        pytest.skip("reference GFA dumper not built (needs the reference sources)")
    sys.path.insert(0, GOLDEN)
    import make_gfa_goldens as mg
    for k in range(40):
        rng = random.Random(9000 + k)
        text, names = mg.synth_gfa(9000 + k, rng.randint(3, 60), rng.choice([0.0, 0.3, 0.6]),
                                   rng.choice([0.0, 0.5]), rng.random() < 0.7)
        path = tmp_path / f"r{k}.gfa"
        path.write_text(text)
        for st in [None, rng.choice(names)]:
            want = subprocess.run([dump, str(path)] + ([st] if st else []), capture_output=True, text=True,
                                  check=True).stdout
            assert _print(gw, str(path), st, tmp_path) == want, (k, st)
