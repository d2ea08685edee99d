"""Regenerate the golden fixtures of tests/golden (run in the build container).

Inputs are the reference's own fixture files (exposed_testing/, data only):
    graph_1000.gfa, reads_1000.fa.gz  <- exposed_testing/1000_reads/
    grafo_test.gfa                    <- exposed_testing/generated/
The reference ships no expected outputs, so the expected EDs are computed
here by the oracle (the CPU restatement of gwf_ed) and every value is
cross-checked against the independent brute-force DP; any disagreement
aborts.  Outputs:
    reads_1000.ed        one ED per read of reads_1000.fa.gz (v0 = 0)
    grafo_reads.fa       200 reads sampled along random paths of grafo_test.gfa
    grafo_reads.ed       their EDs (v0 = 0)
"""
import gzip
import importlib.util
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)


def load_pkg():
    spec = importlib.util.spec_from_file_location(
        "gwfa_amd", os.path.join(ROOT, "gwfa-cuda-version_amd", "__init__.py"),
        submodule_search_locations=[os.path.join(ROOT, "gwfa-cuda-version_amd")])
    m = importlib.util.module_from_spec(spec)
    sys.modules["gwfa_amd"] = m
    spec.loader.exec_module(m)
    return m


def main():
    gw = load_pkg()
    import oracle
    # 1000_reads fixture
    g = gw.Graph.from_gfa(os.path.join(HERE, "graph_1000.gfa")).index()
    r = gw.Reads.from_fasta(os.path.join(HERE, "reads_1000.fa.gz"))
    eds, _, _ = oracle.batch(g, r)
    for i in range(len(r)):
        d = oracle.dp(g, r.read(i))
        if d != eds[i]:
            raise SystemExit(f"reads_1000 read {i}: oracle {eds[i]} != dp {d}")
    with open(os.path.join(HERE, "reads_1000.ed"), "w") as f:
        f.write("".join(f"{int(x)}\n" for x in eds))
    # grafo_test fixture with sampled reads
    g2 = gw.Graph.from_gfa(os.path.join(HERE, "grafo_test.gfa")).index()
    r2 = gw.Reads.synth(g2, n_reads=200, len_mode=1, len_a=600, len_sigma=0.5, len_min=50, len_max=3000,
                        err_mode=1, err_lo=0.0, err_hi=0.12, start_mode=0, seed=4242)
    with open(os.path.join(HERE, "grafo_reads.fa"), "w") as f:
        for i in range(len(r2)):
            f.write(f">r{i} err={float(r2.err[i]):.4f}\n{r2.read(i).decode()}\n")
    eds2, _, _ = oracle.batch(g2, r2)
    for i in range(len(r2)):
        d = oracle.dp(g2, r2.read(i))
        if d != eds2[i]:
            raise SystemExit(f"grafo read {i}: oracle {eds2[i]} != dp {d}")
    with open(os.path.join(HERE, "grafo_reads.ed"), "w") as f:
        f.write("".join(f"{int(x)}\n" for x in eds2))
    print("reads_1000: n", len(r), "sum ED", int(eds.sum()), "| grafo: n", len(r2), "sum ED", int(eds2.sum()))


if __name__ == "__main__":
    main()
