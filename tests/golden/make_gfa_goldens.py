"""Write the GFA loader's golden fixtures (run here, where the reference
sources are; the fixtures travel, the reference does not).

For a set of seeded synthetic GFA files (both strands, CIGAR overlaps,
cycles, unreachable segments, duplicate links, `*` sequences with LN tags)
plus the two fixture graphs, each with the default and some named start
segments, the reference's own GFA layer (oracle/_ref/gfa_dump, built by
`make -C oracle ref` from gfa-io.c / gfa-base.c / gfa-sub.c) prints the graph
gwf_gfa2gwf builds.  Output: gfa_ref/<name>.gfa and gfa_ref/<name>[@<start>].txt.
"""
import os
import random
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "gfa_ref")
DUMP = os.path.join(ROOT, "oracle", "_ref", "gfa_dump")


def synth_gfa(seed, n, p_minus, p_ovl, cyclic, p_star=0.0, lower=False):
    rng = random.Random(seed)
    names = [f"s{rng.randrange(1000):03d}_{i}" for i in range(n)]
    seqs = []
    for _ in range(n):
        L = rng.randint(1, 40)
        s = "".join(rng.choice("ACGT") for _ in range(L))
        seqs.append(s.lower() if lower and rng.random() < 0.3 else s)
    lines = ["H\tVN:Z:1.0"]
    order = list(range(n))
    rng.shuffle(order)   # S lines in shuffled order, some L lines before their S lines
    links = []
    for v in range(n):
        for _ in range(rng.randint(0, 3)):
            w = rng.randrange(n) if cyclic else rng.randint(v + 1, n - 1) if v + 1 < n else None
            if w is None:
                continue
            ov = 0
            if rng.random() < p_ovl:
                ov = rng.randint(0, max(0, min(len(seqs[v]), len(seqs[w])) - 1))
            o1 = "-" if rng.random() < p_minus else "+"
            o2 = "-" if rng.random() < p_minus else "+"
            links.append(f"L\t{names[v]}\t{o1}\t{names[w]}\t{o2}\t{ov}M")
            if rng.random() < 0.1:
                links.append(links[-1])   # duplicate
    for i, v in enumerate(order):
        if p_star and rng.random() < p_star:
            lines.append(f"S\t{names[v]}\t*\tLN:i:{len(seqs[v])}")
        else:
            lines.append(f"S\t{names[v]}\t{seqs[v]}")
        if i == n // 2:
            lines.extend(links[: len(links) // 2])
    lines.extend(links[len(links) // 2:])
    return "\n".join(lines) + "\n", names


def edge_gfa():
    """Hand-written lines for the parser's corner cases: ov:ow and :ow
    overlaps (one side only: filled from the complement link when exactly
    one matches, else dropped), '*' overlaps, a missing overlap field,
    CIGARs with I/D, an invalid orientation, links to segments never
    defined, a self loop, a long overlap.  (A segment given as '*' is left
    out: the reference's conversion dereferences its null sequence.)"""
    segs = {"a": "ACGTACGTAC", "b": "GGCATTACAG", "c": "TTAGGC", "d": "CATCATCATCAT", "e": "A",
            "f": "GATTACAGATTACA", "g": "CCGGTTAA"}
    lines = ["H\tVN:Z:1.0"]
    lines += [f"S\t{k}\t{v}" for k, v in segs.items()]
    lines += ["L\ta\t+\tb\t+\t3:3", "L\tb\t-\ta\t-\t3:3",        # explicit pair: complement found
              "L\tb\t+\tc\t+\t:2", "L\tc\t-\tb\t-\t2:4",          # semi overlap filled from complement
              "L\tc\t+\td\t+\t:1",                                    # semi overlap, no complement: dropped
              "L\td\t+\te\t+\t*", "L\te\t+\tf\t+",                # '*' and missing overlap field
              "L\tf\t+\tg\t+\t2M1I1M", "L\tg\t+\ta\t-\t1M2D",     # CIGARs with I and D
              "L\ta\t+\tc\tx\t0M",                                    # invalid orientation: ignored
              "L\td\t+\tghost\t+\t0M",                                # never defined: dropped
              "L\tg\t+\tg\t+\t0M",                                    # self loop
              "L\tb\t+\td\t-\t20M"]                                   # overlap longer than b: clamped
    return "\n".join(lines) + "\n"


def fasta_gfa():
    """The reference reader also takes FASTA records as segments (named s1,
    s2, ... in record order, the header name is not a segment name), with
    GFA L lines after them."""
    return (">x1 first\nACGTAC\nGGT\n>x2\nTTTTGCA\n>x3\nGATTACA\nL\ts1\t+\ts2\t+\t0M\n"
            "L\ts2\t+\ts3\t-\t2M\nL\ts3\t+\ts1\t-\t1M\nL\tx1\t+\ts2\t+\t0M\n")


def main():
    if not os.path.exists(DUMP):
        sys.exit(f"{DUMP} missing: run `make -C oracle ref` first (needs the reference sources)")
    os.makedirs(OUT, exist_ok=True)
    cases = []
    specs = [("dag_plus", 11, 30, 0.0, 0.0, False), ("dag_strands", 12, 30, 0.4, 0.3, False),
             ("cyc_plus", 13, 25, 0.0, 0.2, True), ("cyc_strands", 14, 25, 0.5, 0.3, True),
             ("cyc_big", 15, 60, 0.3, 0.3, True), ("dag_lower", 16, 20, 0.2, 0.0, False)]
    for name, seed, n, pm, po, cyc in specs:
        text, names = synth_gfa(seed, n, pm, po, cyc, lower=name.endswith("lower"))
        path = os.path.join(OUT, name + ".gfa")
        with open(path, "w") as f:
            f.write(text)
        rng = random.Random(seed + 100)
        cases.append((name, path, [None] + rng.sample(names, 2)))
    for k in range(12):   # more seeds, mixed features
        name = f"rnd{k:02d}"
        seed = 500 + k
        rng0 = random.Random(seed)
        text, names = synth_gfa(seed, rng0.randint(8, 45), rng0.choice([0.0, 0.2, 0.5]), rng0.choice([0.0, 0.4]),
                                rng0.random() < 0.6, lower=k % 4 == 0)
        path = os.path.join(OUT, name + ".gfa")
        with open(path, "w") as f:
            f.write(text)
        cases.append((name, path, [None] + random.Random(seed + 7).sample(names, 2)))
    for name, text, starts in (("edges", edge_gfa(), [None, "b", "c", "f"]), ("fasta", fasta_gfa(), [None, "s2", "x1"])):
        path = os.path.join(OUT, name + ".gfa")
        with open(path, "w") as f:
            f.write(text)
        cases.append((name, path, starts))
    g1000 = os.path.join(HERE, "graph_1000.gfa")
    gtest = os.path.join(HERE, "grafo_test.gfa")
    cases.append(("graph_1000", g1000, [None, "seq1", "seq40"]))
    cases.append(("grafo_test", gtest, [None, "100", "400"]))
    for name, path, starts in cases:
        for st in starts:
            cmd = [DUMP, path] + ([st] if st else [])
            out = subprocess.run(cmd, capture_output=True, text=True, check=True).stdout
            tag = name + (f"@{st}" if st else "")
            with open(os.path.join(OUT, tag + ".txt"), "w") as f:
                f.write(out)
            print(tag, out.count("\nS") + out.startswith("S"), "vertices", out.count("\nL"), "arcs")


if __name__ == "__main__":
    main()
