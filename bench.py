#!/usr/bin/env python3
"""Throughput benchmark: aligned query bases/s of GWFA graph edit distance
on one MI355X, with every ED checked bit-exact on the CPU-baseline sample.

A step is one pass of the alignment over the whole read set of the workload
(default: the uneven batch, 100,000 reads with log-normal lengths of median
3 kb clipped to 500 b-20 kb and error rates uniform on 1-10 %, against a
5 Mb-backbone graph with nested bubbles and indel alleles up to 50 bp, every
read starting at offset 0 of a sampled vertex).  Graph and reads are resident
in HBM before the timed region; the timed region is K calls of gwf_ed_run
(launches + the 4 B/read ED copy-back).  Outside it the same reads are aligned
once by the plain one-wavefront-per-read launch (the persistent/plain speed-up)
and every read by the CPU oracle on the host cores (the CPU baseline, and the
parity check of every ED).

--serve measures the serving loop instead (the uneven-length workload's
graph resident, Poisson arrivals at swept fractions of the saturated request
rate, p50/p99 latency and sustained bases/s per rate, every ED checked against
the batch path).

Prints ONE JSON line.  Multi-GPU runs are not part of this build: --gpus N
with N > 1 exits with a message.
"""
import argparse
import importlib.util
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
METRIC = "aligned query bases/sec on 1 MI355X (bit-exact ED); p50/p99 read latency"
HBM_PEAK_GBS = 8000.0


def load_package():
    pdir = os.path.join(ROOT, "gwfa-cuda-version_amd")
    spec = importlib.util.spec_from_file_location("gwfa_amd", os.path.join(pdir, "__init__.py"),
                                                  submodule_search_locations=[pdir])
    m = importlib.util.module_from_spec(spec)
    sys.modules["gwfa_amd"] = m
    spec.loader.exec_module(m)
    return m


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_cores():
    """Cores this process may use: its CPU affinity, limited to the thread
    count the host allots to it (OMP_NUM_THREADS, when set)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    lim = os.environ.get("OMP_NUM_THREADS", "")
    if lim.isdigit() and int(lim) > 0:
        n = min(n, int(lim))
    return max(1, n)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(gw, graph, reads, sample, sample_1t, gpu_eds):
    """Oracle (CPU port of gwf_ed): the first `sample` reads on every host core
    this process may use (also the parity check of those reads' EDs), then the
    first `sample_1t` of them on one thread (0: as many as take ~8 s on one
    thread at the rate the multi-thread run measured per thread)."""
    sys.path.insert(0, ROOT)
    import oracle
    idx = np.arange(min(sample, len(reads)))
    sub = reads.subset(idx)
    threads = host_cores()
    t0 = time.perf_counter()
    eds, st, used = oracle.batch(graph, sub, threads=threads)
    dt = time.perf_counter() - t0
    mism = int((eds != gpu_eds[idx]).sum())
    if sample_1t <= 0:
        per_thread = sub.bases() / dt / max(used, 1)
        cum = np.cumsum(sub.ql.astype(np.int64))
        sample_1t = max(1, int(np.searchsorted(cum, 8.0 * per_thread)))
    sub1 = reads.subset(np.arange(min(sample_1t, len(idx))))
    t1 = time.perf_counter()
    eds1, _, _ = oracle.batch(graph, sub1, threads=1)
    dt1 = time.perf_counter() - t1
    mism += int((eds1 != gpu_eds[:len(sub1)]).sum())
    log(f"cpu baseline: {sub.bases() / dt / 1e6:.2f} M bases/s on {used} threads ({dt:.1f} s), "
        f"{sub1.bases() / dt1 / 1e6:.3f} M bases/s on 1 thread ({dt1:.1f} s)")
    return {
        "value": sub.bases() / dt,
        "unit": "bases/s",
        "cores": int(used),
        "kind": "port",
        "sample": f"first {len(sub)} reads of the workload ({sub.bases()} bases), "
                  f"{dt:.2f} s wall on {used} threads",
        "value_1t": sub1.bases() / dt1,
        "sample_1t": f"first {len(sub1)} reads ({sub1.bases()} bases), {dt1:.2f} s wall on 1 thread",
        "model": cpu_model(),
        "host_cpus": os.cpu_count(),
        "threads_rule": "every core of this process's CPU affinity, limited to OMP_NUM_THREADS when set",
    }, {"reads": len(sub), "mismatches": mism}, st


def serve_bench(gw, args):
    """Serving sweep on the uneven-length workload's graph."""
    cfg = gw.CONFIGS[5]
    n = args.reads or 40_000
    t = time.perf_counter()
    graph, reads = gw.make_workload(5, n)
    graph.index()
    log(f"workload {cfg['name']}: {graph.n_vtx} vertices, {n} reads, {reads.bases()} bases "
        f"({time.perf_counter() - t:.1f} s to generate)")
    dg = gw.DeviceGraph(graph)
    dr = gw.DeviceReads(reads)
    want, _ = gw.ed_run(dg, dr, kernel=gw.KERNEL_PERSISTENT)   # parity reference (oracle-checked path)
    dr.free()
    ql = reads.ql.astype(np.int64)
    fracs = [float(x) for x in args.rates.split(",")]
    out = {"metric": "serving: sustained aligned bases/sec and p50/p99 request latency under Poisson arrivals "
                     "(1 MI355X, bit-exact ED)",
           "unit": "bases/s", "n_gpus": 1, "higher_is_better": True, "vs_baseline": None, "dtype": "u8+int32",
           "data": "synthetic", "scaling": "weak",
           "config": {"workload": cfg["name"], "graph_vertices": graph.n_vtx, "n_reads_pool": n,
                      "graph_seed": cfg["graph"]["seed"], "read_seed": cfg["reads"]["seed"],
                      "arrivals": "Poisson, rate = fraction x saturated reads/s, seed 5"}}
    mism = 0
    shape = {}
    if args.serve_nw:
        shape["nw"] = args.serve_nw
    if args.serve_lds:
        shape["lds_states"] = args.serve_lds
    out["config"]["shape"] = {"nw": args.serve_nw or 4, "lds_states": args.serve_lds or 1024}
    bounds = [0, 2000, 5000, 10000, 1 << 30]
    with gw.Server(dg, state_budget_bytes=args.budget << 20, **shape) as srv:
        # warm-up, then the saturated rate: every request released at once
        srv.submit(reads.subset(np.arange(min(n, 2000))))
        srv.collect(min(n, 2000))
        t0 = srv.clock_ns()
        ids = srv.submit(reads)
        res = srv.collect(n)
        mism += int((res["ed"] != want[ids - ids[0]]).sum())
        sat_s = (int(res["t_done_ns"].max()) - t0) / 1e9
        # the capacity: completions per second while the queue is full (between
        # the 10th and 90th percentile completion times); the makespan rate also
        # holds the first-in-first-out tail of the longest reads
        done = np.sort(res["t_done_ns"].astype(np.int64))
        lo, hi = done[int(0.1 * (n - 1))], done[int(0.9 * (n - 1))]
        sat_rps = float((int(0.9 * (n - 1)) - int(0.1 * (n - 1))) / max((hi - lo) / 1e9, 1e-9))
        bpr = int(ql.sum()) / n
        out["saturated"] = {"reads_per_s": round(sat_rps, 1), "bases_per_s": round(sat_rps * bpr, 1),
                            "definition": "completions per second between the 10th and 90th percentile completion "
                                          "times, every request released at once",
                            "makespan_reads_per_s": round(n / sat_s, 1), "seconds": round(sat_s, 3),
                            "stats": srv.stats()}
        log(f"saturated: {sat_rps:.0f} reads/s ({sat_rps * bpr / 1e6:.1f} M bases/s) while the queue is full, "
            f"{n / sat_s:.0f} reads/s over the makespan")
        rng = np.random.default_rng(5)
        sweep = []
        for f in fracs:
            lam = f * sat_rps
            m = int(min(n, max(2000, lam * args.serve_seconds)))
            idx = rng.choice(n, m, replace=False)
            start = srv.clock_ns() + 300_000_000   # the whole schedule is queued before the first release
            rel = start + np.cumsum(rng.exponential(1e9 / lam, m)).astype(np.int64)
            ids = srv.submit(reads.subset(idx), rel)
            res = srv.collect(m, timeout_s=600)
            mism += int((res["ed"] != want[idx]).sum())   # results come back sorted by id = submission order
            lat = (res["t_done_ns"] - res["t_arrival_ns"]) / 1e6
            span = (int(res["t_done_ns"].max()) - int(res["t_arrival_ns"].min())) / 1e9
            row = {"load": f, "offered_reads_per_s": round(lam, 1), "requests": m,
                   "sustained_bases_per_s": round(float(ql[idx].sum()) / span, 1),
                   "p50_ms": round(float(np.percentile(lat, 50)), 3),
                   "p99_ms": round(float(np.percentile(lat, 99)), 3),
                   "service_p50_ms": round(float(np.percentile(res["service_ns"], 50)) / 1e6, 3),
                   "service_p99_ms": round(float(np.percentile(res["service_ns"], 99)) / 1e6, 3),
                   "queue_p99_ms": round(float(np.percentile((res["t_admit_ns"] - res["t_arrival_ns"]) / 1e6, 99)), 3)}
            qs = ql[idx]
            row["p99_ms_by_length"] = {
                f"{lo}-{hi if hi < (1 << 30) else 'max'}": round(float(np.percentile(lat[(qs >= lo) & (qs < hi)], 99)), 3)
                for lo, hi in zip(bounds[:-1], bounds[1:]) if ((qs >= lo) & (qs < hi)).any()}
            sweep.append(row)
            log(f"load {f}: {row}")
        out["sweep"] = sweep
        out["stats"] = srv.stats()
    dg.free()
    # the headline: the highest offered load served without queueing (admission
    # queue p99 under 1 ms); the loads above it show where the queue builds
    unq = [r for r in sweep if r["queue_p99_ms"] <= 1.0]
    top = max(unq or sweep, key=lambda r: r["load"])
    out["value"] = top["sustained_bases_per_s"]
    out["value_rule"] = "sustained bases/s at the highest offered load whose admission queue p99 is under 1 ms"
    out["latency_ms"] = {"p50": top["p50_ms"], "p99": top["p99_ms"], "at_load": top["load"]}
    out["parity"] = {"requests_checked": int(out["stats"]["completed"]), "mismatches": mism,
                     "against": "gwf_ed_run persistent batch EDs of the same reads"}
    print(json.dumps(out), flush=True)
    if mism:
        raise SystemExit(f"bench.py --serve: {mism} EDs differ")


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, default=3, help="workload 1-4 (default 3, the uneven batch)")
    ap.add_argument("--kernel", choices=["plain", "persistent"], default="persistent")
    ap.add_argument("--reads", type=int, default=None, help="use only the first N reads")
    ap.add_argument("--cpu-sample", type=int, default=None,
                    help="reads timed on the CPU oracle on all host cores and parity-checked (default: every read)")
    ap.add_argument("--cpu-sample-1t", type=int, default=0,
                    help="of those, reads timed again on one thread (default 0: about 8 s of them)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-pcie", action="store_true", help="skip the host-buffers-to-host-EDs (gwf_ed_batch) run")
    ap.add_argument("--no-plain", action="store_true",
                    help="persistent kernel: skip the plain-launch comparison run")
    ap.add_argument("--tune", default=None, help="persistent kernel: load this bin table (no autotuning)")
    ap.add_argument("--save-tune", default=None, help="persistent kernel: write the table used to this file")
    ap.add_argument("--autotune", type=int, default=-1,
                    help="persistent kernel: autotune the bin table on the first N reads before timing, untimed "
                         "(-1: all reads; 0: use the built-in table)")
    ap.add_argument("--serve", action="store_true", help="serving sweep (Poisson arrivals) instead of batch throughput")
    ap.add_argument("--rates", default="0.25,0.5,0.75,0.9,1.0,1.25,1.5",
                    help="--serve: loads as fractions of the saturated rate")
    ap.add_argument("--serve-seconds", type=float, default=1.5, help="--serve: arrival window per load")
    ap.add_argument("--budget", type=int, default=4096, help="--serve: device state budget, MiB")
    ap.add_argument("--serve-nw", type=int, default=0, help="--serve: waves per request (default: the server's, 4)")
    ap.add_argument("--serve-lds", type=int, default=0, help="--serve: LDS states per request (default: the server's, 1024)")
    ap.add_argument("--traffic", default=None,
                    help="JSON with calibrated HBM bytes per step from rocprofv3 --pmc (default: "
                         "profiles/traffic_<kernel>_<workload>.json when present)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build measures one GPU only (reads shard independently; "
              "no multi-GPU path is part of it). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    gw = load_package()
    if args.serve:
        return serve_bench(gw, args)
    kernel = gw.KERNEL_PLAIN if args.kernel == "plain" else gw.KERNEL_PERSISTENT
    cfg = gw.CONFIGS[args.config]
    t = time.perf_counter()
    graph, reads = gw.make_workload(args.config, args.reads)
    graph.index()
    log(f"workload {cfg['name']}: {graph.n_vtx} vertices, {graph.n_arc} arcs, {len(reads)} reads, "
        f"{reads.bases()} bases ({time.perf_counter() - t:.1f} s to generate)")
    dg, dr = gw.DeviceGraph(graph), gw.DeviceReads(reads)
    tuning = None
    if kernel == gw.KERNEL_PERSISTENT and args.tune:
        gw.tune_load(args.tune)
        tuning = gw.tune_get()
        args.autotune = 0
    elif kernel == gw.KERNEL_PERSISTENT and args.autotune != 0:
        t = time.perf_counter()
        if args.autotune < 0 or args.autotune >= len(reads):
            tuning, report = gw.autotune(dg, dr)
        else:
            dsample = gw.DeviceReads(reads.subset(np.arange(args.autotune)))
            tuning, report = gw.autotune(dg, dsample)
            dsample.free()
        log(report.rstrip())
        log(f"autotune: {time.perf_counter() - t:.1f} s (untimed)")
    elif kernel == gw.KERNEL_PERSISTENT:
        tuning = gw.tune_get()

    if args.save_tune and kernel == gw.KERNEL_PERSISTENT:
        gw.tune_save(args.save_tune)

    for i in range(args.warmup):
        eds0, st = gw.ed_run(dg, dr, kernel=kernel)
        log(f"warmup {i}: kernel {st['kernel_ms']:.1f} ms, wall {st['wall_ms']:.1f} ms")
    runs = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        eds, st = gw.ed_run(dg, dr, kernel=kernel, timed=True)
        runs.append((eds, st))
        log(f"step {i}: kernel {st['kernel_ms']:.1f} ms, wall {st['wall_ms']:.1f} ms, launches {st['n_launch']}, "
            f"retries {st['n_retry']}")
    elapsed = time.perf_counter() - t0
    for e, _ in runs[1:]:
        if not np.array_equal(e, runs[0][0]):
            raise SystemExit("bench.py: EDs differ between timed steps")
    eds, st = runs[-1]

    bases = reads.bases()
    value = bases * args.steps / elapsed
    # Algorithmic bytes (per read: ql + 4 + 32 W + 2 C, W = wavefront states
    # entering each score step, C = characters compared by extension).
    alg_bytes = bases + 4 * len(reads) + 32 * st["sum_states"] + 2 * st["sum_chars"]
    kern_s = np.mean([s["kernel_ms"] for _, s in runs]) / 1e3
    achieved = alg_bytes / kern_s / 1e9
    roof = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None,
            "kernel": "gwf_plain_kernel" if args.kernel == "plain" else "gwf_persistent_kernel",
            "kernel_ms_per_step": round(kern_s * 1e3, 3), "launches_per_step": st["n_launch"],
            "alg_bytes_per_step": int(alg_bytes),
            "definition": "achieved = (sum ql + 4 n + 32 W + 2 C) / summed HIP-event time of the step's launches; "
                          "W, C = wavefront states and compared characters counted by the kernel"}
    tpath = args.traffic or os.path.join(ROOT, "profiles", f"traffic_{args.kernel}_{cfg['name']}.json")
    if os.path.exists(tpath) and (args.reads is None):
        with open(tpath) as f:
            tr = json.load(f)
        roof["traffic"] = tr.get("hbm_bytes_per_step")
        roof["traffic_source"] = os.path.relpath(tpath, ROOT)
    lat = st["t_end_ns"] / 1e6
    svc = (st["t_end_ns"] - st["t_start_ns"]) / 1e6
    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "bases/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(elapsed / args.steps * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u8+int32", "data": "synthetic",
        "config": {"workload": cfg["name"], "n_reads": len(reads), "bases": bases, "kernel": args.kernel,
                   "graph_vertices": graph.n_vtx, "graph_seed": cfg["graph"]["seed"],
                   "read_seed": cfg["reads"]["seed"], "parallelism": "reads across workgroups, 1 GPU",
                   "bins": tuning, "autotuned": bool(tuning is not None and args.autotune != 0),
                   "graph_replay": kernel == gw.KERNEL_PERSISTENT and os.environ.get("GWF_GRAPH", "1") != "0"},
        "reads_per_s": round(len(reads) * args.steps / elapsed, 1),
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3), "p99": round(float(np.percentile(lat, 99)), 3),
                       "service_p50": round(float(np.percentile(svc, 50)), 3),
                       "service_p99": round(float(np.percentile(svc, 99)), 3),
                       "definition": "batch mode: p = read's result time since the run's first read started; "
                                     "service = result time - start time of that read"},
        "roofline": roof,
        "gpu_counters": {"states": st["sum_states"], "chars": st["sum_chars"], "steps": st["sum_steps"]},
        "spill": ({"reads_spilled": st["n_spill"], "read_fraction": round(st["n_spill"] / max(len(reads), 1), 4),
                   "hbm_step_fraction": round(st["sum_hbm_steps"] / max(st["sum_steps"], 1), 4),
                   "definition": "reads whose state outgrew the LDS capacity and moved to HBM; share of all "
                                 "score steps run with the state in HBM"}
                  if kernel == gw.KERNEL_PERSISTENT else
                  {"definition": "plain kernel: every read's state is in HBM throughout"}),
    }
    if kernel == gw.KERNEL_PERSISTENT and not args.no_plain:
        # the plain one-wavefront-per-read launch on the same resident reads,
        # outside the timed region: one warm-up run, one measured run
        gw.ed_run(dg, dr, kernel=gw.KERNEL_PLAIN)
        tp = time.perf_counter()
        peds, pst = gw.ed_run(dg, dr, kernel=gw.KERNEL_PLAIN)
        pdt = time.perf_counter() - tp
        if not np.array_equal(peds, eds):
            raise SystemExit("bench.py: plain and persistent EDs differ")
        out["vs_plain"] = {"plain_bases_per_s": round(bases / pdt, 1), "plain_ms": round(pdt * 1e3, 3),
                           "plain_kernel_ms": round(pst["kernel_ms"], 3), "plain_launches": pst["n_launch"],
                           "speedup": round(value / (bases / pdt), 3),
                           "definition": "persistent bases/s over the plain launch's bases/s (wall time of one "
                                         "gwf_ed_run each, same resident reads); plain = one 64-lane wavefront per "
                                         "read, batches of the budget spread over 16 streams"}
        log(f"plain: {pdt * 1e3:.1f} ms ({pst['n_launch']} launches); persistent/plain {out['vs_plain']['speedup']}x")
    if not args.no_pcie:
        # reads in host memory to EDs in host memory (gwf_ed_batch: upload of the
        # read set, the run, copy-back, free), outside the timed region; not `value`
        gw.ed_batch(dg, reads, kernel=kernel)
        tb = time.perf_counter()
        beds, bst = gw.ed_batch(dg, reads, kernel=kernel)
        bdt = time.perf_counter() - tb
        if not np.array_equal(beds, eds):
            raise SystemExit("bench.py: gwf_ed_batch and gwf_ed_run EDs differ")
        out["pcie_inclusive"] = {"bases_per_s": round(bases / bdt, 1), "ms": round(bdt * 1e3, 3),
                                 "kernel_ms": round(bst["kernel_ms"], 3),
                                 "definition": "one gwf_ed_batch call from host read buffers: read-set upload "
                                               "over PCIe, the same run (captured anew: new read handle), "
                                               "EDs copied back, handle freed"}
        log(f"pcie-inclusive: {bdt * 1e3:.1f} ms per call, {bases / bdt / 1e6:.1f} M bases/s")
    if not args.no_cpu:
        sample = len(reads) if args.cpu_sample is None else args.cpu_sample
        cb, par, ost = cpu_baseline(gw, graph, reads, sample, args.cpu_sample_1t, eds)
        out["cpu_baseline"] = cb
        out["parity_sample"] = par
        n = par["reads"]
        out["parity_sample"]["oracle_states"] = ost["states"]
        if par["mismatches"]:
            print(json.dumps(out), flush=True)
            raise SystemExit(f"bench.py: {par['mismatches']} EDs differ from the oracle")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
