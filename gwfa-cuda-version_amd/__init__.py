"""Host-side Python mirror of the GWFA interface (ctypes over the C-ABI).

The product is native: libgwfa_hip.so (HIP kernels + the gwfa.h C-ABI) and
libgwfa_host.so (GFA/FASTA readers, synthetic data).  This module only binds
them for tests, benchmarks and scripting:

* ``Graph``      — a ``gwf_graph_t`` (reference layout, gwfa.h:11-18) loaded
                   from GFA (gwf_gfa2gwf, main.cu:294-345) or generated;
                   ``index()`` calls ``gwf_ed_index`` (main.cu:265-270).
* ``Reads``      — packed reads (FASTA via kseq semantics, or synthetic).
* ``DeviceGraph``/``DeviceReads``/``ed_run`` — the GPU path.  There is no CPU
  fallback: if libgwfa_hip.so is missing or no GPU is visible these raise.
* ``CONFIGS``    — the five benchmark workloads (graph + read generators).
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(HERE, "lib")

GWF_OK = 0
GWF_ERR = {
    -2: "GWF_ERR_ARG", -3: "GWF_ERR_HIP", -4: "GWF_ERR_NOMEM", -5: "GWF_ERR_NODEV",
    -6: "GWF_ERR_CAPACITY", -7: "GWF_ERR_STATE", -8: "GWF_ERR_BUDGET",
}
KERNEL_PLAIN = 0
KERNEL_PERSISTENT = 1


class GwfError(RuntimeError):
    pass


class GwfArc(C.Structure):
    _fields_ = [("a", C.c_uint64), ("o", C.c_int32)]


class GwfGraph(C.Structure):
    _fields_ = [("n_vtx", C.c_uint32), ("n_arc", C.c_uint32),
                ("len", C.POINTER(C.c_uint32)), ("src", C.POINTER(C.c_uint32)),
                ("seq", C.POINTER(C.c_char_p)), ("arc", C.POINTER(GwfArc)),
                ("aux", C.POINTER(C.c_uint64))]


class SynthGraphOpts(C.Structure):
    _fields_ = [("backbone_len", C.c_int64), ("seg_min", C.c_int32), ("seg_max", C.c_int32),
                ("bubble_every", C.c_int32), ("p_snp", C.c_double), ("indel_max", C.c_int32),
                ("p_multi", C.c_double), ("p_nested", C.c_double), ("p_tandem", C.c_double),
                ("tr_motif_min", C.c_int32), ("tr_motif_max", C.c_int32),
                ("tr_copy_min", C.c_int32), ("tr_copy_max", C.c_int32), ("seed", C.c_uint64)]


class SynthReadsOpts(C.Structure):
    _fields_ = [("n_reads", C.c_int32), ("len_mode", C.c_int32), ("len_a", C.c_int32),
                ("len_sigma", C.c_double), ("len_min", C.c_int32), ("len_max", C.c_int32),
                ("err_mode", C.c_int32), ("n_err_levels", C.c_int32),
                ("err_levels", C.c_double * 8), ("err_lo", C.c_double), ("err_hi", C.c_double),
                ("start_mode", C.c_int32), ("seed", C.c_uint64)]


class RunStats(C.Structure):
    _fields_ = [("kernel_ms", C.c_double), ("wall_ms", C.c_double), ("n_launch", C.c_int64),
                ("n_retry", C.c_int64), ("sum_steps", C.c_int64), ("sum_states", C.c_int64),
                ("sum_chars", C.c_int64), ("n_spill", C.c_int64), ("sum_hbm_steps", C.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_libs = {}


class TuneBin(C.Structure):
    _fields_ = [("est_max", C.c_int32), ("nw", C.c_int32), ("lds_states", C.c_int32), ("merge_runs", C.c_int32)]


class Tune(C.Structure):
    _fields_ = [("n_bins", C.c_int32), ("bin", TuneBin * 8)]

    def as_list(self):
        return [dict(est_max=b.est_max, nw=b.nw, lds_states=b.lds_states, merge_runs=b.merge_runs)
                for b in self.bin[:self.n_bins]]

    @classmethod
    def from_list(cls, bins):
        t = cls()
        t.n_bins = len(bins)
        for i, b in enumerate(bins):
            t.bin[i] = TuneBin(b["est_max"], b["nw"], b["lds_states"], b["merge_runs"])
        return t


class ServeOpts(C.Structure):
    _fields_ = [("state_budget_bytes", C.c_int64), ("staging_bytes", C.c_int64), ("ring_slots", C.c_int32),
                ("nw", C.c_int32), ("lds_states", C.c_int32), ("wgs", C.c_int32), ("park_ms", C.c_int32),
                ("watchdog_ms", C.c_int32)]


# gwf_serve_result_t as a numpy record (72 bytes, no padding holes)
SERVE_RESULT = np.dtype([("id", "<i8"), ("ed", "<i4"), ("status", "<i4"), ("t_arrival_ns", "<i8"),
                         ("t_admit_ns", "<i8"), ("t_done_ns", "<i8"), ("service_ns", "<i8"), ("steps", "<i8"),
                         ("states", "<i8"), ("attempts", "<i4"), ("spilled", "<i4")])


class ServeStats(C.Structure):
    _fields_ = [(n, C.c_int64) for n in (
        "submitted", "completed", "queued", "inflight", "readmitted", "rejected", "launches", "budget_bytes",
        "budget_used", "budget_peak", "inflight_peak", "stalled", "bases_done")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


def _load(name):
    if name in _libs:
        return _libs[name]
    path = os.path.join(LIB_DIR, name)
    if not os.path.exists(path):
        raise GwfError(f"{path} is not built; run __graft_entry__.build() "
                       f"(or python gwfa-cuda-version_amd/build.py)")
    lib = C.CDLL(path)
    _libs[name] = lib
    return lib


def host_lib():
    lib = _load("libgwfa_host.so")
    if not getattr(lib, "_typed", False):
        P = C.POINTER
        lib.gwf_gfa_load.argtypes = [C.c_char_p, C.c_char_p, P(P(GwfGraph))]
        lib.gwf_free.argtypes = [P(GwfGraph)]
        lib.gwf_free.restype = None
        lib.gwf_graph_build.argtypes = [C.c_uint32, C.c_char_p, C.c_void_p, C.c_uint32, C.c_void_p,
                                        C.c_void_p, C.c_void_p, P(P(GwfGraph))]
        lib.gwf_graph_print_file.argtypes = [P(GwfGraph), C.c_char_p]
        lib.gwf_graph_write_gfa.argtypes = [P(GwfGraph), C.c_char_p]
        lib.gwf_fasta_load.argtypes = [C.c_char_p, P(C.c_int32), P(C.c_void_p), P(C.c_void_p), P(C.c_void_p)]
        lib.gwf_host_free.argtypes = [C.c_void_p]
        lib.gwf_host_free.restype = None
        lib.gwf_synth_graph.argtypes = [P(SynthGraphOpts), P(P(GwfGraph))]
        lib.gwf_synth_reads.argtypes = [P(GwfGraph), P(SynthReadsOpts), P(C.c_int32), P(C.c_void_p),
                                        P(C.c_void_p), P(C.c_void_p), P(C.c_void_p), P(C.c_void_p)]
        lib._typed = True
    return lib


def hip_lib():
    """libgwfa_hip.so; raises GwfError if it has not been built."""
    lib = _load("libgwfa_hip.so")
    if not getattr(lib, "_typed", False):
        P = C.POINTER
        lib.gwf_ed_index.argtypes = [P(GwfGraph)]
        lib.gwf_ed_index.restype = None
        lib.gwf_cleanup.argtypes = [P(GwfGraph)]
        lib.gwf_cleanup.restype = None
        lib.gwf_ed.argtypes = [P(GwfGraph), C.c_int32, C.c_char_p, C.c_int32, C.c_int32, C.c_uint32,
                               C.c_void_p, C.c_void_p]
        lib.gwf_ed.restype = C.c_int32
        lib.gwf_dgraph_upload.argtypes = [P(GwfGraph), P(C.c_void_p)]
        lib.gwf_dgraph_free.argtypes = [C.c_void_p]
        lib.gwf_dgraph_free.restype = None
        lib.gwf_dreads_upload.argtypes = [C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                          P(C.c_void_p)]
        lib.gwf_dreads_free.argtypes = [C.c_void_p]
        lib.gwf_dreads_free.restype = None
        lib.gwf_ed_run_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_uint32, C.c_void_p,
                                           P(RunStats)]
        lib.gwf_dgraph_generation.argtypes = [C.c_void_p]
        lib.gwf_dgraph_generation.restype = C.c_uint64
        lib.gwf_dreads_generation.argtypes = [C.c_void_p]
        lib.gwf_dreads_generation.restype = C.c_uint64
        lib.gwf_ed_run.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_uint32, C.c_int32,
                                   C.c_void_p, P(RunStats)]
        lib.gwf_ed_run_timed.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_uint32, C.c_int32,
                                         C.c_void_p, P(RunStats), C.c_void_p, C.c_void_p]
        lib.gwf_ed_batch.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p,
                                     C.c_int32, C.c_uint32, C.c_int32, C.c_void_p, P(RunStats)]
        lib.gwf_last_error.restype = C.c_char_p
        lib.gwf_serve_clock_ns.restype = C.c_int64
        lib.gwf_tune_get.argtypes = [P(Tune)]
        lib.gwf_tune_set.argtypes = [P(Tune)]
        lib.gwf_tune_save.argtypes = [C.c_char_p]
        lib.gwf_tune_load.argtypes = [C.c_char_p]
        lib.gwf_autotune.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, P(Tune), C.c_char_p, C.c_int64]
        lib.gwf_serve_start.argtypes = [C.c_void_p, P(ServeOpts), P(C.c_void_p)]
        lib.gwf_serve_submit.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p,
                                         C.c_void_p, C.c_void_p, C.c_void_p]
        lib.gwf_serve_submit_v1.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p,
                                            C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        lib.gwf_serve_poll.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32]
        lib.gwf_serve_drain.argtypes = [C.c_void_p, C.c_int32]
        lib.gwf_serve_stats.argtypes = [C.c_void_p, P(ServeStats)]
        lib.gwf_serve_stop.argtypes = [C.c_void_p]
        lib._typed = True
    return lib


def _check(rc, what):
    if rc != GWF_OK:
        msg = ""
        if "libgwfa_hip.so" in _libs:
            msg = _libs["libgwfa_hip.so"].gwf_last_error().decode(errors="replace")
        raise GwfError(f"{what} failed: {GWF_ERR.get(rc, rc)} {msg}")


class Graph:
    """Owns a gwf_graph_t* allocated by libgwfa_host.so."""

    def __init__(self, ptr):
        self.ptr = ptr
        self.indexed = False

    @classmethod
    def from_gfa(cls, path, start=None):
        p = C.POINTER(GwfGraph)()
        rc = host_lib().gwf_gfa_load(os.fsencode(path), start.encode() if start else None, C.byref(p))
        _check(rc, f"gwf_gfa_load({path})")
        return cls(p)

    @classmethod
    def from_lists(cls, seqs, arcs):
        """seqs: list of str/bytes; arcs: list of (src, dst) or (src, dst, overlap)."""
        bs = [s if isinstance(s, bytes) else s.encode() for s in seqs]
        lens = np.array([len(s) for s in bs], dtype=np.uint32)
        src = np.array([a[0] for a in arcs], dtype=np.uint32)
        dst = np.array([a[1] for a in arcs], dtype=np.uint32)
        ovl = np.array([a[2] if len(a) > 2 else 0 for a in arcs], dtype=np.int32)
        p = C.POINTER(GwfGraph)()
        _check(host_lib().gwf_graph_build(len(bs), b"".join(bs), lens.ctypes.data, len(arcs), src.ctypes.data,
                                          dst.ctypes.data, ovl.ctypes.data, C.byref(p)), "gwf_graph_build")
        return cls(p)

    @classmethod
    def synth(cls, **kw):
        o = SynthGraphOpts()
        for k, v in kw.items():
            setattr(o, k, v)
        p = C.POINTER(GwfGraph)()
        _check(host_lib().gwf_synth_graph(C.byref(o), C.byref(p)), "gwf_synth_graph")
        return cls(p)

    def index(self):
        """gwf_ed_index (main.cu:265-270), from libgwfa_hip.so."""
        if not self.indexed:
            hip_lib().gwf_ed_index(self.ptr)
            self.indexed = True
        return self

    @property
    def n_vtx(self):
        return self.ptr.contents.n_vtx

    @property
    def n_arc(self):
        return self.ptr.contents.n_arc

    def total_len(self):
        g = self.ptr.contents
        return int(np.ctypeslib.as_array(g.len, (g.n_vtx,)).sum()) if g.n_vtx else 0

    def lens(self):
        g = self.ptr.contents
        return np.ctypeslib.as_array(g.len, (g.n_vtx,)).copy()

    def seq(self, v):
        return self.ptr.contents.seq[v]

    def arcs(self):
        g = self.ptr.contents
        return [(int(g.arc[i].a >> 32), int(g.arc[i].a & 0xffffffff), int(g.arc[i].o)) for i in range(g.n_arc)]

    def aux(self):
        g = self.ptr.contents
        return np.ctypeslib.as_array(g.aux, (g.n_vtx,)).copy()

    def write_gfa(self, path):
        _check(host_lib().gwf_graph_write_gfa(self.ptr, os.fsencode(path)), "gwf_graph_write_gfa")

    def print_file(self, path):
        _check(host_lib().gwf_graph_print_file(self.ptr, os.fsencode(path)), "gwf_graph_print_file")

    def free(self):
        if self.ptr:
            if self.indexed:
                hip_lib().gwf_cleanup(self.ptr)
            host_lib().gwf_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Reads:
    """Packed reads: ``seqs`` (uint8, NUL after each read), ``off`` int64,
    ``ql`` int32, ``v0`` int32 start vertices, ``err`` float32 (synthetic)."""

    def __init__(self, seqs, off, ql, v0=None, err=None):
        self.seqs = np.ascontiguousarray(seqs, dtype=np.uint8)
        self.off = np.ascontiguousarray(off, dtype=np.int64)
        self.ql = np.ascontiguousarray(ql, dtype=np.int32)
        self.v0 = np.ascontiguousarray(v0 if v0 is not None else np.zeros(len(self.ql)), dtype=np.int32)
        self.err = None if err is None else np.asarray(err, dtype=np.float32)

    def __len__(self):
        return len(self.ql)

    def read(self, i):
        o = int(self.off[i])
        return self.seqs[o:o + int(self.ql[i])].tobytes()

    def subset(self, idx):
        idx = np.asarray(idx)
        parts, offs, o = [], [], 0
        for i in idx:
            r = self.read(int(i))
            parts.append(r + b"\0")
            offs.append(o)
            o += len(r) + 1
        seqs = np.frombuffer(b"".join(parts), dtype=np.uint8) if parts else np.zeros(0, np.uint8)
        return Reads(seqs, offs, self.ql[idx], self.v0[idx], None if self.err is None else self.err[idx])

    @classmethod
    def from_list(cls, seqs, v0=None):
        parts, offs, o = [], [], 0
        for s in seqs:
            s = s if isinstance(s, bytes) else s.encode()
            parts.append(s + b"\0")
            offs.append(o)
            o += len(s) + 1
        buf = np.frombuffer(b"".join(parts), dtype=np.uint8) if parts else np.zeros(1, np.uint8)
        return cls(buf, offs, [len(p) - 1 for p in parts], v0)

    @classmethod
    def _adopt(cls, lib, n, seqs_p, off_p, ql_p, v0_p=None, err_p=None):
        n = n.value
        off = np.ctypeslib.as_array(C.cast(off_p, C.POINTER(C.c_int64)), (max(n, 1),))[:n].copy()
        ql = np.ctypeslib.as_array(C.cast(ql_p, C.POINTER(C.c_int32)), (max(n, 1),))[:n].copy()
        total = int(off[-1] + ql[-1] + 1) if n else 1
        seqs = np.ctypeslib.as_array(C.cast(seqs_p, C.POINTER(C.c_uint8)), (total,)).copy()
        v0 = err = None
        if v0_p is not None:
            v0 = np.ctypeslib.as_array(C.cast(v0_p, C.POINTER(C.c_int32)), (max(n, 1),))[:n].copy()
        if err_p is not None:
            err = np.ctypeslib.as_array(C.cast(err_p, C.POINTER(C.c_float)), (max(n, 1),))[:n].copy()
        for p in (seqs_p, off_p, ql_p, v0_p, err_p):
            if p is not None:
                lib.gwf_host_free(p)
        return cls(seqs, off, ql, v0, err)

    @classmethod
    def from_fasta(cls, path):
        lib = host_lib()
        n = C.c_int32()
        s, o, q = C.c_void_p(), C.c_void_p(), C.c_void_p()
        _check(lib.gwf_fasta_load(os.fsencode(path), C.byref(n), C.byref(s), C.byref(o), C.byref(q)),
               f"gwf_fasta_load({path})")
        return cls._adopt(lib, n, s, o, q)

    @classmethod
    def synth(cls, graph, **kw):
        o = SynthReadsOpts()
        for k, v in kw.items():
            if k == "err_levels":
                o.n_err_levels = len(v)
                for i, x in enumerate(v):
                    o.err_levels[i] = x
            else:
                setattr(o, k, v)
        lib = host_lib()
        n = C.c_int32()
        s, of, q, v0, e = C.c_void_p(), C.c_void_p(), C.c_void_p(), C.c_void_p(), C.c_void_p()
        _check(lib.gwf_synth_reads(graph.ptr, C.byref(o), C.byref(n), C.byref(s), C.byref(of), C.byref(q),
                                   C.byref(v0), C.byref(e)), "gwf_synth_reads")
        return cls._adopt(lib, n, s, of, q, v0, e)

    def bases(self):
        return int(self.ql.astype(np.int64).sum())


# ---------------------------------------------------------------- workloads

def _graph_opts(backbone, bubble_every, p_multi=0.0, p_nested=0.0, p_tandem=0.0, indel_max=10,
                seed=1000):
    return dict(backbone_len=backbone, seg_min=20, seg_max=200, bubble_every=bubble_every, p_snp=0.7,
                indel_max=indel_max, p_multi=p_multi, p_nested=p_nested, p_tandem=p_tandem,
                tr_motif_min=2, tr_motif_max=6, tr_copy_min=5, tr_copy_max=40, seed=seed)


# Seeds: graph 1000+c, reads 2000+c for config c.
CONFIGS = {
    1: dict(name="c1-plain-50kb-1k-x1kb-1pct",
            graph=_graph_opts(50_000, 500, seed=1001),
            reads=dict(n_reads=1000, len_mode=0, len_a=1000, err_mode=0, err_levels=[0.01],
                       start_mode=0, seed=2001)),
    2: dict(name="c2-1Mb-20k-x5kb-err1-5-10",
            graph=_graph_opts(1_000_000, 200, p_multi=0.15, seed=1002),
            reads=dict(n_reads=20_000, len_mode=0, len_a=5000, err_mode=0, err_levels=[0.01, 0.05, 0.10],
                       start_mode=1, seed=2002)),
    3: dict(name="c3-5Mb-100k-lognormal3kb-err1-10",
            graph=_graph_opts(5_000_000, 200, p_multi=0.15, p_nested=0.1, indel_max=50, seed=1003),
            reads=dict(n_reads=100_000, len_mode=1, len_a=3000, len_sigma=0.6, len_min=500, len_max=20_000,
                       err_mode=1, err_lo=0.01, err_hi=0.10, start_mode=1, seed=2003)),
    4: dict(name="c4-2Mb-bubble50-tandem-2k-x10kb-10pct",
            graph=_graph_opts(2_000_000, 50, p_tandem=0.05, seed=1004),
            reads=dict(n_reads=2000, len_mode=0, len_a=10_000, err_mode=0, err_levels=[0.10],
                       start_mode=1, seed=2004)),
}
# Config 5 (serving) reuses config 3's graph and read distribution.
CONFIGS[5] = dict(name="c5-serving-on-c3", graph=CONFIGS[3]["graph"], reads=CONFIGS[3]["reads"])


def make_workload(c, n_reads=None):
    """(Graph, Reads) of benchmark config c (1-5), optionally with fewer reads
    (the first n_reads of the same seeded stream)."""
    cfg = CONFIGS[c]
    g = Graph.synth(**cfg["graph"])
    ro = dict(cfg["reads"])
    if n_reads is not None:
        ro["n_reads"] = n_reads
    r = Reads.synth(g, **ro)
    return g, r


# ---------------------------------------------------------------- GPU path

class DeviceGraph:
    def __init__(self, graph):
        graph.index()
        self.lib = hip_lib()
        self.ptr = C.c_void_p()
        _check(self.lib.gwf_dgraph_upload(graph.ptr, C.byref(self.ptr)), "gwf_dgraph_upload")

    def free(self):
        if self.ptr:
            self.lib.gwf_dgraph_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


    @property
    def generation(self):
        return int(self.lib.gwf_dgraph_generation(self.ptr))


class DeviceReads:
    def __init__(self, reads):
        self.lib = hip_lib()
        self.n = len(reads)
        self.ptr = C.c_void_p()
        _check(self.lib.gwf_dreads_upload(self.n, reads.seqs.ctypes.data, reads.off.ctypes.data,
                                          reads.ql.ctypes.data, reads.v0.ctypes.data, C.byref(self.ptr)),
               "gwf_dreads_upload")

    @property
    def generation(self):
        return int(self.lib.gwf_dreads_generation(self.ptr))

    def free(self):
        if self.ptr:
            self.lib.gwf_dreads_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def ed_batch(dgraph, reads, kernel=KERNEL_PLAIN, v1=-1):
    """gwf_ed_batch: upload `reads` (a Reads) from host pointers, align, free;
    returns (EDs, stats)."""
    n = len(reads)
    ptrs = (C.c_char_p * max(n, 1))()
    base = reads.seqs.ctypes.data
    for i in range(n):
        ptrs[i] = C.cast(base + int(reads.off[i]), C.c_char_p)
    out = np.empty(max(n, 1), dtype=np.int32)
    st = RunStats()
    _check(hip_lib().gwf_ed_batch(dgraph.ptr, n, ptrs, reads.ql.ctypes.data, reads.v0.ctypes.data, v1, 0, kernel,
                                  out.ctypes.data, C.byref(st)), "gwf_ed_batch")
    return out[:n], st.as_dict()


def ed_run(dgraph, dreads, kernel=KERNEL_PLAIN, v1=-1, timed=False):
    """EDs of every read of dreads (np.int32) and the run statistics; with
    timed=True the stats also carry per-read device start/end times (ns)."""
    n = max(dreads.n, 1)
    out = np.empty(n, dtype=np.int32)
    st = RunStats()
    if timed:
        t0, t1 = np.zeros(n, np.int64), np.zeros(n, np.int64)
        rc = hip_lib().gwf_ed_run_timed(dgraph.ptr, dreads.ptr, v1, 0, kernel, out.ctypes.data, C.byref(st),
                                        t0.ctypes.data, t1.ctypes.data)
    else:
        rc = hip_lib().gwf_ed_run(dgraph.ptr, dreads.ptr, v1, 0, kernel, out.ctypes.data, C.byref(st))
    _check(rc, "gwf_ed_run")
    d = st.as_dict()
    if timed:
        d["t_start_ns"], d["t_end_ns"] = t0[:dreads.n], t1[:dreads.n]
    return out[:dreads.n], d


def ed_run_noalloc(dgraph, dreads, v1=-1):
    """gwf_ed_run_noalloc: the plain-kernel run with no device allocation."""
    n = max(dreads.n, 1)
    out = np.empty(n, dtype=np.int32)
    st = RunStats()
    _check(hip_lib().gwf_ed_run_noalloc(dgraph.ptr, dreads.ptr, v1, 0, out.ctypes.data, C.byref(st)),
           "gwf_ed_run_noalloc")
    return out[:dreads.n], st.as_dict()


def gwf_ed(graph, q, v0=0, v1=-1):
    """The reference-signature single-read call (gwfa.h:95) through the C-ABI."""
    graph.index()
    q = q if isinstance(q, bytes) else q.encode()
    r = hip_lib().gwf_ed(graph.ptr, len(q), q, v0, v1, 0, None, None)
    if r < -1:
        _check(r, "gwf_ed")
    return r


class Server:
    """Continuous-batching server (gwf_serve_*): the device graph stays
    resident, requests are admitted by predicted state footprint under a fixed
    device budget, and a resident device loop aligns them as they arrive."""

    def __init__(self, dgraph, **opts):
        self.lib = hip_lib()
        self.dgraph = dgraph
        o = ServeOpts(**opts)
        self.ptr = C.c_void_p()
        _check(self.lib.gwf_serve_start(dgraph.ptr, C.byref(o), C.byref(self.ptr)), "gwf_serve_start")

    @staticmethod
    def clock_ns():
        return hip_lib().gwf_serve_clock_ns()

    def submit(self, reads, release_ns=None, v1=None):
        """Queue every read of `reads` (a Reads); release_ns (int64 array,
        server clock) delays each one's visibility; v1 (int, or int32 array:
        end vertex per read, < 0 = anywhere; None = anywhere).  Returns the ids."""
        n = len(reads)
        ids = np.empty(max(n, 1), np.int64)
        rel = None
        if release_ns is not None:
            rel = np.ascontiguousarray(release_ns, dtype=np.int64)
            assert len(rel) == n
        if v1 is None:
            _check(self.lib.gwf_serve_submit(self.ptr, n, reads.seqs.ctypes.data, reads.off.ctypes.data,
                                             reads.ql.ctypes.data, reads.v0.ctypes.data,
                                             rel.ctypes.data if rel is not None else None, ids.ctypes.data),
                   "gwf_serve_submit")
        else:
            ends = np.ascontiguousarray(np.broadcast_to(np.asarray(v1, dtype=np.int32), (n,)))
            _check(self.lib.gwf_serve_submit_v1(self.ptr, n, reads.seqs.ctypes.data, reads.off.ctypes.data,
                                                reads.ql.ctypes.data, reads.v0.ctypes.data, ends.ctypes.data,
                                                rel.ctypes.data if rel is not None else None, ids.ctypes.data),
                   "gwf_serve_submit_v1")
        return ids[:n]

    def poll(self, max_n=4096, wait_ms=0):
        buf = np.zeros(max_n, SERVE_RESULT)
        k = self.lib.gwf_serve_poll(self.ptr, buf.ctypes.data, max_n, wait_ms)
        if k < 0:
            _check(k, "gwf_serve_poll")
        return buf[:k]

    def drain(self, timeout_ms=-1):
        _check(self.lib.gwf_serve_drain(self.ptr, timeout_ms), "gwf_serve_drain")

    def collect(self, n, timeout_s=600.0):
        """Poll until n results have been returned; results sorted by id."""
        import time
        out, got, t_end = [], 0, time.monotonic() + timeout_s
        while got < n:
            if time.monotonic() > t_end:
                raise GwfError(f"Server.collect: {n - got} of {n} results missing after {timeout_s} s")
            r = self.poll(8192, 50)
            out.append(r)
            got += len(r)
        res = np.concatenate(out) if out else np.zeros(0, SERVE_RESULT)
        return np.sort(res, order="id")

    def stats(self):
        st = ServeStats()
        _check(self.lib.gwf_serve_stats(self.ptr, C.byref(st)), "gwf_serve_stats")
        return st.as_dict()

    def stop(self):
        if self.ptr:
            rc = self.lib.gwf_serve_stop(self.ptr)
            self.ptr = None
            _check(rc, "gwf_serve_stop")

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.stop()

    def __del__(self):
        try:
            self.stop()
        except Exception:
            pass


def tune_get():
    """The current tuning table as a list of bins (dicts)."""
    t = Tune()
    _check(hip_lib().gwf_tune_get(C.byref(t)), "gwf_tune_get")
    return t.as_list()


def tune_set(bins=None):
    """Install a tuning table (list of bin dicts); None restores the built-in one."""
    if bins is None:
        _check(hip_lib().gwf_tune_set(None), "gwf_tune_set")
    else:
        t = Tune.from_list(bins)
        _check(hip_lib().gwf_tune_set(C.byref(t)), "gwf_tune_set")


def tune_save(path):
    _check(hip_lib().gwf_tune_save(path.encode()), "gwf_tune_save")


def tune_load(path):
    _check(hip_lib().gwf_tune_load(path.encode()), "gwf_tune_load")


def autotune(dgraph, dreads, max_evals=0):
    """Run the autotuner on dreads (a workload sample); returns (table, report text)."""
    t = Tune()
    buf = C.create_string_buffer(1 << 16)
    _check(hip_lib().gwf_autotune(dgraph.ptr, dreads.ptr, max_evals, C.byref(t), buf, len(buf)), "gwf_autotune")
    return t.as_list(), buf.value.decode()
