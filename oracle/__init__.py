"""TEST INFRASTRUCTURE ONLY: ctypes binding of liboracle.so (the CPU
restatement of the reference's gwf_ed, gwf-ed.cu:80-1337, and the
brute-force DP checker).  Importable only by tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg.  Graphs are passed as the gwf_graph_t
pointers owned by the package's ``Graph`` objects."""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SO = os.path.join(HERE, "liboracle.so")


class OracleStats(C.Structure):
    _fields_ = [("steps", C.c_int64), ("states", C.c_int64), ("chars", C.c_int64),
                ("max_na", C.c_int64), ("a_events", C.c_int64), ("crossings", C.c_int64),
                ("vis_lossy", C.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


def build():
    subprocess.run(["make", "-C", HERE, "-s", "liboracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(SO):
            build()
        _lib = C.CDLL(SO)
        _lib.oracle_gwf_ed.argtypes = [C.c_void_p, C.c_int32, C.c_char_p, C.c_int32, C.c_int32, C.c_uint32,
                                       C.POINTER(OracleStats)]
        _lib.oracle_gwf_ed.restype = C.c_int32
        _lib.oracle_dp_ed.argtypes = [C.c_void_p, C.c_int32, C.c_char_p, C.c_int32, C.c_int32]
        _lib.oracle_dp_ed.restype = C.c_int32
        _lib.oracle_ed_index.argtypes = [C.c_void_p, C.c_void_p]
        _lib.oracle_ed_index.restype = None
        _lib.oracle_batch.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.c_int32, C.c_void_p, C.c_int, C.POINTER(OracleStats)]
        _lib.oracle_batch.restype = C.c_int
    return _lib


def _gptr(graph):
    return C.cast(graph.ptr, C.c_void_p)


def ed(graph, q, v0=0, v1=-1, stats=None):
    q = q if isinstance(q, bytes) else q.encode()
    st = OracleStats()
    r = lib().oracle_gwf_ed(_gptr(graph), len(q), q, v0, v1, 0, C.byref(st))
    if stats is not None:
        stats.update(st.as_dict())
    return r


def dp(graph, q, v0=0, v1=-1):
    q = q if isinstance(q, bytes) else q.encode()
    return lib().oracle_dp_ed(_gptr(graph), len(q), q, v0, v1)


def batch(graph, reads, v1=-1, threads=0):
    """(EDs, stats, threads used) for every read of a Reads object."""
    out = np.empty(max(len(reads), 1), dtype=np.int32)
    st = OracleStats()
    t = lib().oracle_batch(_gptr(graph), len(reads), reads.seqs.ctypes.data, reads.off.ctypes.data,
                           reads.ql.ctypes.data, reads.v0.ctypes.data, v1, out.ctypes.data, threads,
                           C.byref(st))
    return out[:len(reads)], st.as_dict(), t
