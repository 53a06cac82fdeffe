"""ctypes wrapper of oracle/liboracle.so (the CPU restatement and the OpenMP
baseline).  Test infrastructure only."""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "liboracle.so")
REF_CONVERTER = os.path.join(ORACLE_DIR, "_ref", "converter")
INF = 4294967294

if not os.path.exists(ORACLE_LIB):
    subprocess.run(["make", "-C", ORACLE_DIR, "liboracle.so"], check=True, capture_output=True)

_lib = ctypes.CDLL(ORACLE_LIB)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_f32p = ctypes.POINTER(ctypes.c_float)
_f64p = ctypes.POINTER(ctypes.c_double)
_u64 = ctypes.c_uint64

_lib.oracle_bfs.argtypes = [_u64, _u64p, _u32p, ctypes.c_uint32, _u32p]
_lib.oracle_sssp.argtypes = [_u64, _u64p, _u32p, _u32p, ctypes.c_uint32, _u32p]
_lib.oracle_sswp.argtypes = [_u64, _u64p, _u32p, _u32p, ctypes.c_uint32, _u32p]
_lib.oracle_cc.argtypes = [_u64, _u64p, _u32p, _u32p]
_lib.oracle_pr.argtypes = [_u64, _u64p, _u32p, ctypes.c_double, ctypes.c_double, ctypes.c_double, _f64p, _f64p]
_lib.oracle_pr.restype = _u64
_lib.oracle_generate.argtypes = [_u64, _u8p, _u32p, _u32p, _u64p]
_lib.oracle_generate.restype = _u64
_lib.oracle_partition.argtypes = [_u64, _u64p, _u64, _u64, _u64p, _u64p, _u64p, _u64p]
_lib.oracle_partition.restype = _u64
_lib.omp_label_run.argtypes = [ctypes.c_int, _u64, _u64p, _u32p, _u32p, ctypes.c_uint32, _u32p, _u64p]
_lib.omp_label_run.restype = _u64
_lib.omp_label_run_w8.argtypes = [ctypes.c_int, _u64, _u64p, _u32p, _u8p, ctypes.c_uint32, _u32p, _u64p]
_lib.omp_label_run_w8.restype = _u64
_lib.omp_pr_run.argtypes = [_u64, _u64p, _u32p, ctypes.c_float, ctypes.c_float, _f32p, _f32p, _u64p]
_lib.omp_pr_run.restype = _u64
_lib.omp_num_threads.restype = ctypes.c_int


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def _p(a, t):
    return a.ctypes.data_as(t) if a is not None else None


def bfs(ptr, adj, src):
    ptr, adj = _c(ptr, np.uint64), _c(adj, np.uint32)
    n = len(ptr) - 1
    out = np.empty(n, np.uint32)
    _lib.oracle_bfs(n, _p(ptr, _u64p), _p(adj, _u32p), src, _p(out, _u32p))
    return out


def sssp(ptr, adj, w, src):
    ptr, adj, w = _c(ptr, np.uint64), _c(adj, np.uint32), _c(w, np.uint32)
    n = len(ptr) - 1
    out = np.empty(n, np.uint32)
    _lib.oracle_sssp(n, _p(ptr, _u64p), _p(adj, _u32p), _p(w, _u32p), src, _p(out, _u32p))
    return out


def sswp(ptr, adj, w, src):
    ptr, adj, w = _c(ptr, np.uint64), _c(adj, np.uint32), _c(w, np.uint32)
    n = len(ptr) - 1
    out = np.empty(n, np.uint32)
    _lib.oracle_sswp(n, _p(ptr, _u64p), _p(adj, _u32p), _p(w, _u32p), src, _p(out, _u32p))
    return out


def cc(ptr, adj):
    ptr, adj = _c(ptr, np.uint64), _c(adj, np.uint32)
    n = len(ptr) - 1
    out = np.empty(n, np.uint32)
    _lib.oracle_cc(n, _p(ptr, _u64p), _p(adj, _u32p), _p(out, _u32p))
    return out


def pr(ptr, adj, acc, init=0.15, damping=0.85):
    ptr, adj = _c(ptr, np.uint64), _c(adj, np.uint32)
    n = len(ptr) - 1
    value = np.empty(n, np.float64)
    delta = np.empty(n, np.float64)
    _lib.oracle_pr(n, _p(ptr, _u64p), _p(adj, _u32p), init, damping, acc, _p(value, _f64p), _p(delta, _f64p))
    return value, delta


def generate(active, degree):
    active, degree = _c(active, np.uint8), _c(degree, np.uint32)
    n = len(degree)
    an = np.empty(n, np.uint32)
    ap = np.empty(n + 1, np.uint64)
    a = _lib.oracle_generate(n, _p(active, _u8p), _p(degree, _u32p), _p(an, _u32p), _p(ap, _u64p))
    return an[:a].copy(), ap[: a + 1].copy()


def partition(active_nodes_pointer, budget):
    ap = _c(active_nodes_pointer, np.uint64)
    a = len(ap) - 1
    cap = a + 1
    cols = [np.zeros(cap, np.uint64) for _ in range(4)]
    np_ = _lib.oracle_partition(a, _p(ap, _u64p), budget, cap, *[_p(c, _u64p) for c in cols])
    return np.stack([c[:np_] for c in cols], axis=1)


ALGO = {"bfs": 0, "cc": 1, "sssp": 2, "sswp": 3}


def omp_label(algo, ptr, adj, w, src):
    """OpenMP baseline of one label algorithm.  w: uint32 weights, uint8
    weights (a graph held as split arrays) or None (BFS and CC read none)."""
    ptr, adj = _c(ptr, np.uint64), _c(adj, np.uint32)
    n = len(ptr) - 1
    out = np.empty(n, np.uint32)
    edges = ctypes.c_uint64()
    if w is None and algo in ("sssp", "sswp"):
        raise ValueError(f"{algo} needs weights")
    if w is not None and w.dtype == np.uint8:
        w8 = _c(w, np.uint8)
        rounds = _lib.omp_label_run_w8(ALGO[algo], n, _p(ptr, _u64p), _p(adj, _u32p), _p(w8, _u8p), src,
                                       _p(out, _u32p), ctypes.byref(edges))
    else:
        w32 = None if w is None else _c(w, np.uint32)
        rounds = _lib.omp_label_run(ALGO[algo], n, _p(ptr, _u64p), _p(adj, _u32p), _p(w32, _u32p), src,
                                    _p(out, _u32p), ctypes.byref(edges))
    return out, int(rounds), int(edges.value)


def omp_pr(ptr, adj, acc, init=0.15):
    ptr, adj = _c(ptr, np.uint64), _c(adj, np.uint32)
    n = len(ptr) - 1
    value = np.empty(n, np.float32)
    delta = np.empty(n, np.float32)
    edges = ctypes.c_uint64()
    rounds = _lib.omp_pr_run(n, _p(ptr, _u64p), _p(adj, _u32p), init, acc, _p(value, _f32p), _p(delta, _f32p),
                             ctypes.byref(edges))
    return value, int(rounds), int(edges.value)


def omp_threads():
    return int(_lib.omp_num_threads())


def expected(algo, ptr, adj, w, src, acc=1e-6):
    """Oracle answer for one application."""
    if algo == "bfs":
        return bfs(ptr, adj, src)
    if algo == "sssp":
        return sssp(ptr, adj, w, src)
    if algo == "sswp":
        return sswp(ptr, adj, w, src)
    if algo == "cc":
        return cc(ptr, adj)
    if algo == "pr":
        return pr(ptr, adj, acc)[0]
    raise ValueError(algo)
