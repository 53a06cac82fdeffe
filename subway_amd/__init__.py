"""subway_amd -- Python host mirror of the MI355X-native Subway pipeline.

Everything here calls the C ABI of ``lib/libsubway_amd.so`` (declared in
``include/subway_amd.h``) through ctypes; there is no Python or CPU fallback
for any GPU stage.  Importing works without a GPU (graph construction,
format conversion and R-MAT generation are host code); running a stage
without a GPU raises :class:`SubwayError`.

The entry points mirror the reference's applications
(``subway/{bfs,cc,sssp,sswp,pr}-{sync,async}.cu``): :func:`run` takes the
algorithm, the mode and the reference's ``--source`` meaning, plus the
device edge budget that drives partitioning.
"""
from __future__ import annotations

import atexit
import ctypes
import os
import weakref
from dataclasses import dataclass, field

import numpy as np

from ._lib import lib, LIB_PATH, SubwayError, check

__all__ = [
    "Graph", "RunStats", "Session", "BatchSession", "run", "generate", "generate_pr", "partition", "gather",
    "device_count", "DIST_INFINITY", "ALGOS", "LIB_PATH", "SubwayError", "APPS", "h2d_peak", "probe_copy",
]

DIST_INFINITY = 4294967294  # shared/globals.hpp:27
ALGOS = {"bfs": 0, "cc": 1, "sssp": 2, "sswp": 3, "pr": 4}
MODES = {"sync": 0, "async": 1}
TRANSFERS = {"auto": 0, "zerocopy": 1, "host": 2}
POLICIES = {"plain": 0, "learned": 1}
# the ten applications of the reference (subway/*.cu)
APPS = [f"{a}-{m}" for a in ("bfs", "cc", "sssp", "sswp", "pr") for m in ("sync", "async")]

_u8p = ctypes.POINTER(ctypes.c_uint8)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_f32p = ctypes.POINTER(ctypes.c_float)


class _Options(ctypes.Structure):
    _fields_ = [
        ("algo", ctypes.c_int32), ("mode", ctypes.c_int32), ("source", ctypes.c_uint32),
        ("num_buffers", ctypes.c_int32), ("budget_bytes", ctypes.c_uint64), ("acc", ctypes.c_float),
        ("init_pr", ctypes.c_float), ("transfer", ctypes.c_int32), ("policy", ctypes.c_int32),
        ("verbose", ctypes.c_int32), ("advance_grid", ctypes.c_int32), ("profile", ctypes.c_int32),
        ("rounds_mode", ctypes.c_int32), ("rounds_grid", ctypes.c_int32), ("max_rounds", ctypes.c_int32),
        ("batch", ctypes.c_int32), ("batch_unroll", ctypes.c_int32), ("theta", ctypes.c_float),
        ("spin_limit", ctypes.c_uint32), ("unpacked_edges", ctypes.c_int32), ("join_window", ctypes.c_uint32),
        ("no_resident", ctypes.c_int32), ("resident_min_bytes", ctypes.c_uint32), ("push_only", ctypes.c_int32),
        ("no_blocks", ctypes.c_int32), ("early_readback", ctypes.c_int32),
    ]


class _Stats(ctypes.Structure):
    _fields_ = [
        ("ms_total", ctypes.c_double), ("ms_transfer", ctypes.c_double), ("ms_advance", ctypes.c_double),
        ("ms_frontier", ctypes.c_double), ("ms_generate", ctypes.c_double), ("ms_rounds", ctypes.c_double),
        ("h2d_bytes", ctypes.c_uint64),
        ("h2d_dma_bytes", ctypes.c_uint64), ("h2d_gather_bytes", ctypes.c_uint64),
        ("inner_rounds", ctypes.c_uint64), ("partitions", ctypes.c_uint64), ("updates", ctypes.c_uint64),
        ("advance_launches", ctypes.c_uint64), ("frontier_launches", ctypes.c_uint64),
        ("transfer_ops", ctypes.c_uint64), ("rounds_launches", ctypes.c_uint64), ("work_edges", ctypes.c_uint64),
        ("work_entries", ctypes.c_uint64),
        ("global_iterations", ctypes.c_uint32), ("trace_len", ctypes.c_uint32),
        ("work_items", ctypes.c_uint64), ("mispredictions", ctypes.c_uint32), ("fallback_iteration", ctypes.c_uint32),
        ("edge_bytes", ctypes.c_uint32), ("min_members", ctypes.c_uint32),
        ("lane_work", ctypes.c_uint64), ("lane_entries", ctypes.c_uint64),
        ("pair_work_edges", ctypes.c_uint64), ("pair_work_entries", ctypes.c_uint64), ("pair_updates", ctypes.c_uint64),
        ("resident_visits", ctypes.c_uint32), ("resident_overlaps", ctypes.c_uint32),
        ("pull_rounds", ctypes.c_uint32), ("blocked_units", ctypes.c_uint32), ("pull_found", ctypes.c_uint64),
        ("h2d_blocked_bytes", ctypes.c_uint64), ("blocked_edges", ctypes.c_uint64),
        ("ms_rounds_blocked", ctypes.c_double), ("rounds_launches_blocked", ctypes.c_uint64),
        ("blk_work_edges", ctypes.c_uint64), ("blk_work_entries", ctypes.c_uint64), ("blk_updates", ctypes.c_uint64),
        ("early_readback", ctypes.c_uint64), ("values_patched", ctypes.c_uint64),
    ]


class _IterRecord(ctypes.Structure):
    _fields_ = [
        ("iteration", ctypes.c_uint32), ("partitions", ctypes.c_uint32), ("active_nodes", ctypes.c_uint64),
        ("active_edges", ctypes.c_uint64), ("h2d_bytes", ctypes.c_uint64), ("ms_iteration", ctypes.c_double),
        ("features", ctypes.c_float * 12), ("heads", ctypes.c_float * 3), ("theta", ctypes.c_float),
        ("rounds_cap", ctypes.c_uint32), ("on_plain", ctypes.c_uint32),
    ]


def _check_layouts():
    """The ctypes mirrors above must match the library's structs byte for
    byte; a stale library or binding fails here instead of corrupting runs."""
    got = [ctypes.c_uint32() for _ in range(3)]
    lib.subway_struct_sizes(*(ctypes.cast(ctypes.byref(x), _u32p) for x in got))
    want = [ctypes.sizeof(_Options), ctypes.sizeof(_Stats), ctypes.sizeof(_IterRecord)]
    if [x.value for x in got] != want:
        raise ImportError(f"{LIB_PATH}: struct sizes (options, stats, iter_record) {[x.value for x in got]} "
                          f"do not match this binding's {want}; rebuild the library")


_check_layouts()


def _ptr(a, t):
    return a.ctypes.data_as(t) if a is not None else None


# Every live graph and session, so that the exit hook can free them while the
# HIP runtime is still whole (interpreter teardown would otherwise run their
# finalizers in arbitrary order, possibly after other modules are gone).
_live_sessions = weakref.WeakSet()
_live_graphs = weakref.WeakSet()


@atexit.register
def _shutdown():
    """Close every session, then free every graph, then release the
    library's shared streams (subway_shutdown).  Runs at interpreter exit,
    before any finalizer of module teardown; the objects' own close()/__del__
    are no-ops afterwards."""
    for s in list(_live_sessions):
        s.close()
    for g in list(_live_graphs):
        g.free()
    lib.subway_shutdown()


def device_count() -> int:
    """Number of visible HIP devices (0 when there is no GPU)."""
    return int(lib.subway_device_count())


class Graph:
    """Host CSR in pinned, device-mapped memory (Graph<E> / GraphPR<E>).

    ``weighted`` graphs hold OutEdgeWeighted {end, w8} edges (.bwcsr),
    unweighted ones OutEdge {end} (.bcsr).
    """

    def __init__(self, handle: ctypes.c_void_p):
        self._h = handle
        n, m = ctypes.c_uint64(), ctypes.c_uint64()
        w, md, mv = ctypes.c_int32(), ctypes.c_uint32(), ctypes.c_uint32()
        check(lib.subway_graph_info(self._h, ctypes.byref(n), ctypes.byref(m), ctypes.byref(w),
                                    ctypes.byref(md), ctypes.byref(mv)))
        self.n, self.m, self.weighted = int(n.value), int(m.value), bool(w.value)
        self.max_degree, self.max_degree_vertex = int(md.value), int(mv.value)
        _live_graphs.add(self)

    def free(self):
        """Releases the host CSR (sessions built on it keep their own reference)."""
        h = getattr(self, "_h", None)
        if h:
            self._h = None
            lib.subway_graph_free(h)

    def __del__(self):
        self.free()

    # --- constructors
    @classmethod
    def load(cls, path: str, weighted: bool | None = None, split_layout: bool = False) -> "Graph":
        """Read .bcsr/.bwcsr (and 64-bit variants) or .el/.wel text.
        split_layout: a .bwcsr graph that moves only in the split format (as
        weighted lists of 16 GiB or more that do not pack are read by
        themselves); weights must fit a byte."""
        if weighted is None:
            weighted = os.path.splitext(path)[1] in (".bwcsr", ".bwcsr64", ".wel")
        h = ctypes.c_void_p()
        w = (2 if split_layout else 1) if weighted else 0
        check(lib.subway_load(path.encode(), w, ctypes.byref(h)))
        return cls(h)

    @classmethod
    def from_csr(cls, node_pointer, ends, weights=None) -> "Graph":
        ptr = np.ascontiguousarray(node_pointer, dtype=np.uint64)
        e = np.ascontiguousarray(ends, dtype=np.uint32)
        w = None if weights is None else np.ascontiguousarray(weights, dtype=np.uint32)
        n = len(ptr) - 1
        if n < 1:
            raise SubwayError(-1, "a graph needs at least one vertex")
        h = ctypes.c_void_p()
        check(lib.subway_graph_from_csr(n, len(e), _ptr(ptr, _u64p), _ptr(e, _u32p), _ptr(w, _u32p),
                                        ctypes.byref(h)))
        return cls(h)

    @classmethod
    def from_edges(cls, n: int, src, dst, weights=None) -> "Graph":
        """CSR by a stable sort on the source (tools/converter.cpp:64-83)."""
        src = np.asarray(src, dtype=np.uint32)
        dst = np.asarray(dst, dtype=np.uint32)
        order = np.argsort(src, kind="stable")
        deg = np.bincount(src, minlength=n).astype(np.uint64)
        ptr = np.zeros(n + 1, dtype=np.uint64)
        np.cumsum(deg, out=ptr[1:])
        w = None if weights is None else np.asarray(weights, dtype=np.uint32)[order]
        return cls.from_csr(ptr, dst[order], w)

    @classmethod
    def rmat(cls, scale: int, edge_factor: int = 16, a: float = 0.57, b: float = 0.19, c: float = 0.19,
             seed: int = 42, weighted: bool = False, symmetrize: bool = False, wmin: int = 1,
             wmax: int = 64, split_layout: bool = False, build: str = "auto") -> "Graph":
        """split_layout: lay a weighted graph out as one that moves only in the
        split format (pinned destination and weight arrays), as large weighted
        graphs that do not pack are by themselves.

        build: "host" (threads of this process), "device" (the GPU builds the
        same CSR, byte for byte, and copies it to pinned host memory; a
        split-layout graph then keeps no 8-byte list) or "auto" (the device
        for 2^24 generated edges or more when a GPU is visible and has the
        memory, else the host)."""
        if build not in ("auto", "host", "device"):
            raise ValueError(f"build must be auto, host or device, not {build!r}")
        h = ctypes.c_void_p()
        w = (2 if split_layout else 1) if weighted else 0
        args = (scale, edge_factor, a, b, c, seed, w, int(symmetrize), wmin, wmax, ctypes.byref(h))
        use_device = build == "device" or (
            build == "auto" and (edge_factor << scale) >= (1 << 24) and wmax <= 255 and device_count() > 0)
        if use_device:
            rc = lib.subway_graph_rmat_device(*args)
            if rc == 0:
                return cls(h)
            if build == "device" or rc != -5:      # only a lack of device memory falls back to the host
                check(rc)
        check(lib.subway_graph_rmat(*args))
        return cls(h)

    def reorder(self, order: str = "in-degree") -> None:
        """Renumbers the vertices in descending in-degree order (ties by
        ascending id) on the GPU, in place (subway_graph_reorder), so that
        the most-targeted vertices' values share cache lines.  Runs, sessions
        and batches keep taking sources and returning values in the original
        numbering; csr(), split_view(), node_pointer(), out_degree() and
        save() then describe the renumbered graph, and order() gives the
        maps.  Call it before creating any session on the graph."""
        if order != "in-degree":
            raise ValueError(f"order must be 'in-degree', not {order!r}")
        check(lib.subway_graph_reorder(self._h, 1))

    def build_blocks(self, unpin_split: bool = False) -> None:
        """Sorts every row by destination and builds the block-compressed
        copy of the rows on the GPU, in place (subway_graph_build_blocks).
        Sessions then move whole id ranges in that layout unless created
        with blocks=False; every result is unchanged.  csr(), split_view()
        and save() then see the sorted rows.  unpin_split: move the split
        arrays to ordinary memory (done by itself when the words cannot be
        pinned beside them); compacted units are then gathered by host
        threads.  Call it before creating any session on the graph."""
        check(lib.subway_graph_build_blocks(self._h, 1 if unpin_split else 0))

    def split_tail(self) -> int:
        """First edge of the pinned copy of the low-degree id suffix's split
        edges (kept when the split arrays left pinned memory), m if none."""
        t = ctypes.c_uint64()
        check(lib.subway_graph_split_tail(self._h, ctypes.byref(t)))
        return t.value

    def blocks_view(self):
        """(row_ptr u64[n+1], words u32[num_words], weight_bits, weight_base)
        of the block layout, read-only zero-copy views that keep the graph
        alive, or None when the graph has no block layout."""
        rp, w, nw, code = _u64p(), _u32p(), ctypes.c_uint64(), ctypes.c_uint32()
        rc = lib.subway_graph_blocks_view(self._h, ctypes.byref(rp), ctypes.byref(w), ctypes.byref(nw),
                                          ctypes.byref(code))
        if rc != 0:
            return None
        out = []
        for ptr, count, ct, dt in ((rp, self.n + 1, ctypes.c_uint64, np.uint64),
                                   (w, nw.value, ctypes.c_uint32, np.uint32)):
            buf = (ct * max(count, 1)).from_address(ctypes.addressof(ptr.contents))
            buf._graph = self
            a = np.frombuffer(buf, dtype=dt)[:count]
            a.flags.writeable = False
            out.append(a)
        return out[0], out[1], code.value & 63, code.value >> 6

    def order(self):
        """(perm, iperm): read-only u32[n] views with perm[v] the internal id
        of original vertex v and iperm[p] the original id of internal vertex
        p, or None for a graph that keeps its own numbering.  Each view keeps
        the graph alive."""
        p, ip = _u32p(), _u32p()
        check(lib.subway_graph_order(self._h, ctypes.byref(p), ctypes.byref(ip)))
        if not p:
            return None
        out = []
        for ptr in (p, ip):
            buf = (ctypes.c_uint32 * self.n).from_address(ctypes.addressof(ptr.contents))
            buf._graph = self
            a = np.frombuffer(buf, dtype=np.uint32)
            a.flags.writeable = False
            out.append(a)
        return tuple(out)

    @property
    def reordered(self) -> bool:
        return self.order() is not None

    # --- accessors
    def save(self, path: str) -> None:
        check(lib.subway_graph_save(self._h, path.encode()))

    def csr(self):
        """(node_pointer[n+1] u64, ends[m] u32, weights[m] u32 or None)."""
        ptr = np.empty(self.n + 1, dtype=np.uint64)
        ends = np.empty(self.m, dtype=np.uint32)
        w = np.empty(self.m, dtype=np.uint32) if self.weighted else None
        check(lib.subway_graph_export(self._h, _ptr(ptr, _u64p), _ptr(ends, _u32p), _ptr(w, _u32p)))
        return ptr, ends, w

    def split_view(self):
        """Zero-copy read-only views (node_pointer u64[n+1], ends u32[m],
        weights u8[m]) of a split-layout graph's host arrays; each view keeps
        the graph alive."""
        p, e, w = _u64p(), _u32p(), _u8p()
        check(lib.subway_graph_split_view(self._h, ctypes.byref(p), ctypes.byref(e), ctypes.byref(w)))
        out = []
        for ptr, count, ct, dt in ((p, self.n + 1, ctypes.c_uint64, np.uint64), (e, self.m, ctypes.c_uint32, np.uint32),
                                   (w, self.m, ctypes.c_uint8, np.uint8)):
            buf = (ct * max(count, 1)).from_address(ctypes.addressof(ptr.contents))
            buf._graph = self
            a = np.frombuffer(buf, dtype=dt)[:count]
            a.flags.writeable = False
            out.append(a)
        return tuple(out)

    def node_pointer(self):
        """A copy of the CSR offsets (u64[n+1]), without the edges."""
        ptr = np.empty(self.n + 1, dtype=np.uint64)
        check(lib.subway_graph_export(self._h, _ptr(ptr, _u64p), None, None))
        return ptr

    def out_degree(self):
        return np.diff(self.node_pointer()).astype(np.uint32)


@dataclass
class RunStats:
    ms_total: float
    ms_transfer: float
    ms_advance: float
    ms_frontier: float
    ms_generate: float
    ms_rounds: float
    h2d_bytes: int
    h2d_dma_bytes: int
    h2d_gather_bytes: int
    inner_rounds: int
    partitions: int
    updates: int
    advance_launches: int
    frontier_launches: int
    transfer_ops: int
    rounds_launches: int
    work_edges: int
    work_entries: int
    global_iterations: int
    work_items: int = 0
    mispredictions: int = 0
    fallback_iteration: int = 0
    edge_bytes: int = 0
    min_members: int = 0
    lane_work: int = 0
    lane_entries: int = 0
    pair_work_edges: int = 0
    pair_work_entries: int = 0
    pair_updates: int = 0
    resident_visits: int = 0
    resident_overlaps: int = 0
    pull_rounds: int = 0
    blocked_units: int = 0
    pull_found: int = 0
    h2d_blocked_bytes: int = 0
    blocked_edges: int = 0
    ms_rounds_blocked: float = 0.0
    rounds_launches_blocked: int = 0
    blk_work_edges: int = 0
    blk_work_entries: int = 0
    blk_updates: int = 0
    early_readback: int = 0
    values_patched: int = 0
    trace: list = field(default_factory=list)


ROUNDS = {"persistent": 0, "per-round": 1}


def _set_policy(h, params):
    if params is None:
        check(lib.subway_session_set_policy(h, None))
        return
    p = np.ascontiguousarray(params, np.float32)
    if p.size != 5211:
        raise ValueError(f"scheduler parameters: expected 5211 floats, got {p.size}")
    check(lib.subway_session_set_policy(h, p.ctypes.data_as(ctypes.c_void_p)))


def _options(algo, mode, source, budget, buffers, threshold, init_pr, transfer, policy, verbose, advance_grid,
             profile, rounds="persistent", rounds_grid=0, max_rounds=0, theta=0.5, spin_limit=0, packed=True,
             join_window=0, resident=True, resident_min_bytes=0, push_only=False, blocks=True, early_readback=0):
    o = _Options()
    lib.subway_default_options(ctypes.byref(o))
    o.algo = ALGOS[algo]
    o.mode = MODES[mode]
    o.source = source
    o.budget_bytes = budget
    o.num_buffers = buffers
    o.acc = threshold
    o.init_pr = init_pr
    o.transfer = TRANSFERS[transfer]
    o.policy = POLICIES[policy]
    o.verbose = int(verbose)
    o.advance_grid = advance_grid
    o.profile = int(profile)
    o.rounds_mode = ROUNDS[rounds]
    o.rounds_grid = rounds_grid
    o.max_rounds = max_rounds
    o.theta = theta
    o.spin_limit = spin_limit
    o.unpacked_edges = 0 if packed else 1
    o.join_window = join_window
    o.no_resident = 0 if resident else 1
    o.resident_min_bytes = resident_min_bytes
    o.push_only = int(push_only)
    o.no_blocks = 0 if blocks else 1
    o.early_readback = early_readback
    return o


def _stats(st, rec):
    trace = [dict(iteration=r.iteration, partitions=r.partitions, active_nodes=r.active_nodes,
                  active_edges=r.active_edges, h2d_bytes=r.h2d_bytes, ms=r.ms_iteration,
                  features=list(r.features), heads=list(r.heads), theta=r.theta, rounds_cap=r.rounds_cap,
                  on_plain=r.on_plain)
             for r in rec[: st.trace_len]]
    names = [f.name for f in RunStats.__dataclass_fields__.values() if f.name != "trace"]
    return RunStats(**{k: getattr(st, k) for k in names}, trace=trace)


_TRACE_CAP = 4096


class Session:
    """Device state and pipeline buffers of one (graph, algorithm, options)
    built once; :meth:`run` answers one query per call."""

    def __init__(self, graph: Graph, algo: str, mode: str = "sync", budget: int = 0, buffers: int = 2,
                 threshold: float = 0.01, init_pr: float = 0.15, transfer: str = "auto", policy: str = "plain",
                 verbose: bool = False, advance_grid: int = 0, profile: bool = False, rounds: str = "persistent",
                 rounds_grid: int = 0, max_rounds: int = 0, theta: float = 0.5, spin_limit: int = 0,
                 packed: bool = True, params=None, join_window: int = 0, resident: bool = True,
                 resident_min_bytes: int = 0, push_only: bool = False, blocks: bool = True,
                 early_readback: int = 0):
        """``resident``: a graph whose edges fit one buffer stays there, and
        the learned schedule visits in place the ranges whose whole edge slice
        a buffer still holds, when they hold ``resident_min_bytes`` (0: 8 MiB)
        of active edges or more.  ``push_only``: BFS on a resident graph
        takes top-down steps only (no bottom-up steps for dense frontiers).
        ``blocks``: on a graph with its block layout (Graph.build_blocks),
        whole id ranges move in that layout (False: as plain edges).
        ``early_readback``: 0 lets runs on graphs of 2^26 vertices or more
        copy their results to the host while their last iterations run
        (then patch what changed), -1 copies them at the end only, N > 0
        copies them after the first iteration on any graph with a patch
        list of N values (RunStats.early_readback says which happened)."""
        self.graph, self.algo = graph, algo
        o = _options(algo, mode, 0, budget, buffers, threshold, init_pr, transfer, policy, verbose, advance_grid,
                     profile, rounds, rounds_grid, max_rounds, theta, spin_limit, packed, join_window, resident,
                     resident_min_bytes, push_only, blocks, early_readback)
        h = ctypes.c_void_p()
        check(lib.subway_session_create(graph._h, ctypes.byref(o), ctypes.byref(h)))
        self._h = h
        _live_sessions.add(self)
        self._rec = (_IterRecord * _TRACE_CAP)()
        if params is not None:
            self.set_policy(params)

    def set_policy(self, params):
        """Scheduler network parameters for policy="learned" (None: the fixed
        theta and round cap of the options)."""
        _set_policy(self._h, params)

    def run(self, source: int = 0, copy: bool = True):
        """One query.  Returns (values, RunStats).  With copy=False the values
        are a read-only view of the session's host result array (no copy of V
        values): the view keeps the session alive, holds this run's values
        until the session's next run, and is invalid once close() has been
        called explicitly."""
        if not getattr(self, "_h", None):
            raise SubwayError(-1, "the session is closed")
        st = _Stats()
        dt = np.float32 if self.algo == "pr" else np.uint32
        if copy:
            vals = np.empty(self.graph.n, dtype=dt)
            out = vals.ctypes.data_as(ctypes.c_void_p)
        else:
            out = None
        check(lib.subway_session_run(self._h, source, out, ctypes.byref(st), self._rec, _TRACE_CAP))
        if not copy:
            ptr = lib.subway_session_values(self._h)
            buf = (ctypes.c_uint8 * (self.graph.n * 4)).from_address(ptr)
            buf._session = self   # the array's base keeps the session (and the result array) alive
            vals = np.frombuffer(buf, dtype=dt)
            vals.flags.writeable = False
        return vals, _stats(st, self._rec)

    def close(self):
        if getattr(self, "_h", None):
            lib.subway_session_free(self._h)
            self._h = None

    def __del__(self):
        self.close()


class BatchSession:
    """k <= 64 sources of BFS/SSSP/SSWP advancing together (async), one
    persistent kernel per transferred partition; :meth:`run` returns a
    (k, n) array (row i = values for sources[i])."""

    def __init__(self, graph: Graph, algo: str = "sssp", budget: int = 0, buffers: int = 2, transfer: str = "auto",
                 unroll: int = 0, rounds_grid: int = 0, profile: bool = False, policy: str = "plain",
                 theta: float = 0.5, packed: bool = True, params=None, join_window: int = 0,
                 resident: bool = True, resident_min_bytes: int = 0):
        self.graph, self.algo = graph, algo
        o = _options(algo, "async", 0, budget, buffers, 0.01, 0.15, transfer, policy, False, 0, profile,
                     "persistent", rounds_grid, 0, theta, 0, packed, join_window, resident, resident_min_bytes)
        o.batch = 1
        o.batch_unroll = unroll
        h = ctypes.c_void_p()
        check(lib.subway_session_create(graph._h, ctypes.byref(o), ctypes.byref(h)))
        self._h = h
        _live_sessions.add(self)
        if params is not None:
            _set_policy(self._h, params)

    def run(self, sources, rows_out: bool = True):
        src = np.ascontiguousarray(sources, dtype=np.uint32)
        k = len(src)
        rows = np.empty(self.graph.n * 64, dtype=np.uint32) if rows_out else None
        st = _Stats()
        rec = (_IterRecord * _TRACE_CAP)()
        check(lib.subway_session_run_batch(self._h, _ptr(src, _u32p), k, _ptr(rows, _u32p), ctypes.byref(st), rec,
                                           _TRACE_CAP))
        vals = rows.reshape(self.graph.n, 64)[:, :k].T.copy() if rows_out else None
        return vals, _stats(st, rec)

    def tune(self, sources):
        src = np.ascontiguousarray(sources, dtype=np.uint32)
        best = ctypes.c_int32()
        times = (ctypes.c_double * 4)()
        check(lib.subway_session_tune_batch(self._h, _ptr(src, _u32p), len(src), ctypes.byref(best), times))
        return int(best.value), list(times)

    def close(self):
        if getattr(self, "_h", None):
            lib.subway_session_free(self._h)
            self._h = None

    def __del__(self):
        self.close()


def run(graph: Graph, algo: str, mode: str = "sync", source: int = 0, budget: int = 0, buffers: int = 2,
        threshold: float = 0.01, init_pr: float = 0.15, transfer: str = "auto", policy: str = "plain",
        verbose: bool = False, advance_grid: int = 0, profile: bool = False, rounds: str = "persistent",
        rounds_grid: int = 0, max_rounds: int = 0, theta: float = 0.5, spin_limit: int = 0, packed: bool = True,
        params=None, join_window: int = 0, resident: bool = True, push_only: bool = False, blocks: bool = True):
    """Run one application ("sssp", mode "async" = sssp-async) to its fixpoint.

    Returns (values, RunStats); values are uint32 (BFS/CC/SSSP/SSWP) or
    float32 (PR).  ``budget`` is the device edge budget in bytes (0 = 90% of
    device memory, the reference's sizing).  ``params``: scheduler network
    for policy="learned".  ``join_window``: how long (100 MHz ticks) a
    persistent launch waits for its workgroups before closing membership.
    ``resident``: a graph whose edges fit one device buffer moves once per
    run and stays there (False: the active edges move every iteration).
    ``push_only``: BFS on a resident graph without bottom-up steps.
    ``blocks``: whole id ranges of a graph with its block layout move in it.
    """
    if params is not None:
        s = Session(graph, algo, mode=mode, budget=budget, buffers=buffers, threshold=threshold, init_pr=init_pr,
                    transfer=transfer, policy=policy, verbose=verbose, advance_grid=advance_grid, profile=profile,
                    rounds=rounds, rounds_grid=rounds_grid, max_rounds=max_rounds, theta=theta,
                    spin_limit=spin_limit, packed=packed, params=params, join_window=join_window,
                    resident=resident, push_only=push_only, blocks=blocks)
        try:
            return s.run(source)
        finally:
            s.close()
    o = _options(algo, mode, source, budget, buffers, threshold, init_pr, transfer, policy, verbose, advance_grid,
                 profile, rounds, rounds_grid, max_rounds, theta, spin_limit, packed, join_window, resident,
                 push_only=push_only, blocks=blocks)
    vals = np.empty(graph.n, dtype=np.float32 if algo == "pr" else np.uint32)
    st = _Stats()
    rec = (_IterRecord * _TRACE_CAP)()
    check(lib.subway_run(graph._h, ctypes.byref(o), vals.ctypes.data_as(ctypes.c_void_p), ctypes.byref(st),
                         rec, _TRACE_CAP))
    return vals, _stats(st, rec)


def generate(label1, label2, out_degree):
    """Subgraph generation on the GPU: (activeNodes, activeNodesPointer[A+1])."""
    l1 = np.ascontiguousarray(label1, dtype=np.uint8)
    l2 = np.ascontiguousarray(label2, dtype=np.uint8)
    deg = np.ascontiguousarray(out_degree, dtype=np.uint32)
    n = len(deg)
    an = np.empty(n, dtype=np.uint32)
    ap = np.empty(n + 1, dtype=np.uint64)
    cnt = ctypes.c_uint64()
    check(lib.subway_generate(n, _ptr(l1, _u8p), _ptr(l2, _u8p), _ptr(deg, _u32p), _ptr(an, _u32p),
                              _ptr(ap, _u64p), ctypes.byref(cnt)))
    a = int(cnt.value)
    return an[:a].copy(), ap[: a + 1].copy()


def generate_pr(delta, acc, out_degree):
    d = np.ascontiguousarray(delta, dtype=np.float32)
    deg = np.ascontiguousarray(out_degree, dtype=np.uint32)
    n = len(deg)
    an = np.empty(n, dtype=np.uint32)
    ap = np.empty(n + 1, dtype=np.uint64)
    cnt = ctypes.c_uint64()
    check(lib.subway_generate_pr(n, _ptr(d, _f32p), ctypes.c_float(acc), _ptr(deg, _u32p), _ptr(an, _u32p),
                                 _ptr(ap, _u64p), ctypes.byref(cnt)))
    a = int(cnt.value)
    return an[:a].copy(), ap[: a + 1].copy()


def partition(active_nodes_pointer, active_nodes, budget_edges: int):
    """Device partitioner: rows of (fromNode, nodeCount, fromEdge, edgeCount, firstId, lastId)."""
    ap = np.ascontiguousarray(active_nodes_pointer, dtype=np.uint64)
    an = np.ascontiguousarray(active_nodes, dtype=np.uint32)
    a = len(an)
    cap = max(16, a + 1)
    rows = np.zeros((cap, 6), dtype=np.uint64)
    nr = ctypes.c_uint64()
    check(lib.subway_partition(a, _ptr(ap, _u64p), _ptr(an, _u32p), budget_edges, cap, _ptr(rows, _u64p),
                               ctypes.byref(nr)))
    return rows[: int(nr.value)].copy()


def gather(graph: Graph, active_nodes):
    """GPU gather of the active vertices' adjacency (ends, weights or None)."""
    an = np.ascontiguousarray(active_nodes, dtype=np.uint32)
    ne = ctypes.c_uint64()
    check(lib.subway_gather(graph._h, len(an), _ptr(an, _u32p), None, ctypes.byref(ne)))
    m = int(ne.value)
    width = 2 if graph.weighted else 1
    out = np.empty(m * width, dtype=np.uint32)
    if m:
        check(lib.subway_gather(graph._h, len(an), _ptr(an, _u32p), out.ctypes.data_as(ctypes.c_void_p),
                                ctypes.byref(ne)))
    if graph.weighted:
        return out[0::2].copy(), out[1::2].copy()
    return out, None


def h2d_peak(nbytes: int = 256 << 20, reps: int = 10) -> float:
    """Measured pinned host-to-device copy peak in GB/s (best of reps)."""
    out = ctypes.c_double()
    check(lib.subway_h2d_peak(nbytes, reps, ctypes.byref(out)))
    return float(out.value)


def probe_copy(nbytes: int = 256 << 20) -> float:
    """One launch of a plain copy kernel of nbytes read + nbytes written (ms)."""
    out = ctypes.c_double()
    check(lib.subway_probe_copy(nbytes, ctypes.byref(out)))
    return float(out.value)
