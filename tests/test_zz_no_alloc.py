"""subway_generate_noalloc (the generator over caller-owned device arrays,
with its scratch in a fixed file-scope device array) against
subway_generate (host arrays in, device copies allocated per call): the
same active list, edge offsets and count, on two calls in a row, the second
with a smaller input, and again after other launches.  The test itself
allocates its device arrays through the HIP runtime the library uses."""
import ctypes
import os

import numpy as np
import pytest

from tests import _oracle as O

pytestmark = pytest.mark.gpu

_hip = None


def hip():
    global _hip
    if _hip is None:
        _hip = ctypes.CDLL(os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so"))
        _hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        _hip.hipFree.argtypes = [ctypes.c_void_p]
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
    return _hip


H2D, D2H = 1, 2


class Dev:
    """A device array of `nbytes`, freed by close()."""

    def __init__(self, nbytes):
        self.p = ctypes.c_void_p()
        assert hip().hipMalloc(ctypes.byref(self.p), max(int(nbytes), 1)) == 0
        self.nbytes = int(nbytes)

    def put(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        assert hip().hipMemcpy(self.p, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, H2D) == 0

    def get(self, dtype, count):
        out = np.empty(count, dtype)
        assert hip().hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.p, out.nbytes, D2H) == 0
        return out

    def close(self):
        if self.p:
            hip().hipFree(self.p)
            self.p = ctypes.c_void_p()


def _inputs(n, density, seed):
    rng = np.random.default_rng(seed)
    l1 = (rng.random(n) < density).astype(np.uint8)
    l2 = (rng.random(n) < density / 2).astype(np.uint8)
    deg = rng.integers(0, 300, n).astype(np.uint32)
    return l1, l2, deg


def _noalloc(sw, l1, l2, deg):
    n = len(deg)
    bufs = [Dev(n), Dev(n), Dev(4 * n), Dev(4 * n), Dev(8 * (n + 1))]
    try:
        bufs[0].put(l1)
        bufs[1].put(l2)
        bufs[2].put(deg)
        cnt = ctypes.c_uint64()
        rc = sw._lib.lib.subway_generate_noalloc(n, bufs[0].p, bufs[1].p, bufs[2].p, bufs[3].p, bufs[4].p,
                                                   ctypes.byref(cnt))
        sw._lib.check(rc)
        a = int(cnt.value)
        return bufs[3].get(np.uint32, a), bufs[4].get(np.uint64, a + 1)
    finally:
        for b in bufs:
            b.close()


def test_noalloc_generate_matches_generate(sw, gpu):
    # two calls in a row, the second smaller; then a different size again
    # after an ordinary run has used the device in between
    for n, density, seed in ((1_048_583, 0.3, 1), (70_001, 0.05, 2)):
        l1, l2, deg = _inputs(n, density, seed)
        an, ap = _noalloc(sw, l1, l2, deg)
        en, ep = sw.generate(l1, l2, deg)
        assert np.array_equal(an, en) and np.array_equal(ap, ep), n
        # and the closed form: the active ids ascending, exclusive degree prefix
        act = np.flatnonzero((l1 | l2) != 0).astype(np.uint32)
        assert np.array_equal(an, act)
        assert np.array_equal(ap[:-1], np.concatenate([[0], np.cumsum(deg[act], dtype=np.uint64)[:-1]]))
        assert int(ap[-1]) == int(deg[act].sum(dtype=np.uint64))
    g = sw.Graph.rmat(12, 16, seed=3)
    ptr, adj, _ = g.csr()
    vals, _ = sw.run(g, "bfs", mode="sync", source=0, budget=64 << 10)
    assert np.array_equal(vals, O.bfs(ptr, adj, 0))
    l1, l2, deg = _inputs(5, 1.0, 3)
    an, ap = _noalloc(sw, l1, l2, deg)
    en, ep = sw.generate(l1, l2, deg)
    assert np.array_equal(an, en) and np.array_equal(ap, ep)


def test_noalloc_generate_rejects_oversize(sw, gpu):
    cnt = ctypes.c_uint64()
    rc = sw._lib.lib.subway_generate_noalloc(1 << 32, ctypes.c_void_p(16), ctypes.c_void_p(16),
                                               ctypes.c_void_p(16), ctypes.c_void_p(16), ctypes.c_void_p(16),
                                               ctypes.byref(cnt))
    assert rc == -1
    assert "at most" in sw._lib.lib.subway_last_error().decode()
