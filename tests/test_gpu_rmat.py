"""The GPU build of the synthetic R-MAT inputs: the same CSR as the host
generator, byte for byte, for every layout the host generator produces
(unweighted, 8-byte weighted, symmetrised, split), and the files written from
a split-only graph equal the files of the same graph built on the host."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _same_csr(a, b):
    pa, ea, wa = a.csr()
    pb, eb, wb = b.csr()
    assert (a.n, a.m, a.max_degree, a.max_degree_vertex) == (b.n, b.m, b.max_degree, b.max_degree_vertex)
    assert np.array_equal(pa, pb)
    assert np.array_equal(ea, eb)
    if wa is None:
        assert wb is None
    else:
        assert np.array_equal(wa, wb)


@pytest.mark.parametrize("scale,ef", [(1, 4), (5, 3), (10, 16), (14, 16), (17, 8)])
@pytest.mark.parametrize("weighted", [False, True])
@pytest.mark.parametrize("symmetrize", [False, True])
def test_device_build_equals_host_build(sw, gpu, scale, ef, weighted, symmetrize):
    kw = dict(seed=scale * 7 + ef, weighted=weighted, symmetrize=symmetrize)
    host = sw.Graph.rmat(scale, ef, build="host", **kw)
    dev = sw.Graph.rmat(scale, ef, build="device", **kw)
    _same_csr(host, dev)


def test_device_build_other_parameters(sw, gpu):
    """Quadrant probabilities, seed and weight range reach the device build."""
    kw = dict(a=0.45, b=0.25, c=0.15, seed=99, weighted=True, wmin=3, wmax=200)
    _same_csr(sw.Graph.rmat(15, 12, build="host", **kw), sw.Graph.rmat(15, 12, build="device", **kw))


def test_device_build_split_layout(sw, gpu, tmp_path):
    """A split-layout graph built on the device holds only the pinned split
    arrays; its export, its zero-copy views and the .bwcsr file written from
    it equal those of the host build, and SSSP runs on it unchanged."""
    host = sw.Graph.rmat(16, 16, seed=5, weighted=True, split_layout=True, build="host")
    dev = sw.Graph.rmat(16, 16, seed=5, weighted=True, split_layout=True, build="device")
    _same_csr(host, dev)
    p, e, w = dev.split_view()
    hp, he, hw = host.csr()
    assert np.array_equal(p, hp) and np.array_equal(e, he) and np.array_equal(w.astype(np.uint32), hw)
    assert not e.flags.writeable
    fa, fb = tmp_path / "h.bwcsr", tmp_path / "d.bwcsr"
    host.save(str(fa))
    dev.save(str(fb))
    assert fa.read_bytes() == fb.read_bytes()
    src = dev.max_degree_vertex
    vh, _ = sw.run(host, "sssp", mode="async", source=src, budget=1 << 20)
    vd, st = sw.run(dev, "sssp", mode="async", source=src, budget=1 << 20)
    assert np.array_equal(vh, vd) and st.edge_bytes == 5


def test_device_build_scale22(sw, gpu):
    """A graph of 2^26 generated edges (16 tiles of the sort per digit pass
    and then some): still the host's CSR."""
    kw = dict(seed=42, weighted=True)
    _same_csr(sw.Graph.rmat(22, 16, build="host", **kw), sw.Graph.rmat(22, 16, build="device", **kw))
