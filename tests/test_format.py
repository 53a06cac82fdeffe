"""On-disk formats (CPU): .bcsr/.bwcsr byte layout, the 64-bit variant,
.el/.wel ingestion and the converter, checked against the golden files and
against the reference's own tools/converter.cpp compiled from source."""
import os
import shutil
import struct
import subprocess

import numpy as np
import pytest

from tests import _oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
ROOT = os.path.dirname(HERE)
CONVERTER = os.path.join(ROOT, "subway_amd", "bin", "converter")


@pytest.mark.parametrize("ext,weighted", [("el", False), ("wel", True)])
def test_readme_graph_converts_byte_exact(sw, tmp_path, ext, weighted):
    src = tmp_path / f"Graph.{ext}"
    shutil.copy(os.path.join(GOLD, f"readme_graph.{ext}"), src)
    g = sw.Graph.load(str(src))
    out = tmp_path / ("Graph.bwcsr" if weighted else "Graph.bcsr")
    g.save(str(out))
    exp = open(os.path.join(GOLD, "readme_graph.expected." + ("bwcsr" if weighted else "bcsr")), "rb").read()
    assert out.read_bytes() == exp


def test_converter_binary_matches_golden(tmp_path):
    if not os.path.exists(CONVERTER):
        pytest.skip("converter not built")
    src = tmp_path / "Graph.wel"
    shutil.copy(os.path.join(GOLD, "readme_graph.wel"), src)
    subprocess.run([CONVERTER, str(src)], check=True, capture_output=True)
    assert (tmp_path / "Graph.bwcsr").read_bytes() == open(os.path.join(GOLD, "readme_graph.expected.bwcsr"),
                                                           "rb").read()


def write_el(path, src, dst, w=None):
    with open(path, "w") as f:
        if w is None:
            f.writelines(f"{a} {b}\n" for a, b in zip(src, dst))
        else:
            f.writelines(f"{a} {b} {c}\n" for a, b, c in zip(src, dst, w))


@pytest.mark.parametrize("weighted", [False, True])
def test_matches_reference_converter(sw, tmp_path, weighted):
    """Byte-for-byte against the reference's converter built from its own
    source (oracle/_ref); it needs more than ~32k vertices to run at all."""
    if not os.path.exists(O.REF_CONVERTER):
        pytest.skip("reference converter not built (reference tree absent)")
    rng = np.random.default_rng(4)
    n, m = 40_000, 200_000
    s = rng.integers(0, n, m)
    d = rng.integers(0, n, m)
    s[0], d[0] = n - 1, n - 2  # make max id = n-1
    w = rng.integers(1, 65, m) if weighted else None
    ext = "wel" if weighted else "el"
    ref_in = tmp_path / "ref" / f"g.{ext}"
    our_in = tmp_path / "our" / f"g.{ext}"
    ref_in.parent.mkdir()
    our_in.parent.mkdir()
    write_el(ref_in, s, d, w)
    shutil.copy(ref_in, our_in)
    subprocess.run([O.REF_CONVERTER, str(ref_in)], check=True, capture_output=True)
    out_ext = "bwcsr" if weighted else "bcsr"
    g = sw.Graph.load(str(our_in))
    g.save(str(our_in.with_suffix("." + out_ext)))
    assert our_in.with_suffix("." + out_ext).read_bytes() == ref_in.with_suffix("." + out_ext).read_bytes()


@pytest.mark.parametrize("weighted", [False, True])
def test_64bit_variant_round_trip(sw, tmp_path, weighted):
    g = sw.Graph.rmat(11, 8, seed=5, weighted=weighted)
    ext = "bwcsr64" if weighted else "bcsr64"
    p = tmp_path / f"g.{ext}"
    g.save(str(p))
    data = p.read_bytes()
    n, m = struct.unpack("<QQ", data[:16])
    assert (n, m) == (g.n, g.m)
    assert len(data) == 16 + 8 * n + m * (8 if weighted else 4)
    g2 = sw.Graph.load(str(p))
    for a, b in zip(g.csr(), g2.csr()):
        if a is not None:
            assert np.array_equal(a, b)


@pytest.mark.parametrize("ext", ["bwcsr", "bwcsr64"])
def test_split_layout_read_keeps_the_csr(sw, tmp_path, ext):
    """A .bwcsr file read for the split layout goes through ordinary memory
    first; the CSR read back is the file's (without a GPU nothing is pinned;
    the split arrays themselves are checked by the GPU tests).  Text files
    refuse the layout."""
    g = sw.Graph.rmat(11, 8, seed=5, weighted=True)
    p = tmp_path / f"g.{ext}"
    g.save(str(p))
    g2 = sw.Graph.load(str(p), split_layout=True)
    for a, b in zip(g.csr(), g2.csr()):
        assert np.array_equal(a, b)
    t = tmp_path / "g.wel"
    t.write_text("0 1 3\n1 2 4\n")
    with pytest.raises(sw.SubwayError, match="split layout"):
        sw.Graph.load(str(t), split_layout=True)


def test_legacy_layout_fields(sw, tmp_path):
    g = sw.Graph.rmat(10, 4, seed=9)
    p = tmp_path / "g.bcsr"
    g.save(str(p))
    data = p.read_bytes()
    n, m = struct.unpack("<II", data[:8])
    ptr, ends, _ = g.csr()
    assert (n, m) == (g.n, g.m)
    assert np.array_equal(np.frombuffer(data[8:8 + 4 * n], np.uint32), ptr[:n].astype(np.uint32))
    assert np.array_equal(np.frombuffer(data[8 + 4 * n:], np.uint32), ends)


def test_text_parsing_comments_and_blank_lines(sw, tmp_path):
    p = tmp_path / "g.wel"
    p.write_text("# a comment\n0 1 5\n\n% another\n1 2 7\r\n2 0 1\n")
    g = sw.Graph.load(str(p))
    ptr, ends, w = g.csr()
    assert g.n == 3 and g.m == 3
    assert ends.tolist() == [1, 2, 0] and w.tolist() == [5, 7, 1]


def test_text_missing_weight_is_an_error(sw, tmp_path):
    p = tmp_path / "g.wel"
    p.write_text("0 1 5\n1 2\n")
    with pytest.raises(sw.SubwayError):
        sw.Graph.load(str(p))


def test_stable_in_source_order(sw, tmp_path):
    """Edges of a vertex keep their input order (tools/converter.cpp:76-83)."""
    p = tmp_path / "g.el"
    p.write_text("2 9\n0 5\n2 1\n0 3\n2 4\n")
    g = sw.Graph.load(str(p))
    ptr, ends, _ = g.csr()
    assert ptr.tolist() == [0, 2, 2, 5, 5, 5, 5, 5, 5, 5, 5]
    assert ends.tolist() == [5, 3, 9, 1, 4]


def test_outdegree_of_every_vertex(sw):
    """All degrees, including vertex n-2 (never written by the reference,
    shared/graph.cu:177-181)."""
    p = np.array([0, 1, 3, 6, 6], np.uint64)
    e = np.array([1, 0, 2, 0, 1, 3], np.uint32)
    g = sw.Graph.from_csr(p, e)
    assert g.out_degree().tolist() == [1, 2, 3, 0]


# ------------------------------------------------------------------ R-MAT
def test_rmat_deterministic_and_thread_invariant(sw):
    code = ("import subway_amd as sw, hashlib;g=sw.Graph.rmat(14,8,seed=77,weighted=True);p,e,w=g.csr();"
            "print(hashlib.sha256(p.tobytes()+e.tobytes()+w.tobytes()).hexdigest())")
    outs = set()
    for t in ("1", "3", "8"):
        env = dict(os.environ, SUBWAY_HOST_THREADS=t)
        r = subprocess.run(["python", "-c", code], cwd=ROOT, env=env, capture_output=True, text=True, check=True)
        outs.add(r.stdout.strip())
    assert len(outs) == 1


def test_rmat_shape(sw):
    g = sw.Graph.rmat(12, 16, seed=42, weighted=True)
    ptr, ends, w = g.csr()
    assert g.n == 4096 and 0.98 * 16 * 4096 < g.m <= 16 * 4096
    src = np.repeat(np.arange(g.n), np.diff(ptr).astype(np.int64))
    assert not np.any(src == ends)                     # self-loops dropped
    assert w.min() >= 1 and w.max() <= 64
    top = src >= g.n // 2                               # source top bit set: prob c + d = 0.24
    assert abs(top.mean() - 0.24) < 0.02
    assert g.max_degree_vertex == 0


def test_rmat_symmetrize(sw):
    g = sw.Graph.rmat(10, 8, seed=3, symmetrize=True)
    ptr, ends, _ = g.csr()
    src = np.repeat(np.arange(g.n), np.diff(ptr).astype(np.int64))
    fwd = sorted(zip(src.tolist(), ends.tolist()))
    rev = sorted(zip(ends.tolist(), src.tolist()))
    assert fwd == rev
