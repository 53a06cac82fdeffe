"""CPU tests of the C ABI boundary: the library loads, exports every symbol
declared in include/subway_amd.h, and reports host-side errors as codes."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "subway_amd.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(subway_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_the_path():
    syms = declared_symbols()
    for s in ("subway_load", "subway_generate", "subway_partition", "subway_run", "subway_gather"):
        assert s in syms


def test_library_exports_every_declared_symbol(sw):
    lib = ctypes.CDLL(sw.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_python_bindings_cover_the_header(sw):
    from subway_amd import _lib
    assert sorted(_lib.EXPORTED) == declared_symbols()


def test_struct_layouts_match_the_library(sw):
    """The ctypes mirrors of subway_options / subway_stats /
    subway_iter_record have the sizes the library was built with (the import
    refuses a mismatch; this names the sizes)."""
    from subway_amd import _Options, _Stats, _IterRecord, _lib
    got = [ctypes.c_uint32() for _ in range(3)]
    _lib.lib.subway_struct_sizes(*(ctypes.cast(ctypes.byref(x), ctypes.POINTER(ctypes.c_uint32)) for x in got))
    assert [x.value for x in got] == [ctypes.sizeof(_Options), ctypes.sizeof(_Stats), ctypes.sizeof(_IterRecord)]
    assert ctypes.sizeof(_Stats) == _Stats.values_patched.offset + 8


def test_missing_file_is_an_io_error(sw, tmp_path):
    with pytest.raises(sw.SubwayError) as e:
        sw.Graph.load(str(tmp_path / "nope.bcsr"))
    assert e.value.code == -2


def test_unsupported_extension(sw, tmp_path):
    p = tmp_path / "g.txt"
    p.write_text("0 1\n")
    with pytest.raises(sw.SubwayError) as e:
        sw.Graph.load(str(p))
    assert e.value.code == -3


def test_truncated_file(sw, tmp_path):
    g = sw.Graph.rmat(8, 4, seed=1)
    p = tmp_path / "g.bcsr"
    g.save(str(p))
    data = p.read_bytes()
    p.write_bytes(data[:-3])
    with pytest.raises(sw.SubwayError) as e:
        sw.Graph.load(str(p))
    assert e.value.code == -3


def test_extension_must_match_edge_type(sw, tmp_path):
    g = sw.Graph.rmat(8, 4, seed=1, weighted=True)
    with pytest.raises(sw.SubwayError):
        g.save(str(tmp_path / "g.bcsr"))


def test_bad_csr_rejected(sw):
    with pytest.raises(sw.SubwayError):
        sw.Graph.from_csr(np.array([0, 2], np.uint64), np.array([0, 5], np.uint32))  # endpoint >= n
    with pytest.raises(sw.SubwayError):
        sw.Graph.from_csr(np.array([0, 3], np.uint64), np.array([0, 0], np.uint32))  # ptr[n] != m


def test_run_needs_a_gpu_when_there_is_none(sw):
    if sw.device_count() > 0:
        pytest.skip("a GPU is visible")
    g = sw.Graph.rmat(8, 4, seed=1)
    with pytest.raises(sw.SubwayError):
        sw.run(g, "bfs")
