"""CPU-side checks: the C ABI library loads and exports every declared entry point; host logic
(construction, persistence, error behaviour) matches the reference."""
import ctypes
import os
import re

import numpy as np
import pytest

from helpers import ROOT, built_index


def declared_functions():
    text = open(os.path.join(ROOT, "include", "hnswgpu.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hg_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(hg):
    from hnsw_rabitq_amd import _lib
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = declared_functions()
    assert len(names) > 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # every ctypes signature refers to a declared function
    assert set(_lib.SIGNATURES) <= set(names)


def test_save_load_roundtrip_bytes(hg, tmp_path):
    p, space, data, queries = built_index("c1")
    path = str(tmp_path / "c1.bin")
    p.save_index(path)
    q = hg.Index("l2", 16)
    q.load_index(path)
    assert q.get_current_count() == data.shape[0]
    assert q.serialize().tobytes() == open(path, "rb").read()


def test_corrupted_file_is_rejected(hg, tmp_path):
    p, space, data, queries = built_index("c1")
    blob = p.serialize().tobytes()
    path = str(tmp_path / "bad.bin")
    with open(path, "wb") as f:
        f.write(blob[:-7])
    q = hg.Index("l2", 16)
    with pytest.raises(RuntimeError, match="corrupted or unsupported"):
        q.load_index(path)
    with pytest.raises(RuntimeError, match="Cannot open file"):
        q.load_index(str(tmp_path / "missing.bin"))


def test_capacity_and_argument_errors(hg):
    with pytest.raises(RuntimeError, match="Space name must be one of l2, ip, or cosine."):
        hg.Index("hamming", 4)
    p = hg.Index("l2", 4)
    p.init_index(3)
    with pytest.raises(RuntimeError, match="The index is already initiated."):
        p.init_index(3)
    with pytest.raises(RuntimeError, match="Wrong dimensionality of the vectors"):
        p.add_items(np.zeros((2, 5), np.float32))
    with pytest.raises(RuntimeError, match="exceeds the specified limit"):
        p.add_items(np.random.default_rng(0).random((4, 4), dtype=np.float32))
    with pytest.raises(TypeError, match="filter must be a callable"):
        p.knn_query(np.zeros(4, np.float32), k=1, filter=[1, 2])


def test_multithreaded_build_is_valid(hg, orc):
    """Parallel insertion (schedule dependent, as in the reference) still yields a valid graph
    that the oracle searches with high self-recall."""
    p, space, data, queries = built_index("c1", num_threads=8)
    o = orc.OracleHnsw(blob=p.serialize())
    lab, dis, _, _, _ = o.search(0, data[:500], 1, 50, mode=orc.MODE_HEAP)
    assert (lab[:, 0] == np.arange(500)).mean() > 0.99


def test_cosine_normalisation_matches_reference_bits(hg, orc):
    ref = orc.ref_module()
    if ref is None:
        pytest.skip("reference module not built here")
    rng = np.random.default_rng(3)
    for dim in (3, 5, 16, 21, 40):
        x = rng.standard_normal((50, dim)).astype(np.float32)
        p = ref.Index("cosine", dim)
        p.init_index(50)
        p.add_items(x, num_threads=1)
        stored = np.asarray(p.get_items(np.arange(50)), dtype=np.float32)
        np.testing.assert_array_equal(hg.normalize_rows(x), stored)
