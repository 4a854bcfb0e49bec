"""Versioned GPU-layout cache (hg_gpu_index_save / hg_gpu_index_load): a saved export loads
straight to the device and searches bit-identically; a damaged or foreign file is refused."""
import numpy as np
import pytest

from helpers import built_index

pytestmark = pytest.mark.gpu


def test_layout_cache_roundtrip(hg, tmp_path):
    p, _, data, queries = built_index("c2s")
    print("case=c2s layout cache", flush=True)
    gi = p.gpu()
    path = str(tmp_path / "layout.hggl")
    gi.save(path)
    g2 = hg.GpuHnsw.load(path)
    l0, d0, c0, _ = gi.search(queries, 10, 40, counters=True)
    l1, d1, c1, _ = g2.search(queries, 10, 40, counters=True)
    np.testing.assert_array_equal(l0, l1)
    np.testing.assert_array_equal(d0.view(np.uint32), d1.view(np.uint32))
    np.testing.assert_array_equal(c0, c1)
    np.testing.assert_array_equal(gi.labels(), g2.labels())
    g2.free()


def test_layout_cache_keeps_delete_marks(hg, tmp_path):
    p, _, data, queries = built_index("c1")
    print("case=c1 layout cache deletions", flush=True)
    q = hg.Index("l2", 16)
    blob_path = str(tmp_path / "c1.bin")
    p.save_index(blob_path)
    q.load_index(blob_path)
    for lab in range(0, 10000, 3):
        q.mark_deleted(lab)
    q.set_ef(50)
    lab0, dis0 = q.knn_query(queries, k=10)
    path = str(tmp_path / "del.hggl")
    q.gpu().save(path)
    g2 = hg.GpuHnsw.load(path)
    lab1, dis1, _, short = g2.search(queries, 10, 50)
    assert not short
    np.testing.assert_array_equal(lab0, lab1)
    np.testing.assert_array_equal(dis0.view(np.uint32), dis1.view(np.uint32))
    assert (lab1 % 3 != 0).all()
    g2.free()


def test_layout_cache_refuses_damaged_files(hg, tmp_path):
    p, _, _, _ = built_index("c1")
    print("case=c1 layout cache damaged", flush=True)
    path = str(tmp_path / "ok.hggl")
    p.gpu().save(path)
    raw = bytearray(open(path, "rb").read())
    bad = bytearray(raw)
    bad[len(bad) // 2] ^= 0x40
    open(str(tmp_path / "bad.hggl"), "wb").write(bytes(bad))
    with pytest.raises(RuntimeError, match="corrupted"):
        hg.GpuHnsw.load(str(tmp_path / "bad.hggl"))
    other = bytearray(raw)
    other[4] = 99  # format version
    open(str(tmp_path / "v99.hggl"), "wb").write(bytes(other))
    with pytest.raises(RuntimeError, match="another version"):
        hg.GpuHnsw.load(str(tmp_path / "v99.hggl"))
    open(str(tmp_path / "short.hggl"), "wb").write(bytes(raw[:1000]))
    with pytest.raises(RuntimeError, match="corrupted"):
        hg.GpuHnsw.load(str(tmp_path / "short.hggl"))
