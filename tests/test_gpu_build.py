"""GPU level-0 construction (hg_hnsw_build_gpu): the graph is a valid HNSW graph of the same
kind as the reference's addPoint builds, checked statistically against the CPU build (the
reference's insertion order cannot be kept within a batch, so the graph is not the same one):
structure invariants, identical upper levels, mean degree and recall@10 at fixed ef within a
stated margin of the CPU-built graph's, and the GPU search on it equal to the oracle's."""
import numpy as np
import pytest

from hnsw_rabitq_amd import datasets
from helpers import built_index, workload

pytestmark = pytest.mark.gpu

HEADER = 96


def _links0(p):
    blob = np.frombuffer(p.serialize(), np.uint8)
    hdr = blob[:HEADER]
    n = int(hdr[16:24].view(np.uint64)[0])
    per = int(hdr[24:32].view(np.uint64)[0])
    maxm0 = int(hdr[64:72].view(np.uint64)[0])
    lv0 = blob[HEADER:HEADER + int(hdr[8:16].view(np.uint64)[0]) * per].reshape(-1, per)[:n]
    cnt = lv0[:, 0:2].copy().view(np.uint16)[:, 0].astype(np.int64)
    ids = lv0[:, 4:4 + 4 * maxm0].copy().view(np.uint32)
    return blob, n, per, cnt, ids, maxm0


def _recall(lab, gt):
    return float(np.mean([len(np.intersect1d(a, b)) for a, b in zip(lab, gt)])) / gt.shape[1]


def _check_graph(p):
    blob, n, per, cnt, ids, maxm0 = _links0(p)
    assert (cnt <= maxm0).all()
    for i in range(n):
        row = ids[i, :cnt[i]]
        assert (row < n).all() and (row != i).all()
        assert len(np.unique(row)) == len(row)
    return cnt


@pytest.mark.parametrize("case,ef", [("c1", 50), ("c2s", 40)])
def test_gpu_build_matches_cpu_build_statistically(orc, hg, case, ef):
    space, data, queries = workload(case)
    print(f"case={case} gpu build n={data.shape[0]} dim={data.shape[1]}", flush=True)
    cpu, _, _, _ = built_index(case)
    gpu = hg.Index(space, data.shape[1])
    gpu.init_index(data.shape[0], M=16, ef_construction=200, random_seed=100)
    gpu.add_items_gpu(data, num_threads=1)
    cnt_g = _check_graph(gpu)
    blob_c, n, per, cnt_c, _, _ = _links0(cpu)
    blob_g = _links0(gpu)[0]
    # levels above 0, element levels, entry point, vectors and labels: identical
    np.testing.assert_array_equal(blob_c[:HEADER], blob_g[:HEADER])
    np.testing.assert_array_equal(blob_c[HEADER + int(blob_c[8:16].view(np.uint64)[0]) * per:],
                                  blob_g[HEADER + int(blob_g[8:16].view(np.uint64)[0]) * per:])
    print(f"mean level-0 degree: cpu {cnt_c.mean():.2f} gpu {cnt_g.mean():.2f}", flush=True)
    assert abs(cnt_g.mean() - cnt_c.mean()) <= 0.1 * cnt_c.mean()
    sp = 0 if space == "l2" else 1
    gt, _ = orc.bf_search(sp, data, queries, 10)
    cpu.set_ef(ef)
    gpu.set_ef(ef)
    rc = _recall(cpu.knn_query(queries, k=10)[0], gt)
    rg = _recall(gpu.knn_query(queries, k=10)[0], gt)
    print(f"recall@10 at ef={ef}: cpu-built {rc:.4f} gpu-built {rg:.4f}", flush=True)
    assert rg >= rc - 0.01
    # the search on the GPU-built graph is still the oracle's
    o = orc.OracleHnsw(blob=gpu.serialize())
    lt, dt, _, _, _ = o.search(sp, queries, 10, ef, mode=orc.MODE_TOTAL)
    lab, dis = gpu.knn_query(queries, k=10)
    np.testing.assert_array_equal(lab, lt)
    np.testing.assert_array_equal(dis.view(np.uint32), dt.view(np.uint32))


def test_gpu_build_incremental(orc, hg):
    """Rows added to a CPU-built index: the existing lists gain links, nothing breaks."""
    space, data, queries = workload("c1")
    print("case=c1 incremental gpu build", flush=True)
    p = hg.Index(space, data.shape[1])
    p.init_index(data.shape[0], M=16, ef_construction=200, random_seed=100)
    half = data.shape[0] // 2
    p.add_items(data[:half], num_threads=1)
    p.add_items_gpu(data[half:], ids=np.arange(half, data.shape[0]), num_threads=1)
    _check_graph(p)
    assert p.get_current_count() == data.shape[0]
    gt, _ = orc.bf_search(0, data, queries, 10)
    p.set_ef(50)
    r = _recall(p.knn_query(queries, k=10)[0], gt)
    print(f"recall@10 at ef=50 after an incremental GPU build: {r:.4f}", flush=True)
    assert r >= 0.95


def _upper_lists(p):
    """Per element: its level and its lists at levels 1..level, from the serialized index."""
    blob = np.frombuffer(p.serialize(), np.uint8)
    hdr = blob[:HEADER]
    n = int(hdr[16:24].view(np.uint64)[0])
    per = int(hdr[24:32].view(np.uint64)[0])
    maxlevel = int(hdr[48:52].view(np.int32)[0])
    entry = int(hdr[52:56].view(np.uint32)[0])
    maxm = int(hdr[56:64].view(np.uint64)[0])
    size_links = maxm * 4 + 4
    off = HEADER + int(hdr[8:16].view(np.uint64)[0]) * per
    levels, lists = [], []
    for _ in range(n):
        sz = int(blob[off:off + 4].view(np.uint32)[0])
        off += 4
        lv = sz // size_links
        ll = []
        for l in range(lv):
            rec = blob[off + l * size_links: off + (l + 1) * size_links]
            c = int(rec[0:2].copy().view(np.uint16)[0])
            ll.append(rec[4:4 + 4 * c].copy().view(np.uint32))
        off += sz
        levels.append(lv)
        lists.append(ll)
    return np.array(levels), lists, maxlevel, entry, maxm


def test_gpu_build_all_levels(orc, hg):
    """With the row threshold lowered, the levels above 0 are built on the GPU as well (twice
    recursively on this 20k-row set: level >= 1 has ~1,250 elements, level >= 2 ~80, inserted on
    the CPU).  Same element levels as the CPU build; every upper list valid (ids of elements at
    that level, no self links, no repeats, at most M); mean level-1 degree and recall@10 within a
    stated margin of the CPU-built graph's; the GPU search on it equal to the oracle's."""
    from hnsw_rabitq_amd import _lib
    space, data, queries = workload("c2s")
    print(f"gpu build of all levels n={data.shape[0]}", flush=True)
    cpu, _, _, _ = built_index("c2s")
    _lib.call("hg_hnsw_build_gpu_upper_min_rows", 1000)
    try:
        gpu = hg.Index(space, data.shape[1])
        gpu.init_index(data.shape[0], M=16, ef_construction=200, random_seed=100)
        gpu.add_items_gpu(data, num_threads=1)
    finally:
        _lib.call("hg_hnsw_build_gpu_upper_min_rows", 250000)
    _check_graph(gpu)
    lv_c, ll_c, ml_c, _, _ = _upper_lists(cpu)
    lv_g, ll_g, ml_g, entry, maxm = _upper_lists(gpu)
    np.testing.assert_array_equal(lv_g, lv_c)
    assert ml_g == ml_c and lv_g[entry] == ml_g
    for e in range(len(lv_g)):
        for l, row in enumerate(ll_g[e], start=1):
            assert len(row) <= maxm
            assert (row != e).all() and len(np.unique(row)) == len(row)
            assert (lv_g[row] >= l).all()
    deg_c = np.mean([len(ll[0]) for ll in ll_c if ll])
    deg_g = np.mean([len(ll[0]) for ll in ll_g if ll])
    print(f"mean level-1 degree: cpu {deg_c:.2f} gpu {deg_g:.2f}", flush=True)
    assert abs(deg_g - deg_c) <= 0.1 * deg_c
    sp = 0 if space == "l2" else 1
    gt, _ = orc.bf_search(sp, data, queries, 10)
    ef = 40
    cpu.set_ef(ef)
    gpu.set_ef(ef)
    rc = _recall(cpu.knn_query(queries, k=10)[0], gt)
    rg = _recall(gpu.knn_query(queries, k=10)[0], gt)
    print(f"recall@10 at ef={ef}: cpu-built {rc:.4f} gpu-built {rg:.4f}", flush=True)
    assert rg >= rc - 0.01
    o = orc.OracleHnsw(blob=gpu.serialize())
    lt, dt, _, _, _ = o.search(sp, queries, 10, ef, mode=orc.MODE_TOTAL)
    lab, dis = gpu.knn_query(queries, k=10)
    np.testing.assert_array_equal(lab, lt)
    np.testing.assert_array_equal(dis.view(np.uint32), dt.view(np.uint32))
