"""GPU non-bare-bone search (deleted elements and label filters, searchBaseLayerST
hnswalg.h:323-437) against the CPU oracle and the reference's golden vectors.

Bar: bit-exact ids, distance bits, level-0 expansions and upper-layer counters against the
oracle's (distance, id) total-order restatement of the same branch; bit-exact against the
reference's golden lists (tests/golden/c1_nbb_golden.npz) on every query whose search met no
equal distances.  The deleted set includes the entry point's label.
"""
import numpy as np
import pytest

from conftest import golden
from helpers import workload

pytestmark = pytest.mark.gpu

EF, K = 50, 10
_idx = {}


def _deleted_index(hg):
    if "p" not in _idx:
        g = golden("c1_nbb_golden.npz")
        space, data, queries = workload("c1")
        p = hg.Index(space, data.shape[1])
        p.init_index(data.shape[0], M=16, ef_construction=200, random_seed=100)
        p.add_items(data, num_threads=1)
        for lab in g["deleted"]:
            p.mark_deleted(int(lab))
        p.set_ef(EF)
        _idx["p"] = (g, p, queries, data)
    return _idx["p"]


def _filter_fn(case):
    if case == "del_f3":
        return lambda label: label % 3 != 0
    if case == "del_f50":
        return lambda label: label % 50 == 1
    return None


def _allowed_by_id(o, case):
    lab_by_id, _ = o.labels()
    if case == "del_f3":
        return (lab_by_id % 3 != 0).astype(np.uint8)
    if case == "del_f50":
        return (lab_by_id % 50 == 1).astype(np.uint8)
    return None


def _check_against_oracle(orc, p, o, queries, case, gpu_filter=None):
    allowed = _allowed_by_id(o, case)
    lt, dt, ct, _, short = o.search(0, queries, K, EF, mode=orc.MODE_TOTAL, counters=True, allowed=allowed)
    assert short == 0
    flt = gpu_filter
    if flt is None and allowed is not None:
        flt = p.gpu().make_filter(o.labels()[0][allowed.astype(bool)])
    lab, dis, ctr, gshort = p.gpu().search(queries, K, EF, counters=True, filter=flt)
    assert not gshort
    np.testing.assert_array_equal(lab, lt)
    np.testing.assert_array_equal(dis.view(np.uint32), dt.view(np.uint32))
    np.testing.assert_array_equal(ctr[:, 0], ct[:, 0])
    np.testing.assert_array_equal(ctr[:, 2:], ct[:, 2:])
    assert (ctr[:, 1] >= ct[:, 1]).all()
    assert p.gpu().status()[0] == 0


@pytest.mark.parametrize("case", ["del", "del_f3", "del_f50"])
def test_gpu_nbb_matches_oracle_total_order(orc, hg, case):
    g, p, queries, _ = _deleted_index(hg)
    print(f"case={case} ef={EF} k={K} deleted={len(g['deleted'])}", flush=True)
    o = orc.OracleHnsw(blob=p.serialize())
    _check_against_oracle(orc, p, o, queries, case)


@pytest.mark.parametrize("case", ["del", "del_f3", "del_f50"])
def test_gpu_nbb_matches_reference_golden(orc, hg, case):
    g, p, queries, _ = _deleted_index(hg)
    print(f"case={case} golden", flush=True)
    o = orc.OracleHnsw(blob=p.serialize())
    _, _, _, tie, _ = o.search(0, queries, K, EF, mode=orc.MODE_HEAP, ties=True, allowed=_allowed_by_id(o, case))
    lab, dis = p.knn_query(queries, k=K, filter=_filter_fn(case))
    # the tie flag marks any equal pair among a query's level-0 distances; the 2% filter makes a
    # query visit most of the graph, where some equal pair is almost certain
    clean = tie == 0
    assert clean.sum() >= 10
    np.testing.assert_array_equal(lab[clean], g[f"ids_{case}"][clean])
    np.testing.assert_array_equal(dis[clean].view(np.uint32), g[f"dists_{case}"][clean].view(np.uint32))
    same = (lab == g[f"ids_{case}"]).all(1)
    print(f"{case}: tie-free queries {int(clean.sum())}; rows equal to the reference {same.mean():.4f}", flush=True)
    got = lab.reshape(-1).astype(np.int64)
    assert not np.isin(got, g["deleted"].astype(np.int64)).any()


def test_gpu_filter_of_everything_equals_bare_bone(hg):
    """With nothing deleted and every label allowed, the non-bare-bone branch makes the same
    decisions as the bare-bone one (every admitted key is in the result set): same lists, same
    traversal."""
    space, data, queries = workload("c1")
    p = hg.Index(space, data.shape[1])
    p.init_index(data.shape[0], M=16, ef_construction=200, random_seed=100)
    p.add_items(data, num_threads=1)
    print("case=c1 filter-all", flush=True)
    gi = p.gpu()
    l0, d0, c0, _ = gi.search(queries, K, EF, counters=True)
    f = gi.make_filter(gi.labels())
    assert f.allowed == data.shape[0]
    l1, d1, c1, _ = gi.search(queries, K, EF, counters=True, filter=f)
    np.testing.assert_array_equal(l0, l1)
    np.testing.assert_array_equal(d0.view(np.uint32), d1.view(np.uint32))
    np.testing.assert_array_equal(c0[:, 0], c1[:, 0])
    np.testing.assert_array_equal(c0[:, 2:], c1[:, 2:])


def test_gpu_delete_marks_refresh_without_export(orc, hg):
    """markDelete / unmarkDelete after the export: the GPU copy re-reads the marks."""
    space, data, queries = workload("c1")
    p = hg.Index(space, data.shape[1])
    p.init_index(data.shape[0], M=16, ef_construction=200, random_seed=100)
    p.add_items(data, num_threads=1)
    p.set_ef(EF)
    print("case=c1 refresh", flush=True)
    lab0, dis0 = p.knn_query(queries[:200], k=K)
    first = p.gpu()
    # delete every result of the first query batch's first column, then search again
    victims = np.unique(lab0[:, 0])
    for lab in victims:
        p.mark_deleted(int(lab))
    lab1, _ = p.knn_query(queries[:200], k=K)
    assert p.gpu() is first
    assert not np.isin(lab1, victims).any()
    o = orc.OracleHnsw(blob=p.serialize())
    _check_against_oracle(orc, p, o, queries[:200], "del")
    for lab in victims:
        p.unmark_deleted(int(lab))
    lab2, dis2 = p.knn_query(queries[:200], k=K)
    np.testing.assert_array_equal(lab2, lab0)
    np.testing.assert_array_equal(dis2.view(np.uint32), dis0.view(np.uint32))


def test_gpu_filter_errors(hg):
    g, p, queries, _ = _deleted_index(hg)
    print("case=c1 filter errors", flush=True)
    # nothing allowed: the search visits the whole graph and outgrows the largest beam
    with pytest.raises(RuntimeError, match="beam entries"):
        p.knn_query(queries[:4], k=K, filter=lambda label: False)
    # a filter of another export of the same index is refused, in Python and by the C ABI
    from hnsw_rabitq_amd import _lib
    f = p.make_filter(np.arange(100, dtype=np.uint64))
    other = hg.GpuHnsw(p.cpu_handle())
    with pytest.raises(ValueError, match="another GPU index"):
        other.search(queries[:4], K, EF, filter=f)
    lab = np.empty((4, K), np.uint64)
    dis = np.empty((4, K), np.float32)
    q = np.ascontiguousarray(queries[:4])
    code = _lib.load().hg_gpu_search_filtered(other.handle, q.ctypes.data, 4, K, EF, f.handle, lab.ctypes.data,
                                              dis.ctypes.data, None)
    assert code == _lib.HG_ERR_INVALID
    assert b"another GPU index" in _lib.load().hg_last_error()
    other.free()


def test_gpu_filtered_device_path_per_query(orc, hg):
    """Device-resident filtered search with per-query ef and k in one launch."""
    import ctypes

    from hnsw_rabitq_amd import _lib
    g, p, queries, _ = _deleted_index(hg)
    print("case=c1 device filtered per-query", flush=True)
    o = orc.OracleHnsw(blob=p.serialize())
    allowed = _allowed_by_id(o, "del_f3")
    nq = 300
    q = np.ascontiguousarray(queries[:nq])
    rng = np.random.default_rng(5)
    efq = rng.integers(10, 120, nq).astype(np.int32)
    kq = rng.choice(np.array([1, 5, 10, 20], np.int32), nq)
    kmax = int(kq.max())
    gi = p.gpu()
    f = gi.make_filter(o.labels()[0][allowed.astype(bool)])
    dq = hg.DeviceArray.from_numpy(q)
    de = hg.DeviceArray.from_numpy(efq)
    dk = hg.DeviceArray.from_numpy(kq)
    dl = hg.DeviceArray(nq * kmax * 8)
    dd = hg.DeviceArray(nq * kmax * 4)
    ds = hg.DeviceArray(4)
    _lib.call("hg_gpu_search_device_filtered", gi.handle, dq.ptr, nq, de.ptr, dk.ptr, int(efq.max()), kmax, kmax,
              dl.ptr, dd.ptr, None, f.handle, ds.ptr, None)
    _lib.call("hg_device_synchronize")
    lab = dl.to_numpy(np.uint64, (nq, kmax))
    dis = dd.to_numpy(np.float32, (nq, kmax))
    assert int(ds.to_numpy(np.uint32, (1,))[0]) == 0
    lt, dt, _, _, _ = o.search(0, q, kmax, 0, mode=orc.MODE_TOTAL, ef_per_query=efq, k_per_query=kq,
                               allowed=allowed)
    for r in range(nq):
        kk = int(kq[r])
        np.testing.assert_array_equal(lab[r, :kk], lt[r, :kk])
        np.testing.assert_array_equal(dis[r, :kk].view(np.uint32), dt[r, :kk].view(np.uint32))


def test_gpu_filter_correlated_with_position_retries_large_beam(orc, hg):
    """A filter that keeps 30% of the elements, all far from the queries (first coordinate above
    0.7, queries below 0.2): the beam planned from the global fraction overflows for most queries,
    which the second pass runs again at the largest beam; device and host paths end equal to the
    oracle."""
    from hnsw_rabitq_amd import DeviceArray
    space, data, queries = workload("c1")
    p = hg.Index(space, data.shape[1])
    p.init_index(data.shape[0], M=16, ef_construction=200, random_seed=100)
    p.add_items(data, num_threads=1)
    q = np.ascontiguousarray(queries[queries[:, 0] < 0.2])
    nq = q.shape[0]
    print(f"case=c1 correlated filter nq={nq}", flush=True)
    gi = p.gpu()
    keep = np.nonzero(data[:, 0] > 0.7)[0].astype(np.uint64)  # labels = row numbers
    f = gi.make_filter(keep)
    dq = DeviceArray.from_numpy(q)
    dl, dd, ds = DeviceArray(nq * K * 8), DeviceArray(nq * K * 4), DeviceArray(4)
    gi.search_device_filtered(dq, nq, K, EF, dl, dd, filter=f, d_status=ds)
    assert int(ds.to_numpy(np.uint32, (1,))[0]) == 0
    status, second = gi.status()
    print(f"queries run again at the largest beam: {second} of {nq}", flush=True)
    assert second > nq // 2
    dev_lab = dl.to_numpy(np.uint64, (nq, K))
    lab, dis, ctr, short = gi.search(q, K, EF, counters=True, filter=f)
    assert not short
    assert gi.status() == (0, second)
    np.testing.assert_array_equal(dev_lab, lab)
    o = orc.OracleHnsw(blob=p.serialize())
    lab_by_id = o.labels()[0].astype(np.int64)
    allowed = (data[lab_by_id, 0] > 0.7).astype(np.uint8)
    lt, dt, ct, _, _ = o.search(0, q, K, EF, mode=orc.MODE_TOTAL, counters=True, allowed=allowed)
    np.testing.assert_array_equal(lab, lt)
    np.testing.assert_array_equal(dis.view(np.uint32), dt.view(np.uint32))
    np.testing.assert_array_equal(ctr[:, 0], ct[:, 0])
    assert (keep[np.searchsorted(keep, lab.reshape(-1))] == lab.reshape(-1)).all()
