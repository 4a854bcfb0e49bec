"""GPU multi-vector search (searchStopConditionClosest with MultiVectorSearchStopCondition,
hnswalg.h:1326-1377, stop_condition.h:82-200) against the oracle and the reference.

Bar: counts, ids and distance bits equal to the oracle's total-order restatement on every query
(expansions too), and equal to the reference's results (tests/golden/c1_mv_golden.npz) on every
query whose search met no equal distances.
"""
import numpy as np
import pytest

from conftest import golden
from helpers import built_index

pytestmark = pytest.mark.gpu

CASES = ["d10_e15", "d20_e40", "grid_d3_e5"]


@pytest.mark.parametrize("name", CASES)
def test_gpu_mv_matches_oracle_and_reference(orc, hg, name):
    g = golden("c1_mv_golden.npz")
    p, _, data, queries = built_index("c1")
    q = np.ascontiguousarray(queries[:200])
    nd, efc = (int(x) for x in g[f"params_{name}"])
    docs = g[f"docs_{name}"]
    print(f"case={name} num_docs={nd} ef_collection={efc} nq={q.shape[0]}", flush=True)
    gi = p.gpu()
    labs = gi.labels()
    gd = gi.make_docs(labs, docs[labs.astype(np.int64)])
    lab, dis, cnt = gi.search_multivector(q, gd, nd, efc, out_stride=64)
    o = orc.OracleHnsw(blob=p.serialize())
    lt, dt, ct, _, _ = o.search_multivector(0, q, docs, nd, efc, 64, mode=orc.MODE_TOTAL)
    np.testing.assert_array_equal(cnt, ct)
    np.testing.assert_array_equal(lab, lt)
    np.testing.assert_array_equal(dis.view(np.uint32), dt.view(np.uint32))
    _, _, _, _, tie = o.search_multivector(0, q, docs, nd, efc, 64, mode=orc.MODE_HEAP, ties=True)
    gc = g[f"counts_{name}"].astype(np.int64)
    off = np.concatenate([[0], np.cumsum(gc)])
    free = np.nonzero(tie == 0)[0]
    for i in free:
        assert cnt[i] == gc[i]
        np.testing.assert_array_equal(lab[i, :gc[i]], g[f"ids_{name}"][off[i]:off[i + 1]])
        np.testing.assert_array_equal(dis[i, :gc[i]].view(np.uint32),
                                      g[f"dists_{name}"][off[i]:off[i + 1]].view(np.uint32))
    print(f"{name}: tie-free queries {len(free)} of {q.shape[0]}; results per query {cnt.mean():.2f}", flush=True)


def test_gpu_mv_device_counters_filter_and_deletions(orc, hg):
    """Device entry point with a label filter on an index with deleted elements: rows, counts
    and expansion counters equal the oracle's."""
    from hnsw_rabitq_amd import DeviceArray, _lib, datasets
    data = datasets.c1_uniform()
    q = np.ascontiguousarray(datasets.c1_uniform(n=1000, seed=1)[:150])
    p = hg.Index("l2", 16)
    p.init_index(data.shape[0], M=16, ef_construction=200, random_seed=100)
    p.add_items(data, num_threads=1)
    for lab in range(0, data.shape[0], 9):
        p.mark_deleted(lab)
    print("case=c1 mv device filter deletions", flush=True)
    gi = p.gpu()
    labs = gi.labels()
    docs = (labs % 250).astype(np.uint32)
    gd = gi.make_docs(labs, docs)
    flt = gi.make_filter(labs[labs % 4 != 0])
    nq, stride = q.shape[0], 64
    dq = DeviceArray.from_numpy(q)
    dl, dd = DeviceArray(nq * stride * 8), DeviceArray(nq * stride * 4)
    dcnt, dctr, dst = DeviceArray(nq * 4), DeviceArray(nq * 16), DeviceArray(4)
    _lib.call("hg_gpu_search_multivector_device", gi.handle, dq.ptr, nq, gd.handle, 8, 12, flt.handle, stride,
              dl.ptr, dd.ptr, dcnt.ptr, dctr.ptr, dst.ptr, 0, None)
    _lib.call("hg_device_synchronize")
    assert int(dst.to_numpy(np.uint32, (1,))[0]) == 0
    o = orc.OracleHnsw(blob=p.serialize())
    lab_by_id = o.labels()[0]
    allowed = (lab_by_id % 4 != 0).astype(np.uint8)
    lt, dt, ct, ctr_t, _ = o.search_multivector(0, q, (lab_by_id % 250).astype(np.uint32), 8, 12, stride,
                                                mode=orc.MODE_TOTAL, allowed=allowed, counters=True)
    np.testing.assert_array_equal(dcnt.to_numpy(np.uint32, (nq,)), ct)
    np.testing.assert_array_equal(dl.to_numpy(np.uint64, (nq, stride)), lt)
    np.testing.assert_array_equal(dd.to_numpy(np.float32, (nq, stride)).view(np.uint32), dt.view(np.uint32))
    ctr = dctr.to_numpy(np.int32, (nq, 4))
    np.testing.assert_array_equal(ctr[:, 0], ctr_t[:, 0])
    np.testing.assert_array_equal(ctr[:, 2:], ctr_t[:, 2:])
    got = dl.to_numpy(np.uint64, (nq, stride))
    for i in range(nq):
        row = got[i, :ct[i]].astype(np.int64)
        assert (row % 4 != 0).all() and (row % 9 != 0).all()
        assert len(np.unique(row % 250)) == 8


def test_gpu_multivector_query_api(hg):
    p, _, data, queries = built_index("c1")
    print("case=c1 multivector_query", flush=True)
    docs = {i: i // 7 for i in range(data.shape[0])}
    res = p.multivector_query(queries[:5], docs, num_docs=4, ef_collection=8)
    assert len(res) == 5
    for lab, dis in res:
        assert len({docs[int(x)] for x in lab}) == 4 and (np.diff(dis) >= 0).all()
