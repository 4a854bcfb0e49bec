"""GPU range search (searchStopConditionClosest with EpsilonSearchStopCondition,
hnswalg.h:1326-1377, stop_condition.h:207-275) against the oracle and the reference.

Bar: counts, ids and distance bits equal to the oracle's total-order restatement for every
query, and expansions equal to its counters; equal to the reference's results
(tests/golden/c1_eps_golden.npz) on every query whose search met no equal distances.
"""
import numpy as np
import pytest

from conftest import golden
from hnsw_rabitq_amd import datasets
from helpers import built_index

pytestmark = pytest.mark.gpu

CASES = ["eps1", "eps2_max100", "self", "eps1_del_f3"]
_idx = {}


def _index(hg, g, name):
    if not name.endswith("_del_f3"):
        p, _, data, queries = built_index("c1")
        return p, None
    if "del" not in _idx:
        data = datasets.c1_uniform()
        p = hg.Index("l2", 16)
        p.init_index(data.shape[0], M=16, ef_construction=200, random_seed=100)
        p.add_items(data, num_threads=1)
        for lab in g["deleted"]:
            p.mark_deleted(int(lab))
        _idx["del"] = p
    p = _idx["del"]
    labs = p.gpu().labels()
    return p, labs[labs % 3 != 0]


@pytest.mark.parametrize("name", CASES)
def test_gpu_eps_matches_oracle_and_reference(orc, hg, name):
    g = golden("c1_eps_golden.npz")
    p, keep = _index(hg, g, name)
    _, _, data, queries = built_index("c1")
    q = np.ascontiguousarray(data[:300] if name == "self" else queries[:100])
    e, mn, mx = (float(x) for x in g[f"params_{name}"])
    mn, mx = int(mn), int(mx)
    print(f"case={name} eps={e} min={mn} max={mx} nq={q.shape[0]}", flush=True)
    gi = p.gpu()
    flt = gi.make_filter(keep) if keep is not None else None
    stride = 300
    lab, dis, cnt = gi.search_epsilon(q, e, mn, mx, out_stride=stride, filter=flt)
    o = orc.OracleHnsw(blob=p.serialize())
    allowed = None if keep is None else (o.labels()[0] % 3 != 0).astype(np.uint8)
    lt, dt, ct, ctr_t, _ = o.search_epsilon(0, q, e, mn, mx, stride, mode=orc.MODE_TOTAL, allowed=allowed,
                                            counters=True)
    np.testing.assert_array_equal(cnt, ct)
    np.testing.assert_array_equal(lab, lt)
    np.testing.assert_array_equal(dis.view(np.uint32), dt.view(np.uint32))
    # against the reference on tie-free queries
    _, _, _, _, tie = o.search_epsilon(0, q, e, mn, mx, stride, mode=orc.MODE_HEAP, allowed=allowed, ties=True)
    gc = g[f"counts_{name}"].astype(np.int64)
    off = np.concatenate([[0], np.cumsum(gc)])
    free = np.nonzero(tie == 0)[0]
    np.testing.assert_array_equal(cnt, gc)
    for i in free:
        np.testing.assert_array_equal(lab[i, :gc[i]], g[f"ids_{name}"][off[i]:off[i + 1]])
        np.testing.assert_array_equal(dis[i, :gc[i]].view(np.uint32),
                                      g[f"dists_{name}"][off[i]:off[i + 1]].view(np.uint32))
    print(f"{name}: tie-free queries {len(free)} of {q.shape[0]}; results per query {cnt.mean():.1f}", flush=True)
    assert all((dis[i, :cnt[i]] <= e).all() for i in range(q.shape[0]))


def test_gpu_eps_device_counters(orc, hg):
    """Device entry point: counts, rows and expansion counters equal the oracle's."""
    from hnsw_rabitq_amd import DeviceArray, _lib
    p, _, data, queries = built_index("c1")
    q = np.ascontiguousarray(queries[:100])
    nq, stride = q.shape[0], 128
    print("case=c1 eps device", flush=True)
    gi = p.gpu()
    dq = DeviceArray.from_numpy(q)
    dl, dd = DeviceArray(nq * stride * 8), DeviceArray(nq * stride * 4)
    dcnt, dctr, dst = DeviceArray(nq * 4), DeviceArray(nq * 16), DeviceArray(4)
    _lib.call("hg_gpu_search_epsilon_device", gi.handle, dq.ptr, nq, 1.5, 100, 400, None, stride, dl.ptr, dd.ptr,
              dcnt.ptr, dctr.ptr, dst.ptr, None)
    _lib.call("hg_device_synchronize")
    assert int(dst.to_numpy(np.uint32, (1,))[0]) == 0
    cnt = dcnt.to_numpy(np.uint32, (nq,))
    lab = dl.to_numpy(np.uint64, (nq, stride))
    ctr = dctr.to_numpy(np.int32, (nq, 4))
    o = orc.OracleHnsw(blob=p.serialize())
    lt, dt, ct, ctr_t, _ = o.search_epsilon(0, q, 1.5, 100, 400, stride, mode=orc.MODE_TOTAL, counters=True)
    np.testing.assert_array_equal(cnt, ct)
    np.testing.assert_array_equal(lab, lt)
    np.testing.assert_array_equal(ctr[:, 0], ctr_t[:, 0])
    np.testing.assert_array_equal(ctr[:, 2:], ctr_t[:, 2:])
    assert (ctr[:, 1] >= ctr_t[:, 1]).all()
    assert (cnt > stride).any()  # rows keep the closest `stride` of a longer result


def test_gpu_range_query_api(hg):
    p, _, data, queries = built_index("c1")
    print("case=c1 range_query", flush=True)
    res = p.range_query(queries[:5], 1.0, 200, 1000, filter=lambda label: label % 2 == 0)
    assert len(res) == 5
    for lab, dis in res:
        assert (lab % 2 == 0).all() and (dis <= 1.0).all() and (np.diff(dis) >= 0).all()
