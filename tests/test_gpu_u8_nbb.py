"""The non-bare-bone searches on the 8-bit rows (integer data, where the export keeps the lossless
byte copy) against the oracle's total order: deleted elements with and without a label filter,
the range search (EpsilonSearchStopCondition) and the multi-vector search
(MultiVectorSearchStopCondition), on an index of its own with every ninth label deleted."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

_idx = {}


def _index(hg):
    if "p" not in _idx:
        from hnsw_rabitq_amd import datasets
        data = datasets.c2_gamma(n=8000)
        queries = datasets.c2_gamma(n=150, seed=1)
        p = hg.Index("l2", 128)
        p.init_index(data.shape[0], M=16, ef_construction=200, random_seed=100)
        p.add_items(data, num_threads=1)
        for lab in range(0, data.shape[0], 9):
            p.mark_deleted(lab)
        _idx["p"] = (p, data, queries)
    return _idx["p"]


@pytest.mark.parametrize("filt", [None, "f3"])
def test_u8_deleted_and_filtered_match_oracle(orc, hg, filt):
    p, data, queries = _index(hg)
    print(f"u8 non-bare-bone: deleted every 9th, filter={filt}", flush=True)
    g = p.gpu()
    assert g.vector_code == "u8"
    o = orc.OracleHnsw(blob=p.serialize())
    lab_by_id, _ = o.labels()
    allowed = flt = None
    if filt == "f3":
        allowed = (lab_by_id % 3 != 0).astype(np.uint8)
        flt = g.make_filter(lab_by_id[allowed.astype(bool)])
    lt, dt, ct, _, short = o.search(0, queries, 10, 80, mode=orc.MODE_TOTAL, counters=True, allowed=allowed)
    assert short == 0
    lab, dis, ctr, gshort = g.search(queries, 10, 80, counters=True, filter=flt)
    assert not gshort
    np.testing.assert_array_equal(lab, lt)
    np.testing.assert_array_equal(dis.view(np.uint32), dt.view(np.uint32))
    np.testing.assert_array_equal(ctr[:, 0], ct[:, 0])
    np.testing.assert_array_equal(ctr[:, 2:], ct[:, 2:])
    assert not np.isin(lab, np.arange(0, data.shape[0], 9, dtype=np.uint64)).any()
    if flt is not None:
        assert (lab % 3 != 0).all()
        flt.free()


def test_u8_range_search_matches_oracle(orc, hg):
    p, data, queries = _index(hg)
    g = p.gpu()
    q = np.ascontiguousarray(queries[:100])
    o = orc.OracleHnsw(blob=p.serialize())
    _, d10, _, _, _ = o.search(0, q, 10, 80, mode=orc.MODE_TOTAL)
    eps = float(np.median(d10[:, 9]))
    print(f"u8 range search: eps={eps} min=20 max=200", flush=True)
    lab, dis, cnt = g.search_epsilon(q, eps, 20, 200, out_stride=200)
    lt, dt, ct, _, _ = o.search_epsilon(0, q, eps, 20, 200, 200, mode=orc.MODE_TOTAL)
    np.testing.assert_array_equal(cnt, ct)
    np.testing.assert_array_equal(lab, lt)
    np.testing.assert_array_equal(dis.view(np.uint32), dt.view(np.uint32))
    assert all((dis[i, :cnt[i]] <= eps).all() for i in range(q.shape[0]))


def test_u8_multivector_matches_oracle(orc, hg):
    p, data, queries = _index(hg)
    g = p.gpu()
    q = np.ascontiguousarray(queries[:100])
    docs = (np.arange(data.shape[0]) % 500).astype(np.uint32)  # per internal id (= label here)
    labs = g.labels()
    gd = g.make_docs(labs, docs[labs.astype(np.int64)])
    print("u8 multi-vector: 500 documents, num_docs=10, ef_collection=40", flush=True)
    lab, dis, cnt = g.search_multivector(q, gd, 10, 40, out_stride=64)
    o = orc.OracleHnsw(blob=p.serialize())
    lt, dt, ct, _, _ = o.search_multivector(0, q, docs, 10, 40, 64, mode=orc.MODE_TOTAL)
    np.testing.assert_array_equal(cnt, ct)
    np.testing.assert_array_equal(lab, lt)
    np.testing.assert_array_equal(dis.view(np.uint32), dt.view(np.uint32))
