"""The non-bare-bone, range, multi-vector and construction paths at an odd dimension in the
cosine space (21: the generic one-lane-per-vector distance form, inner product after
normalisation), against the oracle's total order."""
import numpy as np
import pytest

from helpers import built_index, workload

pytestmark = pytest.mark.gpu


def _setup(hg, orc):
    p, space, data, queries = built_index("cos21")
    q = hg.normalize_rows(queries)
    o = orc.OracleHnsw(blob=p.serialize())
    return p, q, o


def test_odd_dim_filtered_and_deleted(orc, hg):
    p, q, o = _setup(hg, orc)
    print("case=cos21 filtered", flush=True)
    gi = p.gpu()
    labs = gi.labels()
    f = gi.make_filter(labs[labs % 3 != 0])
    lab, dis, ctr, short = gi.search(q, 5, 30, counters=True, filter=f)
    assert not short
    allowed = (o.labels()[0] % 3 != 0).astype(np.uint8)
    lt, dt, ct, _, _ = o.search(1, q, 5, 30, mode=orc.MODE_TOTAL, counters=True, allowed=allowed)
    np.testing.assert_array_equal(lab, lt)
    np.testing.assert_array_equal(dis.view(np.uint32), dt.view(np.uint32))
    np.testing.assert_array_equal(ctr[:, 0], ct[:, 0])


def test_odd_dim_range(orc, hg):
    p, q, o = _setup(hg, orc)
    print("case=cos21 range", flush=True)
    gi = p.gpu()
    lab, dis, cnt = gi.search_epsilon(q, 0.6, 40, 400, out_stride=128)
    lt, dt, ct, _, _ = o.search_epsilon(1, q, 0.6, 40, 400, 128, mode=orc.MODE_TOTAL)
    np.testing.assert_array_equal(cnt, ct)
    np.testing.assert_array_equal(lab, lt)
    np.testing.assert_array_equal(dis.view(np.uint32), dt.view(np.uint32))


def test_odd_dim_multivector(orc, hg):
    p, q, o = _setup(hg, orc)
    print("case=cos21 multivector", flush=True)
    gi = p.gpu()
    labs = gi.labels()
    docs = (labs // 4).astype(np.uint32)
    gd = gi.make_docs(labs, docs)
    lab, dis, cnt = gi.search_multivector(q, gd, 5, 9, out_stride=64)
    lt, dt, ct, _, _ = o.search_multivector(1, q, (o.labels()[0] // 4).astype(np.uint32), 5, 9, 64,
                                            mode=orc.MODE_TOTAL)
    np.testing.assert_array_equal(cnt, ct)
    np.testing.assert_array_equal(lab, lt)
    np.testing.assert_array_equal(dis.view(np.uint32), dt.view(np.uint32))


def test_odd_dim_gpu_build(orc, hg):
    space, data, queries = workload("cos21")
    print("case=cos21 gpu build", flush=True)
    cpu, _, _, _ = built_index("cos21")
    g = hg.Index(space, data.shape[1])
    g.init_index(data.shape[0], M=16, ef_construction=200, random_seed=100)
    g.add_items_gpu(data, num_threads=1)
    q = hg.normalize_rows(queries)
    nd = hg.normalize_rows(data)
    gt, _ = orc.bf_search(1, nd, q, 5)
    cpu.set_ef(30)
    g.set_ef(30)
    rc = float(np.mean([len(np.intersect1d(a, b)) for a, b in zip(cpu.knn_query(queries, k=5)[0], gt)])) / 5
    rg = float(np.mean([len(np.intersect1d(a, b)) for a, b in zip(g.knn_query(queries, k=5)[0], gt)])) / 5
    print(f"recall@5 at ef=30: cpu-built {rc:.4f} gpu-built {rg:.4f}", flush=True)
    assert rg >= rc - 0.02
    o = orc.OracleHnsw(blob=g.serialize())
    lt, dt, _, _, _ = o.search(1, q, 5, 30, mode=orc.MODE_TOTAL)
    lab, dis = g.knn_query(queries, k=5)
    np.testing.assert_array_equal(lab, lt)
    np.testing.assert_array_equal(dis.view(np.uint32), dt.view(np.uint32))
