"""Epsilon (range) search, searchStopConditionClosest with EpsilonSearchStopCondition
(hnswalg.h:1326-1377, stop_condition.h:207-275), on the CPU: the oracle's heap-mode restatement
reproduces the reference's results in tests/golden/c1_eps_golden.npz (written by
tests/golden/gen_golden.py eps from the reference compiled from its sources), and the
total-order restatement the GPU follows agrees with it wherever no equal distances met."""
import numpy as np
import pytest

from conftest import golden
from hnsw_rabitq_amd import datasets
from helpers import built_index

CASES = ["eps1", "eps2_max100", "self", "eps1_del_f3"]


def _case(hg, g, name):
    p, _, data, queries = built_index("c1")
    q = data[:300] if name == "self" else queries[:100]
    e, mn, mx = g[f"params_{name}"]
    return p, q, float(e), int(mn), int(mx)


def _split(g, name):
    cnt = g[f"counts_{name}"].astype(np.int64)
    off = np.concatenate([[0], np.cumsum(cnt)])
    return cnt, [g[f"ids_{name}"][off[i]:off[i + 1]] for i in range(len(cnt))], \
        [g[f"dists_{name}"][off[i]:off[i + 1]] for i in range(len(cnt))]


def _oracle(hg, orc, name, g):
    p, q, e, mn, mx = _case(hg, g, name)
    blob = np.frombuffer(p.serialize(), np.uint8)
    allowed = None
    if name.endswith("_del_f3"):
        # a copy of the index with every 10th label deleted (labels = row numbers here)
        import hnsw_rabitq_amd as H
        space, data, _ = "l2", datasets.c1_uniform(), None
        p2 = H.Index("l2", 16)
        p2.init_index(data.shape[0], M=16, ef_construction=200, random_seed=100)
        p2.add_items(data, num_threads=1)
        for lab in g["deleted"]:
            p2.mark_deleted(int(lab))
        blob = np.frombuffer(p2.serialize(), np.uint8)
        o = orc.OracleHnsw(blob=blob)
        allowed = (o.labels()[0] % 3 != 0).astype(np.uint8)
    else:
        o = orc.OracleHnsw(blob=blob)
    return o, q, e, mn, mx, allowed


@pytest.mark.parametrize("name", CASES)
def test_eps_oracle_matches_reference(hg, orc, name):
    g = golden("c1_eps_golden.npz")
    o, q, e, mn, mx, allowed = _oracle(hg, orc, name, g)
    stride = int(g[f"counts_{name}"].max()) + 1
    lab, dis, cnt, _, _ = o.search_epsilon(0, q, e, mn, mx, stride, mode=orc.MODE_HEAP, allowed=allowed)
    gc, gl, gd = _split(g, name)
    np.testing.assert_array_equal(cnt, gc)
    for i in range(len(gc)):
        np.testing.assert_array_equal(lab[i, :gc[i]], gl[i])
        np.testing.assert_array_equal(dis[i, :gc[i]].view(np.uint32), gd[i].view(np.uint32))
    assert (dis[:, 0][cnt > 0] <= e).all()
    if name == "self":
        assert (cnt >= 1).all() and (dis[:, 0] == 0).all()


@pytest.mark.parametrize("name", CASES)
def test_eps_total_order_agrees_on_tie_free_queries(hg, orc, name):
    g = golden("c1_eps_golden.npz")
    o, q, e, mn, mx, allowed = _oracle(hg, orc, name, g)
    stride = int(g[f"counts_{name}"].max()) + 1
    lh, dh, ch, _, tie = o.search_epsilon(0, q, e, mn, mx, stride, mode=orc.MODE_HEAP, allowed=allowed, ties=True)
    lt, dt, ct, _, _ = o.search_epsilon(0, q, e, mn, mx, stride, mode=orc.MODE_TOTAL, allowed=allowed)
    free = tie == 0
    assert free.sum() >= 1
    np.testing.assert_array_equal(ch[free], ct[free])
    np.testing.assert_array_equal(lh[free], lt[free])
    np.testing.assert_array_equal(dh[free].view(np.uint32), dt[free].view(np.uint32))
