"""Multi-vector search, searchStopConditionClosest with MultiVectorSearchStopCondition
(hnswalg.h:1326-1377, stop_condition.h:82-200), on the CPU: the oracle's heap-mode restatement
reproduces the reference's results (tests/golden/c1_mv_golden.npz, written by
tests/golden/gen_golden.py mv from the reference compiled from its sources, on an index whose
elements carry their document id), and the total-order restatement agrees with it on tie-free
queries.  The reference's multi-vector graph is the plain L2 graph of the same vectors (the
distance reads the vector part only), so the oracle searches the product's plain build."""
import numpy as np
import pytest

from conftest import golden
from helpers import built_index

CASES = ["d10_e15", "d20_e40", "grid_d3_e5"]


def _split(g, name):
    cnt = g[f"counts_{name}"].astype(np.int64)
    off = np.concatenate([[0], np.cumsum(cnt)])
    return cnt, off


@pytest.mark.parametrize("name", CASES)
def test_mv_oracle_matches_reference(orc, name):
    g = golden("c1_mv_golden.npz")
    p, _, _, queries = built_index("c1")
    o = orc.OracleHnsw(blob=np.frombuffer(p.serialize(), np.uint8))
    nd, efc = (int(x) for x in g[f"params_{name}"])
    docs = g[f"docs_{name}"]  # labels = internal ids in this build
    q = queries[:200]
    lab, dis, cnt, _, _ = o.search_multivector(0, q, docs, nd, efc, 64, mode=orc.MODE_HEAP)
    gc, off = _split(g, name)
    np.testing.assert_array_equal(cnt, gc)
    for i in range(len(gc)):
        np.testing.assert_array_equal(lab[i, :gc[i]], g[f"ids_{name}"][off[i]:off[i + 1]])
        np.testing.assert_array_equal(dis[i, :gc[i]].view(np.uint32),
                                      g[f"dists_{name}"][off[i]:off[i + 1]].view(np.uint32))
    # every row covers exactly num_docs documents when the graph holds enough
    for i in range(len(gc)):
        assert len(np.unique(docs[lab[i, :cnt[i]].astype(np.int64)])) == nd


@pytest.mark.parametrize("name", CASES)
def test_mv_total_order_agrees_on_tie_free_queries(orc, name):
    g = golden("c1_mv_golden.npz")
    p, _, _, queries = built_index("c1")
    o = orc.OracleHnsw(blob=np.frombuffer(p.serialize(), np.uint8))
    nd, efc = (int(x) for x in g[f"params_{name}"])
    docs = g[f"docs_{name}"]
    q = queries[:200]
    lh, dh, ch, _, tie = o.search_multivector(0, q, docs, nd, efc, 64, mode=orc.MODE_HEAP, ties=True)
    lt, dt, ct, _, _ = o.search_multivector(0, q, docs, nd, efc, 64, mode=orc.MODE_TOTAL)
    free = tie == 0
    assert free.sum() >= 20
    np.testing.assert_array_equal(ch[free], ct[free])
    np.testing.assert_array_equal(lh[free], lt[free])
    np.testing.assert_array_equal(dh[free].view(np.uint32), dt[free].view(np.uint32))
