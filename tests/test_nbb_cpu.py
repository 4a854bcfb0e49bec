"""Deletions and label filters (the non-bare-bone branch of searchBaseLayerST, hnswalg.h:323-437)
on the CPU: the product's delete marks give the reference's saveIndex bytes, and the oracle's
heap-mode restatement reproduces the reference's filtered results (tests/golden/c1_nbb_golden.npz,
written by tests/golden/gen_golden.py nbb from the reference compiled from its sources)."""
import hashlib

import numpy as np
import pytest

from conftest import golden
from helpers import workload


def _deleted_index(hg):
    g = golden("c1_nbb_golden.npz")
    space, data, queries = workload("c1")
    p = hg.Index(space, data.shape[1])
    p.init_index(data.shape[0], M=16, ef_construction=200, random_seed=100)
    p.add_items(data, num_threads=1)
    for lab in g["deleted"]:
        p.mark_deleted(int(lab))
    return g, p, queries


def test_delete_marks_give_reference_bytes(hg):
    g, p, _ = _deleted_index(hg)
    blob = np.frombuffer(p.serialize(), np.uint8)
    assert hashlib.sha256(blob.tobytes()).hexdigest() == str(g["index_sha"])


def test_delete_mark_errors(hg):
    g, p, _ = _deleted_index(hg)
    lab = int(g["deleted"][0])
    with pytest.raises(Exception, match="already deleted"):
        p.mark_deleted(lab)
    p.unmark_deleted(lab)
    with pytest.raises(Exception, match="is not deleted"):
        p.unmark_deleted(lab)
    with pytest.raises(Exception, match="Label not found"):
        p.mark_deleted(10 ** 9)


@pytest.mark.parametrize("case", ["del", "del_f3", "del_f50"])
def test_oracle_nbb_matches_reference(hg, orc, case):
    g, p, queries = _deleted_index(hg)
    o = orc.OracleHnsw(blob=np.frombuffer(p.serialize(), np.uint8))
    lab_by_id, del_by_id = o.labels()
    assert int(del_by_id.sum()) == len(g["deleted"])
    allowed = None
    if case == "del_f3":
        allowed = (lab_by_id % 3 != 0).astype(np.uint8)
    elif case == "del_f50":
        allowed = (lab_by_id % 50 == 1).astype(np.uint8)
    lab, dis, _, _, short = o.search(0, queries, 10, 50, mode=orc.MODE_HEAP, allowed=allowed)
    assert short == 0
    np.testing.assert_array_equal(lab, g[f"ids_{case}"])
    np.testing.assert_array_equal(dis.view(np.uint32), g[f"dists_{case}"].view(np.uint32))
    # nothing deleted or filtered is returned
    got = lab.reshape(-1).astype(np.int64)
    assert not np.isin(got, g["deleted"].astype(np.int64)).any()
    if case == "del_f3":
        assert (got % 3 != 0).all()
    if case == "del_f50":
        assert (got % 50 == 1).all()


def test_oracle_total_mode_agrees_with_heap_mode_on_tie_free_queries(hg, orc):
    """Closed-form cross-check of the total-order restatement the GPU follows: wherever a
    query's search meets no equal distances, both orders make the same decisions."""
    g, p, queries = _deleted_index(hg)
    o = orc.OracleHnsw(blob=np.frombuffer(p.serialize(), np.uint8))
    lab_by_id, _ = o.labels()
    allowed = (lab_by_id % 3 != 0).astype(np.uint8)
    lh, dh, _, th, _ = o.search(0, queries, 10, 50, mode=orc.MODE_HEAP, allowed=allowed, ties=True)
    lt, dt, _, _, _ = o.search(0, queries, 10, 50, mode=orc.MODE_TOTAL, allowed=allowed)
    free = th == 0
    assert free.mean() > 0.5
    np.testing.assert_array_equal(lh[free], lt[free])
    np.testing.assert_array_equal(dh[free], dt[free])
