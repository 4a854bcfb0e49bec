"""Brute-force index operations: removePoint (bruteforce.h:85-104), addPoint of new and existing
labels (bruteforce.h:64-83), the filtered searchKnn (bruteforce.h:106-125) and saveIndex
(bruteforce.h:128-141), against the reference's results in tests/golden/c1_bf_ops_golden.npz
(written by tests/golden/gen_golden.py bf_ops from the reference compiled from its sources).

The CPU test replays the operations on a row table and checks the oracle's selection and the
file layout against the reference; the GPU tests run the same operations through the C ABI.
"""
import hashlib

import numpy as np
import pytest

from conftest import golden
from hnsw_rabitq_amd import datasets


def _ops():
    data = datasets.c1_uniform()
    queries = datasets.c1_uniform(n=1000, seed=1)
    extra = datasets.c1_uniform(n=60, seed=2)
    return data, queries, extra


def _replay_rows():
    """Row table after the fixture's operations, following removePoint's move of the last row."""
    data, _, extra = _ops()
    n = data.shape[0]
    rows = [data[i] for i in range(n)]
    labs = list(range(n))
    where = {i: i for i in range(n)}
    for lab in range(0, n, 7):
        r = where.pop(lab)
        last = len(rows) - 1
        if r != last:
            rows[r], labs[r] = rows[last], labs[last]
            where[labs[r]] = r
        rows.pop()
        labs.pop()

    def add(vecs, ids):
        for v, l in zip(vecs, ids):
            if l in where:
                rows[where[l]] = v
            else:
                where[l] = len(rows)
                rows.append(v)
                labs.append(l)
    add(extra[:50], range(20000, 20050))
    add(extra[50:55], [3, 5, 8, 9, 10])
    add(extra[55:60], range(n + 55, n + 60))  # default labels: rows passed in so far
    return np.array(rows, np.float32), np.array(labs, np.uint64)


def test_bf_ops_oracle_matches_reference(orc):
    g = golden("c1_bf_ops_golden.npz")
    data, queries, _ = _ops()
    rows, labs = _replay_rows()
    assert rows.shape[0] == int(g["file_header"][2])
    lab, dis = orc.bf_search(0, rows, queries, 10, labels=labs)
    np.testing.assert_array_equal(lab, g["ids_ops"])
    np.testing.assert_array_equal(dis.view(np.uint32), g["dists_ops"].view(np.uint32))
    keep = labs % 3 != 0
    lab, dis = orc.bf_search(0, rows[keep], queries, 10, labels=labs[keep])
    np.testing.assert_array_equal(lab, g["ids_ops_f3"])
    np.testing.assert_array_equal(dis.view(np.uint32), g["dists_ops_f3"].view(np.uint32))
    # saveIndex layout: three size_t, then [16 f32 | u64 label] records in row order
    rec = np.zeros(rows.shape[0], dtype=[("v", np.float32, 16), ("l", np.uint64)])
    rec["v"], rec["l"] = rows, labs
    blob = np.array([data.shape[0] + 100, 72, rows.shape[0]], np.uint64).tobytes() + rec.tobytes()
    assert hashlib.sha256(blob).hexdigest() == str(g["file_used_sha"])


def _gpu_bf(hg):
    data, queries, extra = _ops()
    n = data.shape[0]
    bf = hg.BFIndex("l2", 16)
    bf.init_index(n + 100)
    bf.add_items(data)
    for lab in range(0, n, 7):
        bf.delete_vector(lab)
    bf.delete_vector(10 ** 7)
    bf.add_items(extra[:50], ids=np.arange(20000, 20050))
    bf.add_items(extra[50:55], ids=np.array([3, 5, 8, 9, 10]))
    bf.add_items(extra[55:60])
    return bf, queries


@pytest.mark.gpu
def test_gpu_bf_ops_match_reference(hg):
    g = golden("c1_bf_ops_golden.npz")
    bf, queries = _gpu_bf(hg)
    print("case=c1 bf ops", flush=True)
    assert bf.get_current_count() == int(g["file_header"][2])
    lab, dis = bf.knn_query(queries, k=10)
    np.testing.assert_array_equal(lab, g["ids_ops"])
    np.testing.assert_array_equal(dis.view(np.uint32), g["dists_ops"].view(np.uint32))
    lab, dis = bf.knn_query(queries, k=10, filter=lambda l: l % 3 != 0)
    np.testing.assert_array_equal(lab, g["ids_ops_f3"])
    np.testing.assert_array_equal(dis.view(np.uint32), g["dists_ops_f3"].view(np.uint32))


@pytest.mark.gpu
def test_gpu_bf_save_load(hg, tmp_path):
    g = golden("c1_bf_ops_golden.npz")
    bf, queries = _gpu_bf(hg)
    print("case=c1 bf save/load", flush=True)
    path = str(tmp_path / "bf.bin")
    bf.save_index(path)
    with open(path, "rb") as f:
        blob = f.read()
    hdr = np.frombuffer(blob[:24], np.uint64)
    np.testing.assert_array_equal(hdr, g["file_header"])
    assert len(blob) == 24 + int(hdr[0]) * int(hdr[1])
    assert hashlib.sha256(blob[:24 + int(hdr[2]) * int(hdr[1])]).hexdigest() == str(g["file_used_sha"])
    b2 = hg.BFIndex("l2", 16)
    b2.load_index(path)
    lab, dis = b2.knn_query(queries, k=10)
    np.testing.assert_array_equal(lab, g["ids_ops"])
    np.testing.assert_array_equal(dis.view(np.uint32), g["dists_ops"].view(np.uint32))
    # the loaded index knows its labels: deleting one removes it from every result
    victim = int(lab[0, 0])
    b2.delete_vector(victim)
    lab2, _ = b2.knn_query(queries, k=10)
    assert not (lab2 == victim).any()


@pytest.mark.gpu
def test_gpu_bf_filter_errors(hg):
    bf, queries = _gpu_bf(hg)
    print("case=c1 bf filter errors", flush=True)
    with pytest.raises(RuntimeError, match="not enough elements"):
        bf.knn_query(queries[:3], k=10, filter=lambda l: l in (1, 2, 4))
    lab, _ = bf.knn_query(queries[:3], k=3, filter=lambda l: l in (1, 2, 4))
    assert set(lab.reshape(-1).tolist()) == {1, 2, 4}
