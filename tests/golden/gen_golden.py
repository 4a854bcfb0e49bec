"""Writes the golden fixtures of tests/golden/ from the reference compiled from its own
sources (oracle/_ref: the reference Python module built from python_bindings/bindings.cpp and
the hnswlib headers with setup.py's flags).  Run in a container where /root/reference exists:

    make -C oracle ref && python tests/golden/gen_golden.py        (all fixtures)
                          python tests/golden/gen_golden.py nbb    (the deletion/filter fixture only)
                          python tests/golden/gen_golden.py bf_ops (the brute-force operations fixture)
                          python tests/golden/gen_golden.py eps    (the epsilon-search fixture)
                          python tests/golden/gen_golden.py mv     (the multi-vector fixture)

Every fixture stores its inputs' SHA-256 (or the inputs themselves when small), the reference's
outputs and the SHA-256 of the reference's saveIndex bytes for single-threaded builds.
"""
import hashlib
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from hnsw_rabitq_amd import datasets  # noqa: E402
from oracle import pyoracle  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def build_ref(hnswlib, space, data, M=16, efc=200, seed=100):
    p = hnswlib.Index(space=space, dim=data.shape[1])
    p.init_index(max_elements=data.shape[0], M=M, ef_construction=efc, random_seed=seed)
    p.add_items(data, num_threads=1)
    path = "/tmp/_golden_index.bin"
    p.save_index(path)
    with open(path, "rb") as f:
        blob = f.read()
    os.remove(path)
    return p, hashlib.sha256(blob).hexdigest()


def hnsw_case(hnswlib, name, space, data, queries, efs, ks):
    p, ghash = build_ref(hnswlib, space, data)
    out = dict(data_sha=sha(data), query_sha=sha(queries), index_sha=ghash, efs=np.array(efs), ks=np.array(ks))
    for ef in efs:
        for k in ks:
            p.set_ef(ef)
            lab, dis = p.knn_query(queries, k=k, num_threads=1)
            out[f"ids_ef{ef}_k{k}"] = lab.astype(np.uint64)
            out[f"dists_ef{ef}_k{k}"] = dis.astype(np.float32)
    np.savez_compressed(os.path.join(OUT, name), **out)
    print(name, ghash[:16])


def main():
    hnswlib = pyoracle.ref_module()
    shim = pyoracle.ref_shim()
    assert hnswlib is not None and shim is not None, "build the reference first: make -C oracle ref"

    # F2: distance bits at every dimension class (SIMD16, SIMD4, residual forms, scalar).
    rng = np.random.default_rng(123)
    dims = list(range(1, 41)) + [48, 64, 96, 100, 128, 161, 768]
    kat = {"dims": np.array(dims)}
    for d in dims:
        a = rng.standard_normal((16, d)).astype(np.float32)
        b = rng.standard_normal((16, d)).astype(np.float32)
        for sp, sname in ((0, "l2"), (1, "ip")):
            out = np.empty(16, np.float32)
            shim.ref_distance_batch(sp, d, a.ctypes.data, b.ctypes.data, 16, out.ctypes.data)
            kat[f"{sname}_{d}"] = out
        kat[f"a_{d}"] = a
        kat[f"b_{d}"] = b
    np.savez_compressed(os.path.join(OUT, "dist_kat.npz"), **kat)

    # F4: level sequence of the insertion RNG for random_seed=100, M=16.
    lv = np.empty(100_000, np.int32)
    shim.ref_random_levels(16, 100, lv.shape[0], lv.ctypes.data)
    np.savez_compressed(os.path.join(OUT, "levels_seed100.npz"), levels=lv)

    # F1: C1 (10k x 16 uniform, L2), HNSW ef=50 k=10 and BF k=10.
    data = datasets.c1_uniform()
    queries = datasets.c1_uniform(n=1000, seed=1)
    hnsw_case(hnswlib, "c1_golden.npz", "l2", data, queries, [50], [10])
    bf = hnswlib.BFIndex(space="l2", dim=16)
    bf.init_index(max_elements=data.shape[0])
    bf.add_items(data)
    lab, dis = bf.knn_query(queries, k=10, num_threads=1)
    np.savez_compressed(os.path.join(OUT, "c1_bf_golden.npz"), ids=lab.astype(np.uint64), dists=dis,
                        data_sha=sha(data), query_sha=sha(queries))

    # F3: reduced C2 (integer-valued Gamma data, many distance ties) and reduced C4 (IP GMM).
    data = datasets.c2_gamma(n=20_000)
    queries = datasets.c2_gamma(n=200, seed=1)
    hnsw_case(hnswlib, "c2s_golden.npz", "l2", data, queries, [10, 40], [10])
    data, queries = datasets.c4_gmm(n=20_000, nq=200, n_centres=64)
    hnsw_case(hnswlib, "c4s_golden.npz", "ip", data, queries, [16, 64], [1, 10, 100])

    # Cosine space at an odd dimension (normalisation + residual kernels).
    rng = np.random.default_rng(7)
    data = rng.standard_normal((2000, 21)).astype(np.float32)
    queries = rng.standard_normal((100, 21)).astype(np.float32)
    hnsw_case(hnswlib, "cos21_golden.npz", "cosine", data, queries, [30], [5])
    np.savez_compressed(os.path.join(OUT, "cos21_inputs.npz"), data=data, queries=queries)


def nbb():
    """F5: the non-bare-bone branch on C1: deletions (every 5th label and the entry point's),
    then a label filter on top (label % 3 != 0), then a 2%-selective filter (label % 50 == 1);
    ef=50, k=10.  Stores the reference's saveIndex hash after the deletions."""
    hnswlib = pyoracle.ref_module()
    assert hnswlib is not None, "build the reference first: make -C oracle ref"
    data = datasets.c1_uniform()
    queries = datasets.c1_uniform(n=1000, seed=1)
    p, _ = build_ref(hnswlib, "l2", data)
    path = "/tmp/_golden_index.bin"
    p.save_index(path)
    entry = pyoracle.OracleHnsw(path=path)
    entry_label = int(pyoracle.lib().or_hnsw_entry(entry.h))  # labels = insertion order
    deleted = sorted(set(range(0, data.shape[0], 5)) | {entry_label})
    for lab in deleted:
        p.mark_deleted(lab)
    p.save_index(path)
    with open(path, "rb") as f:
        blob = f.read()
    os.remove(path)
    out = dict(data_sha=sha(data), query_sha=sha(queries), deleted=np.array(deleted, np.uint64),
               entry_label=entry_label, index_sha=hashlib.sha256(blob).hexdigest())
    p.set_ef(50)
    lab, dis = p.knn_query(queries, k=10, num_threads=1)
    out["ids_del"], out["dists_del"] = lab.astype(np.uint64), dis.astype(np.float32)
    lab, dis = p.knn_query(queries, k=10, num_threads=1, filter=lambda l: l % 3 != 0)
    out["ids_del_f3"], out["dists_del_f3"] = lab.astype(np.uint64), dis.astype(np.float32)
    lab, dis = p.knn_query(queries, k=10, num_threads=1, filter=lambda l: l % 50 == 1)
    out["ids_del_f50"], out["dists_del_f50"] = lab.astype(np.uint64), dis.astype(np.float32)
    np.savez_compressed(os.path.join(OUT, "c1_nbb_golden.npz"), **out)
    print("c1_nbb_golden.npz", out["index_sha"][:16], "entry label", entry_label)


def bf_ops():
    """F6: BFIndex deletions, replacement, appends, a label filter and saveIndex on C1.  Stores
    the reference's results and the SHA-256 of the saved file's header and used records (records
    past the count are uninitialised memory in the reference's file)."""
    hnswlib = pyoracle.ref_module()
    assert hnswlib is not None, "build the reference first: make -C oracle ref"
    data = datasets.c1_uniform()
    queries = datasets.c1_uniform(n=1000, seed=1)
    n = data.shape[0]
    bf = hnswlib.BFIndex(space="l2", dim=16)
    bf.init_index(max_elements=n + 100)
    bf.add_items(data)
    deleted = list(range(0, n, 7))
    for lab in deleted:
        bf.delete_vector(lab)
    bf.delete_vector(10 ** 7)  # unknown label: ignored
    extra = datasets.c1_uniform(n=60, seed=2)
    bf.add_items(extra[:50], ids=np.arange(20000, 20050))
    bf.add_items(extra[50:55], ids=np.array([3, 5, 8, 9, 10]))  # existing labels: rows overwritten
    bf.add_items(extra[55:60])  # default labels continue the counter of rows passed in
    out = dict(data_sha=sha(data), query_sha=sha(queries), deleted=np.array(deleted, np.uint64))
    lab, dis = bf.knn_query(queries, k=10, num_threads=1)
    out["ids_ops"], out["dists_ops"] = lab.astype(np.uint64), dis.astype(np.float32)
    lab, dis = bf.knn_query(queries, k=10, num_threads=1, filter=lambda l: l % 3 != 0)
    out["ids_ops_f3"], out["dists_ops_f3"] = lab.astype(np.uint64), dis.astype(np.float32)
    path = "/tmp/_golden_bf.bin"
    bf.save_index(path)
    with open(path, "rb") as f:
        blob = f.read()
    os.remove(path)
    hdr = np.frombuffer(blob[:24], np.uint64)
    used = 24 + int(hdr[2]) * int(hdr[1])
    out["file_header"] = hdr.copy()
    out["file_used_sha"] = hashlib.sha256(blob[:used]).hexdigest()
    np.savez_compressed(os.path.join(OUT, "c1_bf_ops_golden.npz"), **out)
    print("c1_bf_ops_golden.npz", out["file_used_sha"][:16], "count", int(hdr[2]))


def eps():
    """F7: searchStopConditionClosest with EpsilonSearchStopCondition on C1 (the reference's
    epsilon_search_test.cpp setting: squared radius 1.0, min 2000, max = all elements), a
    radius 2.0 search capped at 100 results, a filtered one (label % 3 != 0, 10% of labels
    deleted) and self-queries at radius 1e-4.  Results are stored flat with per-query counts."""
    shim = pyoracle.ref_shim()
    assert shim is not None, "build the reference first: make -C oracle ref"
    data = datasets.c1_uniform()
    queries = datasets.c1_uniform(n=1000, seed=1)[:100]
    n = data.shape[0]
    h = shim.ref_hnsw_build(0, 16, data.ctypes.data, n, 16, 200, 100)
    out = dict(data_sha=sha(data), query_sha=sha(queries))

    def run(name, q, e, mn, mx, allowed=None):
        q = np.ascontiguousarray(q, np.float32)
        stride = mx
        lab = np.empty((q.shape[0], stride), np.uint64)
        dis = np.empty((q.shape[0], stride), np.float32)
        cnt = np.empty(q.shape[0], np.uint32)
        al = None if allowed is None else np.ascontiguousarray(allowed, np.uint8)
        r = shim.ref_hnsw_search_epsilon(h, q.ctypes.data, q.shape[0], e, mn, mx,
                                         None if al is None else al.ctypes.data, n, stride, lab.ctypes.data,
                                         dis.ctypes.data, cnt.ctypes.data)
        assert r == 0, shim.ref_last_error()
        out[f"counts_{name}"] = cnt
        out[f"ids_{name}"] = np.concatenate([lab[i, :cnt[i]] for i in range(q.shape[0])])
        out[f"dists_{name}"] = np.concatenate([dis[i, :cnt[i]] for i in range(q.shape[0])])
        out[f"params_{name}"] = np.array([e, mn, mx], np.float64)
        print(name, "results per query", float(cnt.mean()), "max", int(cnt.max()))

    run("eps1", queries, 1.0, 2000, n)
    run("eps2_max100", queries, 2.0, 50, 100)
    run("self", data[:300], 1e-4, 500, n)
    for lab in range(0, n, 10):
        assert shim.ref_hnsw_mark_deleted(h, lab) == 0
    allowed = (np.arange(n) % 3 != 0).astype(np.uint8)  # labels = row numbers
    run("eps1_del_f3", queries, 1.0, 300, 2000, allowed)
    out["deleted"] = np.arange(0, n, 10, dtype=np.uint64)
    shim.ref_hnsw_free(h)
    np.savez_compressed(os.path.join(OUT, "c1_eps_golden.npz"), **out)


def mv():
    """F8: searchStopConditionClosest with MultiVectorSearchStopCondition on C1 vectors, each
    element in one of 500 documents (document = label % 500, 20 vectors each), num_docs 10 with
    ef_collection 15 (the reference's multivector_search_test.cpp values), and num_docs 20 with
    ef_collection 40; 200 queries.  Results flat with per-query counts."""
    shim = pyoracle.ref_shim()
    assert shim is not None, "build the reference first: make -C oracle ref"
    data = datasets.c1_uniform()
    queries = datasets.c1_uniform(n=1000, seed=1)[:200]
    n = data.shape[0]
    out = dict(data_sha=sha(data), query_sha=sha(queries))
    # scattered documents (label % 500) and spatially coherent ones (an 8 x 8 grid of the first
    # two coordinates: ~156 nearby vectors per document)
    scattered = (np.arange(n) % 500).astype(np.uint32)
    grid = (np.floor(data[:, 0] * 8).astype(np.uint32) * 8 + np.floor(data[:, 1] * 8).astype(np.uint32))
    for name, docs, nd, efc in (("d10_e15", scattered, 10, 15), ("d20_e40", scattered, 20, 40),
                                ("grid_d3_e5", grid.astype(np.uint32), 3, 5)):
        docs = np.ascontiguousarray(docs)
        h = shim.ref_mv_build(16, data.ctypes.data, docs.ctypes.data, n, 16, 200, 100)
        assert h, shim.ref_last_error()
        out[f"docs_{name}"] = docs
        stride = 2000
        lab = np.empty((queries.shape[0], stride), np.uint64)
        dis = np.empty((queries.shape[0], stride), np.float32)
        cnt = np.empty(queries.shape[0], np.uint32)
        q = np.ascontiguousarray(queries)
        assert shim.ref_mv_search(h, q.ctypes.data, q.shape[0], nd, efc, stride, lab.ctypes.data, dis.ctypes.data,
                                  cnt.ctypes.data) == 0, shim.ref_last_error()
        assert cnt.max() < stride
        out[f"counts_{name}"] = cnt
        out[f"ids_{name}"] = np.concatenate([lab[i, :cnt[i]] for i in range(q.shape[0])])
        out[f"dists_{name}"] = np.concatenate([dis[i, :cnt[i]] for i in range(q.shape[0])])
        out[f"params_{name}"] = np.array([nd, efc], np.int64)
        print(name, "results per query", float(cnt.mean()), "max", int(cnt.max()))
        shim.ref_mv_free(h)
    np.savez_compressed(os.path.join(OUT, "c1_mv_golden.npz"), **out)


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "mv":
        mv()
        sys.exit(0)
    if len(sys.argv) > 1 and sys.argv[1] == "eps":
        eps()
        sys.exit(0)
    if len(sys.argv) > 1 and sys.argv[1] == "nbb":
        nbb()
    elif len(sys.argv) > 1 and sys.argv[1] == "bf_ops":
        bf_ops()
    else:
        main()
        nbb()
        bf_ops()
        eps()
        mv()
