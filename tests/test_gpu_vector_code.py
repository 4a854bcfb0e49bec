"""The 8-bit vector rows of the GPU graph search (hg_gpu_set_vector_code).

The export keeps an 8-bit copy of the vectors only when every coordinate is an integer in 0..255
with exactly that FP32 value, and the kernel converts each byte back to the same FP32 value, so
the search must give the same ids, the same distance bits and the same traversal counters as
with the FP32 rows, and both must equal the oracle's total-order search.
"""
import numpy as np
import pytest

from helpers import built_index

pytestmark = pytest.mark.gpu


def small_index(hg, space, data):
    p = hg.Index(space, data.shape[1])
    p.init_index(data.shape[0], M=16, ef_construction=100, random_seed=100)
    p.add_items(data, num_threads=1)
    return p


def integer_data(n, dim, seed):
    rng = np.random.default_rng(seed)
    return np.clip(np.round(rng.gamma(0.5, 40.0, size=(n, dim))), 0, 255).astype(np.float32)


@pytest.mark.parametrize("what,expect", [
    ("integers", "u8"), ("fraction", "f32"), ("above255", "f32"), ("negzero", "f32"), ("negative", "f32"),
    ("dim100", "f32"), ("dim96", "u8"), ("dim16", "u8"), ("uniform", "f32")])
def test_code_only_when_lossless(hg, what, expect):
    dim = {"dim100": 100, "dim96": 96, "dim16": 16}.get(what, 64)
    data = integer_data(600, dim, 7)
    if what == "fraction":
        data[311, 5] += 0.5
    elif what == "above255":
        data[17, 40] = 256.0
    elif what == "negzero":
        data[599, 63] = -0.0
    elif what == "negative":
        data[0, 0] = -1.0
    elif what == "uniform":
        data = np.random.default_rng(1).random((600, dim), dtype=np.float32)
    print(f"vector code case {what} dim={dim}", flush=True)
    p = small_index(hg, "l2", data)  # the Index owns its GPU copy: keep it alive
    g = p.gpu()
    assert g.vector_code == expect
    g.set_vector_code("f32")
    assert g.vector_code == "f32"
    g.set_vector_code("auto")
    assert g.vector_code == expect


def search_both(g, q, k, ef):
    g.set_vector_code("auto")
    a = g.search(q, k, ef, counters=True)
    g.set_vector_code("f32")
    b = g.search(q, k, ef, counters=True)
    g.set_vector_code("auto")
    return a, b


@pytest.mark.parametrize("space,dim,ef,k", [("l2", 128, 40, 10), ("ip", 96, 64, 100), ("l2", 16, 50, 10),
                                            ("l2", 32, 20, 1), ("l2", 64, 80, 10), ("ip", 128, 40, 10)])
def test_u8_rows_equal_f32_rows_and_oracle(hg, orc, space, dim, ef, k):
    print(f"u8 vs f32 rows: space={space} dim={dim} ef={ef} k={k}", flush=True)
    if space == "l2" and dim == 128:
        p, _, data, queries = built_index("c2s")
    else:
        data = integer_data(4000, dim, dim)
        queries = integer_data(100, dim, dim + 1)
        p = small_index(hg, space, data)
    g = p.gpu()
    assert g.vector_code == "u8"
    (la, da, ca, sa), (lb, db, cb, sb) = search_both(g, queries, k, ef)
    assert not sa and not sb
    np.testing.assert_array_equal(la, lb)
    np.testing.assert_array_equal(da.view(np.uint32), db.view(np.uint32))
    np.testing.assert_array_equal(ca[:, [0, 2, 3]], cb[:, [0, 2, 3]])
    o = orc.OracleHnsw(blob=p.serialize())
    lt, dt, ct, _, short = o.search(0 if space == "l2" else 1, queries, k, ef, mode=orc.MODE_TOTAL, counters=True)
    assert short == 0
    np.testing.assert_array_equal(la, lt)
    np.testing.assert_array_equal(da.view(np.uint32), dt.view(np.uint32))
    np.testing.assert_array_equal(ca[:, 0], ct[:, 0])


def test_u8_filtered_search_equals_f32(hg):
    """The non-bare-bone kernel (label filter and deleted elements) on the 8-bit rows."""
    p, _, data, queries = built_index("c2s")
    g = p.gpu()
    assert g.vector_code == "u8"
    allowed = np.arange(0, data.shape[0], 3, dtype=np.uint64)
    f = hg.GpuFilter(g, allowed)
    try:
        g.set_vector_code("auto")
        la, da, ca, _ = g.search(queries, 10, 60, counters=True, filter=f)
        g.set_vector_code("f32")
        lb, db, cb, _ = g.search(queries, 10, 60, counters=True, filter=f)
    finally:
        g.set_vector_code("auto")
        f.free()
    assert np.isin(la[la != np.uint64(~np.uint64(0))], allowed).all()
    np.testing.assert_array_equal(la, lb)
    np.testing.assert_array_equal(da.view(np.uint32), db.view(np.uint32))
    np.testing.assert_array_equal(ca[:, 0], cb[:, 0])


# The plain kernel on 8-bit rows runs its own level-0 loop (lazy merge: new keys pending in
# registers, the main list rewritten every 32 keys; pipelined expansions; exact integer
# distances).  Its results must equal the oracle's at every beam size and k, per-query ef and k,
# and with a hop cap; the FP32 rows (the eager merge) must give the same.
@pytest.mark.parametrize("ef,k", [(1, 1), (5, 10), (64, 1), (100, 10), (300, 50), (700, 100), (2048, 10),
                                  (3000, 10)])
def test_lazy_merge_matches_oracle(hg, orc, ef, k):
    p, _, data, queries = built_index("c2s")
    print(f"lazy merge: c2s ef={ef} k={k}", flush=True)
    g = p.gpu()
    assert g.vector_code == "u8"
    (la, da, ca, sa), (lb, db, cb, sb) = search_both(g, queries, k, ef)
    o = orc.OracleHnsw(blob=p.serialize())
    lt, dt, ct, _, short = o.search(0, queries, k, ef, mode=orc.MODE_TOTAL, counters=True)
    assert sa == sb == (short != 0)
    np.testing.assert_array_equal(la, lt)
    np.testing.assert_array_equal(da.view(np.uint32), dt.view(np.uint32))
    np.testing.assert_array_equal(ca[:, 0], ct[:, 0])
    np.testing.assert_array_equal(ca[:, 2:], ct[:, 2:])
    np.testing.assert_array_equal(la, lb)
    np.testing.assert_array_equal(ca[:, 0], cb[:, 0])


def test_lazy_merge_fractional_queries(hg, orc):
    """Queries that are not integers take the FP32 form of the distances on the same rows."""
    p, _, data, queries = built_index("c2s")
    q = queries + np.float32(0.25)
    g = p.gpu()
    la, da, ca, _ = g.search(q, 10, 80, counters=True)
    o = orc.OracleHnsw(blob=p.serialize())
    lt, dt, ct, _, short = o.search(0, q, 10, 80, mode=orc.MODE_TOTAL, counters=True)
    assert short == 0
    np.testing.assert_array_equal(la, lt)
    np.testing.assert_array_equal(da.view(np.uint32), dt.view(np.uint32))
    np.testing.assert_array_equal(ca[:, 0], ct[:, 0])


def test_lazy_merge_per_query_ef_k(hg, orc):
    from hnsw_rabitq_amd import DeviceArray, datasets
    p, _, data, queries = built_index("c2s")
    nq = queries.shape[0]
    ef, k = datasets.c4_ef_k(nq, seed=11, ef_lo=1, ef_hi=800, ks=(1, 10, 100))
    kmax = int(k.max())
    print(f"lazy merge uneven: ef in [{ef.min()},{ef.max()}]", flush=True)
    g = p.gpu()
    dq = DeviceArray.from_numpy(queries)
    def_, dk = DeviceArray.from_numpy(ef), DeviceArray.from_numpy(k)
    dl, dd = DeviceArray(nq * kmax * 8), DeviceArray(nq * kmax * 4)
    g.search_device(dq, nq, kmax, int(ef.max()), dl, dd, d_ef=def_, d_k=dk, k_stride=kmax)
    lab = dl.to_numpy(np.uint64, (nq, kmax))
    dis = dd.to_numpy(np.float32, (nq, kmax))
    o = orc.OracleHnsw(blob=p.serialize())
    ol, od, _, _, _ = o.search(0, queries, kmax, int(ef.max()), mode=orc.MODE_TOTAL, ef_per_query=ef,
                               k_per_query=k)
    for i in range(nq):
        np.testing.assert_array_equal(lab[i, :k[i]], ol[i, :k[i]])
        np.testing.assert_array_equal(dis[i, :k[i]].view(np.uint32), od[i, :k[i]].view(np.uint32))


def test_lazy_merge_hop_cap(hg, orc):
    from hnsw_rabitq_amd import DeviceArray, _lib
    p, _, data, queries = built_index("c2s")
    nq, k, ef = queries.shape[0], 10, 400
    caps = np.random.default_rng(4).integers(-5, 450, nq).astype(np.int32)
    g = p.gpu()
    dq, dcap = DeviceArray.from_numpy(queries), DeviceArray.from_numpy(caps)
    dl, dd, dc = DeviceArray(nq * k * 8), DeviceArray(nq * k * 4), DeviceArray(nq * 16)
    _lib.call("hg_gpu_search_device_hopcap", g.handle, dq.ptr, nq, dcap.ptr, ef, k, dl.ptr, dd.ptr, dc.ptr, None)
    _lib.call("hg_device_synchronize")
    lab, dis = dl.to_numpy(np.uint64, (nq, k)), dd.to_numpy(np.float32, (nq, k))
    ctr = dc.to_numpy(np.int32, (nq, 4))
    o = orc.OracleHnsw(blob=p.serialize())
    ol, od, oc, _ = o.search_hopcap(0, queries, k, ef, caps, mode=orc.MODE_TOTAL, counters=True)
    np.testing.assert_array_equal(lab, ol)
    np.testing.assert_array_equal(dis.view(np.uint32), od.view(np.uint32))
    np.testing.assert_array_equal(ctr[:, 0], oc[:, 0])
