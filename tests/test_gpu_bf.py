"""GPU brute force: exact FP32 path bit-exact against the reference's golden vectors and the
oracle; reduced-precision (BF16 / FP8 MFMA + exact re-rank) path against the exact path."""
import numpy as np
import pytest

from conftest import golden
from helpers import workload

pytestmark = pytest.mark.gpu


def _bf(hg, space, data, labels=None):
    b = hg.BFIndex(space, data.shape[1])
    b.init_index(data.shape[0])
    b.add_items(data, labels)
    return b


def test_bf_exact_matches_reference_golden(hg):
    g = golden("c1_bf_golden.npz")
    space, data, queries = workload("c1")
    print("case=c1 bf exact k=10", flush=True)
    lab, dis = _bf(hg, space, data).knn_query(queries, k=10)
    np.testing.assert_array_equal(lab, g["ids"])
    np.testing.assert_array_equal(dis.view(np.uint32), g["dists"].view(np.uint32))


@pytest.mark.parametrize("space,dim,n,k", [("l2", 21, 3000, 7), ("ip", 96, 5000, 100), ("l2", 128, 4000, 10),
                                           ("ip", 768, 3000, 100), ("l2", 5, 500, 50)])
def test_bf_exact_matches_oracle(hg, orc, space, dim, n, k):
    print(f"bf exact space={space} dim={dim} n={n} k={k}", flush=True)
    rng = np.random.default_rng(dim)
    data = rng.standard_normal((n, dim)).astype(np.float32)
    q = rng.standard_normal((37, dim)).astype(np.float32)
    lab, dis = _bf(hg, space, data).knn_query(q, k=k)
    ol, od = orc.bf_search(0 if space == "l2" else 1, data, q, k)
    np.testing.assert_array_equal(lab, ol)
    np.testing.assert_array_equal(dis.view(np.uint32), od.view(np.uint32))


def test_bf_exact_ties_by_label(hg, orc):
    """Integer data: many equal distances; ties resolve by the smaller label, with labels not in
    row order."""
    rng = np.random.default_rng(1)
    data = rng.integers(0, 3, (2000, 16)).astype(np.float32)
    labels = rng.permutation(2000).astype(np.uint64) + 1000
    q = rng.integers(0, 3, (20, 16)).astype(np.float32)
    lab, dis = _bf(hg, "l2", data, labels).knn_query(q, k=40)
    ol, od = orc.bf_search(0, data, q, 40, labels=labels)
    np.testing.assert_array_equal(lab, ol)
    np.testing.assert_array_equal(dis, od)


@pytest.mark.parametrize("prec", ["bf16", "fp8"])
@pytest.mark.parametrize("space", ["ip", "l2"])
def test_bf_lowp_matches_exact(hg, space, prec):
    """BF16/FP8 candidates + exact FP32 re-rank: the final lists equal the exact path's."""
    from hnsw_rabitq_amd import datasets, DeviceArray, _lib
    import ctypes
    print(f"bf lowp space={space} prec={prec}", flush=True)
    data, q = datasets.c3_gmm(n=40_000, nq=300, dim=256, n_centres=64)
    if space == "l2":
        data = data * 3.0
        q = q * 3.0
    b = _bf(hg, space, data)
    k = 100
    el, ed = b.knn_query(q, k=k)
    dq = DeviceArray.from_numpy(q)
    dl = DeviceArray(q.shape[0] * k * 8)
    dd = DeviceArray(q.shape[0] * k * 4)
    stats = np.zeros(4, np.float64)
    P = {"bf16": _lib.PREC_BF16, "fp8": _lib.PREC_FP8}[prec]
    _lib.call("hg_bf_search_device", b._h, dq.ptr, q.shape[0], k, P, 2048, dl.ptr, dd.ptr,
              ctypes.c_void_p(stats.ctypes.data), None)
    _lib.call("hg_device_synchronize")
    ll = dl.to_numpy(np.uint64, (q.shape[0], k))
    ld = dd.to_numpy(np.float32, (q.shape[0], k))
    np.testing.assert_array_equal(ll, el)
    np.testing.assert_array_equal(ld.view(np.uint32), ed.view(np.uint32))
    print(f"stats {stats}", flush=True)


@pytest.mark.parametrize("prec", ["bf16", "fp8"])
def test_bf_lowp_ties_at_rerank_cut(hg, prec):
    """Duplicated rows tie in the approximate distance at the re-rank cut.  Queries 0-31 sit on a
    group of 1,500 copies with c = 1,000: the cut falls inside the group and keeps the smallest
    labels (labels are shuffled), so the lists still equal the exact path's.  Queries 32-63 sit
    on a group of 3,000 copies, more ties than the re-rank resolves: those queries take the
    exact path (reported as retried)."""
    from hnsw_rabitq_amd import datasets, DeviceArray, _lib
    import ctypes
    print(f"bf lowp ties prec={prec}", flush=True)
    data, q = datasets.c3_gmm(n=20_000, nq=64, dim=128, n_centres=16)
    rng = np.random.default_rng(5)
    data[:1500] = data[0]
    data[5000:8000] = data[5000]
    q[:32] = data[0] + 1e-3 * rng.standard_normal((32, 128)).astype(np.float32)
    q[32:] = data[5000] + 1e-3 * rng.standard_normal((32, 128)).astype(np.float32)
    labels = (rng.permutation(20_000) + 7).astype(np.uint64)
    b = _bf(hg, "ip", data, labels)
    k = 100
    el, ed = b.knn_query(q, k=k)
    dq = DeviceArray.from_numpy(q)
    dl = DeviceArray(q.shape[0] * k * 8)
    dd = DeviceArray(q.shape[0] * k * 4)
    stats = np.zeros(4, np.float64)
    P = {"bf16": _lib.PREC_BF16, "fp8": _lib.PREC_FP8}[prec]
    _lib.call("hg_bf_search_device", b._h, dq.ptr, q.shape[0], k, P, 1000, dl.ptr, dd.ptr,
              ctypes.c_void_p(stats.ctypes.data), None)
    _lib.call("hg_device_synchronize")
    ll = dl.to_numpy(np.uint64, (q.shape[0], k))
    ld = dd.to_numpy(np.float32, (q.shape[0], k))
    print(f"stats {stats}", flush=True)
    np.testing.assert_array_equal(ll, el)
    np.testing.assert_array_equal(ld.view(np.uint32), ed.view(np.uint32))
    # the first group's lists are its 100 smallest labels
    np.testing.assert_array_equal(ll[0], np.sort(labels[:1500])[:k])
    assert stats[3] == 32


@pytest.mark.parametrize("prec", ["bf16", "fp8"])
@pytest.mark.parametrize("n,c", [(60_000, 512), (6_000, 4096)])
def test_bf_lowp_candidate_list_paths(hg, prec, n, c):
    """The GEMM epilogue keeps each wave's passes in an LDS list and writes them out after the
    reservation; a tile whose list outgrows its space recomputes them instead.  About 1% of the
    rows pass at (60,000, 512) (lists only) and nearly all at (6,000, 4,096) (every tile takes the
    recompute): both give the exact path's lists."""
    from hnsw_rabitq_amd import datasets, DeviceArray, _lib
    import ctypes
    print(f"bf lowp list paths prec={prec} n={n} c={c}", flush=True)
    data, q = datasets.c3_gmm(n=n, nq=200, dim=192, n_centres=32)
    b = _bf(hg, "ip", data)
    k = 20
    el, ed = b.knn_query(q, k=k)
    dq = DeviceArray.from_numpy(q)
    dl = DeviceArray(q.shape[0] * k * 8)
    dd = DeviceArray(q.shape[0] * k * 4)
    stats = np.zeros(4, np.float64)
    P = {"bf16": _lib.PREC_BF16, "fp8": _lib.PREC_FP8}[prec]
    _lib.call("hg_bf_search_device", b._h, dq.ptr, q.shape[0], k, P, c, dl.ptr, dd.ptr,
              ctypes.c_void_p(stats.ctypes.data), None)
    _lib.call("hg_device_synchronize")
    np.testing.assert_array_equal(dl.to_numpy(np.uint64, (q.shape[0], k)), el)
    np.testing.assert_array_equal(dd.to_numpy(np.float32, (q.shape[0], k)).view(np.uint32), ed.view(np.uint32))
    assert stats[3] == 0  # no query fell back to the exact path


@pytest.mark.parametrize("variant", [0, 1, 2, 3, 4])
def test_bf_lowp_gemm_variants_agree(hg, variant):
    """Every GEMM tile configuration gives the exact path's lists (BF16 and FP8, inner product,
    a row count that is not a multiple of any tile and a dim of 320 = five 64-element slices)."""
    from hnsw_rabitq_amd import datasets, DeviceArray, _lib
    import ctypes
    print(f"bf lowp gemm variant={variant}", flush=True)
    data, q = datasets.c3_gmm(n=30_011, nq=517, dim=320, n_centres=32)
    b = _bf(hg, "ip", data)
    k = 50
    el, ed = b.knn_query(q, k=k)
    dq = DeviceArray.from_numpy(q)
    try:
        _lib.call("hg_bf_gemm_variant", variant)
        for P in (_lib.PREC_BF16, _lib.PREC_FP8):
            dl = DeviceArray(q.shape[0] * k * 8)
            dd = DeviceArray(q.shape[0] * k * 4)
            _lib.call("hg_bf_search_device", b._h, dq.ptr, q.shape[0], k, P, 1024, dl.ptr, dd.ptr, None, None)
            _lib.call("hg_device_synchronize")
            np.testing.assert_array_equal(dl.to_numpy(np.uint64, (q.shape[0], k)), el)
            np.testing.assert_array_equal(dd.to_numpy(np.float32, (q.shape[0], k)).view(np.uint32),
                                          ed.view(np.uint32))
    finally:
        _lib.call("hg_bf_gemm_variant", 2)


def test_bf_gemm_autotune(hg):
    """The tile autotuner times every variant, keeps one per index and precision, and the search
    then runs with it: lists equal the exact path's before and after."""
    from hnsw_rabitq_amd import datasets, DeviceArray, _lib
    import ctypes
    data, q = datasets.c3_gmm(n=20_000, nq=300, dim=128, n_centres=16)
    b = _bf(hg, "ip", data)
    k = 20
    el, ed = b.knn_query(q, k=k)
    dq = DeviceArray.from_numpy(q)
    dl, dd = DeviceArray(q.shape[0] * k * 8), DeviceArray(q.shape[0] * k * 4)
    times = np.zeros(5, np.float64)
    chosen = ctypes.c_int(-1)
    _lib.call("hg_bf_autotune", b._h, dq.ptr, q.shape[0], k, _lib.PREC_BF16, 512, dl.ptr, dd.ptr,
              ctypes.c_void_p(times.ctypes.data), ctypes.byref(chosen), None)
    _lib.call("hg_device_synchronize")
    print(f"autotune times {times} chosen {chosen.value}", flush=True)
    assert 0 <= chosen.value < 5 and (times > 0).all()
    assert times[chosen.value] == times.min()
    np.testing.assert_array_equal(dl.to_numpy(np.uint64, (q.shape[0], k)), el)
    _lib.call("hg_bf_search_device", b._h, dq.ptr, q.shape[0], k, _lib.PREC_BF16, 512, dl.ptr, dd.ptr, None, None)
    _lib.call("hg_device_synchronize")
    np.testing.assert_array_equal(dl.to_numpy(np.uint64, (q.shape[0], k)), el)
    np.testing.assert_array_equal(dd.to_numpy(np.float32, (q.shape[0], k)).view(np.uint32), ed.view(np.uint32))
