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
    rep = b.lowp_report()
    print(f"stats {stats} report {rep}", flush=True)
    np.testing.assert_array_equal(ll, el)
    np.testing.assert_array_equal(ld.view(np.uint32), ed.view(np.uint32))
    # the first group's lists are its 100 smallest labels
    np.testing.assert_array_equal(ll[0], np.sort(labels[:1500])[:k])
    # the second group's ties outnumber what the re-rank resolves; the first group's copies left
    # out at the cut have the k-th distance itself, which the bound cannot rule out
    assert rep["exact_overflow"] == 32
    assert rep["exact_overflow"] + rep["exact_uncertified"] == stats[3]


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
    rep = b.lowp_report()
    print(f"report {rep}", flush=True)
    assert rep["exact_overflow"] == 0  # no candidate list overflowed


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
    times = np.zeros(_lib.GEMM_VARIANTS, np.float64)
    chosen = ctypes.c_int(-1)
    _lib.call("hg_bf_autotune", b._h, dq.ptr, q.shape[0], k, _lib.PREC_BF16, 512, dl.ptr, dd.ptr,
              ctypes.c_void_p(times.ctypes.data), ctypes.byref(chosen), None)
    _lib.call("hg_device_synchronize")
    print(f"autotune times {times} chosen {chosen.value}", flush=True)
    assert 0 <= chosen.value < _lib.GEMM_VARIANTS and (times > 0).all()
    assert times[chosen.value] == times.min()
    np.testing.assert_array_equal(dl.to_numpy(np.uint64, (q.shape[0], k)), el)
    _lib.call("hg_bf_search_device", b._h, dq.ptr, q.shape[0], k, _lib.PREC_BF16, 512, dl.ptr, dd.ptr, None, None)
    _lib.call("hg_device_synchronize")
    np.testing.assert_array_equal(dl.to_numpy(np.uint64, (q.shape[0], k)), el)
    np.testing.assert_array_equal(dd.to_numpy(np.float32, (q.shape[0], k)).view(np.uint32), ed.view(np.uint32))


def _lowp_device(b, q, k, prec, c):
    from hnsw_rabitq_amd import DeviceArray, _lib
    import ctypes
    dq = DeviceArray.from_numpy(q)
    dl = DeviceArray(q.shape[0] * k * 8)
    dd = DeviceArray(q.shape[0] * k * 4)
    stats = np.zeros(4, np.float64)
    P = {"bf16": _lib.PREC_BF16, "fp8": _lib.PREC_FP8}[prec]
    _lib.call("hg_bf_search_device", b._h, dq.ptr, q.shape[0], k, P, c, dl.ptr, dd.ptr,
              ctypes.c_void_p(stats.ctypes.data), None)
    _lib.call("hg_device_synchronize")
    return dq, dl, dd, dl.to_numpy(np.uint64, (q.shape[0], k)), dd.to_numpy(np.float32, (q.shape[0], k)), stats


@pytest.mark.parametrize("prec", ["bf16", "fp8"])
def test_bf_lowp_c3_workload_matches_oracle(hg, orc, prec):
    """The 1M x 768 mixed-precision workload's generator at 60,000 rows (1,024 centres,
    sigma 0.35, unit norm, inner product), 200 queries, c = 4,096, k = 100: BF16 / FP8 GEMM, FP32
    re-rank and certification give the oracle's exact lists (bruteforce.h:106-125) bit for bit.
    Queries the bound cannot certify (FP8's bound is wider than the gap between the 100th and the
    4,096th distance at this row count) take the exact path; the report counts them."""
    from hnsw_rabitq_amd import datasets
    print(f"case=c3 reduced 60000x768 prec={prec}", flush=True)
    data, q = datasets.c3_gmm(n=60_000, nq=200, dim=768, n_centres=1024)
    b = _bf(hg, "ip", data)
    k = 100
    _, _, _, ll, ld, stats = _lowp_device(b, q, k, prec, 4096)
    rep = b.lowp_report()
    print(f"stats {stats} report {rep}", flush=True)
    ol, od = orc.bf_search(1, data, q, k)
    np.testing.assert_array_equal(ll, ol)
    np.testing.assert_array_equal(ld.view(np.uint32), od.view(np.uint32))
    assert rep["certified"] == 1 and rep["exact_overflow"] == 0
    assert rep["exact_uncertified"] == stats[3]
    if prec == "bf16":
        assert rep["exact_uncertified"] == 0  # the BF16 bound (~3e-3) is far below the gap
        # bound pruning: the exact pass reads only the rows within twice the bound of the 101st
        # approximate distance, a small part of the 4,096 candidates
        assert rep["reranked_rows"] < 0.25 * 4096 * q.shape[0]
    # every certified query re-ranks at least its k + 1 best candidates
    assert rep["reranked_rows"] >= (k + 1) * q.shape[0]
    # the bound is an upper bound on the approximate error: never below the BF16 / FP8
    # quantisation error of a unit vector
    assert rep["max_bound"] >= (2.0 ** -9 if prec == "bf16" else 2.0 ** -5)
    # the host-buffer entry point (BFIndex.knn_query at this precision) gives the same lists
    b.set_precision(prec, 4096)
    hl, hd = b.knn_query(q, k=k)
    np.testing.assert_array_equal(hl, ol)
    np.testing.assert_array_equal(hd.view(np.uint32), od.view(np.uint32))


def test_bf_lowp_uncertified_queries_take_exact_path(hg, orc):
    """c = k: the re-rank cut sits at the k-th candidate, so the bound cannot certify a list and
    every query is answered by the exact path -- still the oracle's lists."""
    from hnsw_rabitq_amd import datasets
    data, q = datasets.c3_gmm(n=20_000, nq=50, dim=128, n_centres=16)
    b = _bf(hg, "l2", data * 2.0)
    _, _, _, ll, ld, stats = _lowp_device(b, q * 2.0, 10, "fp8", 10)
    rep = b.lowp_report()
    print(f"stats {stats} report {rep}", flush=True)
    ol, od = orc.bf_search(0, data * 2.0, q * 2.0, 10)
    np.testing.assert_array_equal(ll, ol)
    np.testing.assert_array_equal(ld.view(np.uint32), od.view(np.uint32))
    assert rep["exact_uncertified"] + rep["exact_overflow"] == 50 and stats[3] == 50


def _near_tie_data(space, dim, seed=11):
    """Rows in twins that differ by one FP32 ulp in one coordinate; queries next to twins: the
    twins' FP32 distances differ by ~1e-8 relative or tie, and their FP64 order can disagree."""
    from hnsw_rabitq_amd import datasets
    data, _ = datasets.c3_gmm(n=8_000, nq=1, dim=dim, n_centres=8, seed=seed)
    rng = np.random.default_rng(seed)
    twins = data[:4000].copy()
    j = rng.integers(0, dim, 4000)
    up = rng.random(4000) < 0.5
    twins[np.arange(4000), j] = np.where(up, np.nextafter(twins[np.arange(4000), j], np.float32(np.inf)),
                                         np.nextafter(twins[np.arange(4000), j], np.float32(-np.inf)))
    data[4000:] = twins
    qi = rng.integers(0, 4000, 120)
    q = (data[qi] + 0.02 * rng.standard_normal((120, dim))).astype(np.float32)
    if space == "ip":
        q /= np.linalg.norm(q, axis=1, keepdims=True)
    return data, q


def _expected_near_ties(orc, sp, data, q, lab, dis, rel):
    """(near-tie pairs, FP64 order disagreements, max |d32 - d64|) of closest-first lists whose
    labels are row numbers, from the oracle's FP64 distances."""
    nq, k = lab.shape
    qi = np.repeat(np.arange(nq), k)
    d64 = orc.fp64_pairs(sp, q, data, qi, lab.reshape(-1).astype(np.int64)).reshape(nq, k)
    a, b = dis[:, :-1].astype(np.float64), dis[:, 1:].astype(np.float64)
    near = (b - a) <= rel * np.abs(a)
    dis64 = (d64[:, :-1] > d64[:, 1:]) | ((d64[:, :-1] == d64[:, 1:]) & (lab[:, :-1] > lab[:, 1:]))
    return int(near.sum()), int((near & dis64).sum()), float(np.abs(dis.astype(np.float64) - d64).max())


@pytest.mark.parametrize("space", ["l2", "ip"])
def test_fp64_check_on_planted_near_ties(hg, orc, space):
    """hg_bf_fp64_check against the oracle's FP64 distances (or_fp64_pairs) on lists full of
    planted near ties: the same near-tie pairs, the same FP64 order disagreements, the same max
    error; labels differ from rows (a shuffled label table) to exercise the label -> row map."""
    from hnsw_rabitq_amd import DeviceArray, _lib
    import ctypes
    sp = 0 if space == "l2" else 1
    data, q = _near_tie_data(space, 64)
    rng = np.random.default_rng(3)
    labels = (rng.permutation(data.shape[0]) * 3 + 5).astype(np.uint64)
    b = _bf(hg, space, data, labels)
    k = 16
    lab, dis = b.knn_query(q, k=k)
    rows = np.argsort(labels)[np.searchsorted(np.sort(labels), lab)]
    want = _expected_near_ties(orc, sp, data, q, rows, dis, 1e-6)
    print(f"expected near ties {want}", flush=True)
    assert want[0] > 50, "the twins must produce near ties"
    dq, dl, dd = DeviceArray.from_numpy(q), DeviceArray.from_numpy(lab), DeviceArray.from_numpy(dis)
    out = np.zeros(4, np.float64)
    _lib.call("hg_bf_fp64_check", b._h, dq.ptr, q.shape[0], k, dl.ptr, dd.ptr, ctypes.c_double(1e-6),
              ctypes.c_void_p(out.ctypes.data), None)
    print(f"fp64 check {out}", flush=True)
    assert int(out[0]) == want[0] and int(out[1]) == want[1]
    assert out[2] == pytest.approx(want[2], rel=1e-9, abs=1e-15)
    # a label the index does not hold is an error
    bad = lab.copy()
    bad[0, 0] = 4
    dbad = DeviceArray.from_numpy(bad)
    with pytest.raises(RuntimeError, match="not in the index"):
        _lib.call("hg_bf_fp64_check", b._h, dq.ptr, q.shape[0], k, dbad.ptr, dd.ptr, ctypes.c_double(1e-6),
                  ctypes.c_void_p(out.ctypes.data), None)


def test_rerank_fp64_near_tie_check(hg, orc):
    """The FP64 check inside the re-rank (BF16 path) counts the near-tie pairs of the returned
    lists and their FP64 order disagreements as the oracle's FP64 distances do (pairs inside the
    list; the pair at the k-th boundary is counted apart)."""
    data, q = _near_tie_data("ip", 128, seed=12)
    b = _bf(hg, "ip", data)
    k = 16
    _, _, _, ll, ld, _ = _lowp_device(b, q, k, "bf16", 256)
    rep = b.lowp_report()
    print(f"report {rep}", flush=True)
    ol, od = orc.bf_search(1, data, q, k)
    np.testing.assert_array_equal(ll, ol)
    assert rep["exact_overflow"] == 0 and rep["exact_uncertified"] == 0
    want = _expected_near_ties(orc, 1, data, q, ll.astype(np.int64), ld, 1e-6)
    assert want[0] > 50
    assert rep["near_ties"] - rep["boundary_near_ties"] == want[0]
    assert rep["near_tie_disagree"] - rep["boundary_disagree"] == want[1]
    # the pair at the k-th boundary is (k-th, (k+1)-th) of the exact order: the bound-pruned
    # re-rank keeps the (k+1)-th row, so the boundary counts are the oracle's k + 1 list's
    ol1, od1 = orc.bf_search(1, data, q, k + 1)
    want1 = _expected_near_ties(orc, 1, data, q, ol1.astype(np.int64), od1, 1e-6)
    assert rep["boundary_near_ties"] == want1[0] - want[0]
    assert rep["boundary_disagree"] == want1[1] - want[1]


def test_bf_add_over_capacity_keeps_placed_rows(hg, orc):
    """addPoint row by row (bruteforce.h:64-83): rows placed before the index filled up stay
    added; re-adding and deleting those labels afterwards works on the same rows."""
    rng = np.random.default_rng(4)
    data = rng.standard_normal((130, 16)).astype(np.float32)
    b = hg.BFIndex("l2", 16)
    b.init_index(100)
    b.add_items(data[:90], ids=np.arange(90))
    with pytest.raises(RuntimeError, match="exceeds the specified limit"):
        b.add_items(data[90:120], ids=np.arange(90, 120))
    assert b.get_current_count() == 100
    # re-add labels 90..99 (now existing: rows overwritten in place) and delete two of them
    b.add_items(data[120:130], ids=np.arange(90, 100))
    b.delete_vector(95)
    b.delete_vector(3)
    assert b.get_current_count() == 98
    # the reference's rows after the same operations
    rows = {i: data[i] for i in range(90)}
    rows.update({90 + i: data[120 + i] for i in range(10)})
    del rows[95], rows[3]
    labs = np.array(sorted(rows), np.uint64)
    mat = np.stack([rows[int(l)] for l in labs])
    q = rng.standard_normal((20, 16)).astype(np.float32)
    lab, dis = b.knn_query(q, k=7)
    ol, od = orc.bf_search(0, mat, q, 7, labels=labs)
    np.testing.assert_array_equal(lab, ol)
    np.testing.assert_array_equal(dis.view(np.uint32), od.view(np.uint32))
