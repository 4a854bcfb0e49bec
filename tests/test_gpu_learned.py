"""Learned early termination: the HIP training kernels' loss and gradients against the CPU
oracle (FP64), the in-search network against the forward kernel, and the stop hook's fallback
(a network that never fires leaves the plain search bit-identical)."""
import ctypes

import numpy as np
import pytest

from helpers import built_index

pytestmark = pytest.mark.gpu


def bf16_round(p):
    u = np.ascontiguousarray(p, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def effective(params, I, H):
    """Weights as the kernels use them: BF16-rounded matrices, FP32 biases."""
    p = params.copy()
    o = 0
    for size, is_w in ((H * I, True), (H, False), (H * H, True), (H, False), (H, True), (1, False)):
        if is_w:
            p[o:o + size] = bf16_round(p[o:o + size])
        o += size
    return p


def test_mlp_forward_and_gradients_match_oracle(orc):
    from hnsw_rabitq_amd import DeviceArray, learned_stop
    print("mlp gradient check", flush=True)
    I, H, n = 12, 64, 512
    rng = np.random.default_rng(0)
    m = learned_stop.StopModel()
    m.init_params(3)
    params = m.get_params()
    x = rng.standard_normal((n, I)).astype(np.float32)
    y = (rng.random(n) < 0.4).astype(np.float32)
    dx, dy = DeviceArray.from_numpy(x), DeviceArray.from_numpy(y)
    dlog = DeviceArray(n * 4)
    m.forward(dx, n, dlog)
    logits = dlog.to_numpy(np.float32, (n,))
    eff = effective(params, I, H)
    ref = orc.mlp_forward(eff, I, H, x)
    np.testing.assert_allclose(logits, ref, rtol=1e-4, atol=1e-4)
    g = DeviceArray(m.nparam * 4)
    loss = m.train_step(dx, dy, n, lr=0.0, dgrad=g)
    grad = g.to_numpy(np.float32, (m.nparam,))
    rloss, rgrad = orc.mlp_loss_grad(eff, I, H, x, y)
    assert abs(loss - rloss) < 1e-4 * max(1.0, abs(rloss))
    # relative to the gradient's scale (FP32 ordered sums vs FP64)
    scale = np.abs(rgrad).max()
    assert np.abs(grad - rgrad).max() < 1e-4 * scale
    # a wrong-sign or dropped term would be far outside: check directionally too
    cos = float(grad @ rgrad / (np.linalg.norm(grad) * np.linalg.norm(rgrad)))
    assert cos > 0.99999
    m.free()


def test_adam_step_decreases_loss():
    from hnsw_rabitq_amd import DeviceArray, learned_stop
    rng = np.random.default_rng(1)
    n = 4096
    x = rng.standard_normal((n, 12)).astype(np.float32)
    y = (x[:, 0] + 0.5 * x[:, 3] > 0).astype(np.float32)
    m = learned_stop.StopModel()
    m.init_params(0)
    dx, dy = DeviceArray.from_numpy(x), DeviceArray.from_numpy(y)
    first = m.train_step(dx, dy, n, lr=3e-3)
    for _ in range(200):
        last = m.train_step(dx, dy, n, lr=3e-3)
    assert last < 0.5 * first
    m.free()


@pytest.mark.parametrize("case,ef", [("c1", 50), ("c2s", 300)])
def test_stop_model_that_never_fires_is_plain_search(hg, case, ef):
    """FP32 rows (c1) and 8-bit rows (c2s, whose learned kernel merges lazily and merges its
    pending keys at every evaluation)."""
    from hnsw_rabitq_amd import DeviceArray, learned_stop
    p, space, data, queries = built_index(case)
    gpu = p.gpu()
    assert gpu.vector_code == ("u8" if case == "c2s" else "f32")
    nq, k = queries.shape[0], 10
    dq = DeviceArray.from_numpy(queries)
    l0, d0 = DeviceArray(nq * k * 8), DeviceArray(nq * k * 4)
    l1, d1 = DeviceArray(nq * k * 8), DeviceArray(nq * k * 4)
    gpu.search_device(dq, nq, k, ef, l0, d0)
    m = learned_stop.StopModel()
    params = np.zeros(m.nparam, np.float32)
    params[-1] = -30.0  # logit -30: confidence ~1e-13, never >= 0.9
    m.set_params(params)
    m.config(period=4, threshold=0.9, min_hops=4)
    gpu.search_device(dq, nq, k, ef, l1, d1, stop_model=m.h)
    np.testing.assert_array_equal(l0.to_numpy(np.uint64, (nq, k)), l1.to_numpy(np.uint64, (nq, k)))
    np.testing.assert_array_equal(d0.to_numpy(np.float32, (nq, k)), d1.to_numpy(np.float32, (nq, k)))
    # a network that always fires stops every query at its first check: fewer expansions
    params[-1] = 30.0
    m.set_params(params)
    c0, c1 = DeviceArray(nq * 16), DeviceArray(nq * 16)
    gpu.search_device(dq, nq, k, ef, l0, d0, d_counters=c0)
    gpu.search_device(dq, nq, k, ef, l1, d1, d_counters=c1, stop_model=m.h)
    h0 = c0.to_numpy(np.int32, (nq, 4))[:, 0]
    h1 = c1.to_numpy(np.int32, (nq, 4))[:, 0]
    assert (h1 <= np.maximum(h0, 4)).all() and (h1 <= 4).all()
    m.free()


def test_traces_record_features_and_topk(hg):
    from hnsw_rabitq_amd import DeviceArray, learned_stop
    p, space, data, queries = built_index("c1")
    gpu = p.gpu()
    nq, k, ef = 200, 10, 50
    dq = DeviceArray.from_numpy(queries[:nq])
    f, t, r, fin = learned_stop.collect_traces(gpu, dq, nq, ef, k, period=4, max_records=64)
    lab, dis, ctr, _ = gpu.search(queries[:nq], k, ef, counters=True)
    # the final ids are the search's own result (labels == ids here), as sets
    np.testing.assert_array_equal(np.sort(fin.astype(np.uint64), 1), np.sort(lab, 1))
    # one record every 4 expansions
    np.testing.assert_array_equal(r, np.minimum(ctr[:, 0] // 4, 64))
    assert np.isfinite(f[np.arange(64)[None, :] < r[:, None]]).all()
    # feature 0 is the query norm
    np.testing.assert_allclose(f[:, 0, 0], np.linalg.norm(queries[:nq], axis=1), rtol=1e-5)


def _trained_model(gpu, ef, k, period, seed=7, n_train=4000, train_queries=None):
    """A network trained (GPU forward/backward/Adam) on traces of the plain search over training
    queries (by default of the reduced 1M x 128 generator: 20,000 x 128 integer Gamma, separate
    seed)."""
    from hnsw_rabitq_amd import DeviceArray, datasets, learned_stop
    tq = train_queries if train_queries is not None else datasets.c2_gamma(n=n_train, seed=seed)
    n_train = tq.shape[0]
    dtq = DeviceArray.from_numpy(tq)
    f, t, r, fin = learned_stop.collect_traces(gpu, dtq, n_train, ef, k, period=period, max_records=128)
    eq, valid = learned_stop.label_records(t, r, fin)
    X, Y = f[valid], eq[valid].astype(np.float32)
    mean, std = X.mean(0), X.std(0) + 1e-6
    m = learned_stop.StopModel()
    m.init_params(0)
    m.set_norm(mean, 1.0 / std)
    learned_stop.train(m, (X - mean) / std, Y, steps=400, batch=2048, lr=2e-3)
    return m, mean.astype(np.float32), (1.0 / std).astype(np.float32), float(Y.mean())


@pytest.mark.parametrize("case", ["c2s", "c4s"])  # 8-bit rows (L2, dim 128); FP32 rows (IP, dim 96)
def test_learned_stop_matches_hop_capped_oracle(hg, orc, case):
    """The learned search's rows equal the oracle's total-order search stopped after the same
    number of level-0 expansions (the stop-condition slot, hnswlib.h:134-150, hnswalg.h:343-355),
    and the in-kernel network's stop decisions equal those of the oracle's FP64 network on the
    same traversal features (queries whose logit comes within 0.01 of the threshold excepted: the
    kernel's FP32 arithmetic differs from the oracle's FP64 by ~1e-5 there)."""
    from hnsw_rabitq_amd import DeviceArray, datasets, learned_stop
    p, space, data, queries = built_index(case)
    gpu = p.gpu()
    nq, k, ef, period, thr = queries.shape[0], 10, 160, 8, 0.9
    print(f"case={case} learned stop ef={ef} period={period} rows={gpu.vector_code}", flush=True)
    tq = None
    if case == "c4s":  # training queries from the same mixture (64 centres, seed 0), another seed
        tq = datasets.gmm_points(datasets.gmm_centres(64, 96, 0), 4000, 0.5, 7)
    m, mean, inv_std, pos = _trained_model(gpu, ef, k, period, train_queries=tq)
    m.config(period=period, threshold=thr, min_hops=period)
    dq = DeviceArray.from_numpy(queries)
    dl, dd, dc = DeviceArray(nq * k * 8), DeviceArray(nq * k * 4), DeviceArray(nq * 16)
    gpu.search_device(dq, nq, k, ef, dl, dd, d_counters=dc)
    plain_h0 = dc.to_numpy(np.int32, (nq, 4))[:, 0]
    gpu.search_device(dq, nq, k, ef, dl, dd, d_counters=dc, stop_model=m.h)
    lab = dl.to_numpy(np.uint64, (nq, k))
    dis = dd.to_numpy(np.float32, (nq, k))
    ctr = dc.to_numpy(np.int32, (nq, 4))
    stopped = ctr[:, 0] < plain_h0
    print(f"positive records {pos:.3f}; stopped early {stopped.sum()} of {nq}; mean hops "
          f"{ctr[:, 0].mean():.1f} vs {plain_h0.mean():.1f}", flush=True)
    assert stopped.sum() >= nq // 10, "the trained network should stop a share of the queries early"
    o = orc.OracleHnsw(blob=p.serialize())
    ol, od, oc, short = o.search_hopcap(0 if space == "l2" else 1, queries, k, ef, ctr[:, 0], mode=orc.MODE_TOTAL,
                                        counters=True)
    assert short == 0
    np.testing.assert_array_equal(lab, ol)
    np.testing.assert_array_equal(dis.view(np.uint32), od.view(np.uint32))
    np.testing.assert_array_equal(ctr[:, 0], oc[:, 0])
    np.testing.assert_array_equal(ctr[:, 2:], oc[:, 2:])
    # stop decisions: the oracle network (FP64, the kernel's BF16-rounded weights) on the traced
    # features, normalised as the kernel does (FP32)
    f, _, r, _ = learned_stop.collect_traces(gpu, dq, nq, ef, k, period=period, max_records=512)
    assert (r < 512).all()
    fn = ((f - mean[None, None, :]) * inv_std[None, None, :]).astype(np.float32)
    eff = effective(m.get_params(), 12, 64)
    logits = orc.mlp_forward(eff, 12, 64, fn.reshape(-1, 12)).reshape(nq, 512).astype(np.float64)
    lthr = np.log(thr / (1 - thr))
    checked = 0
    for q in range(nq):
        fire = np.nonzero((logits[q, :r[q]] >= lthr))[0]
        want = (fire[0] + 1) * period if fire.size else plain_h0[q]
        upto = fire[0] + 1 if fire.size else r[q]
        if np.abs(logits[q, :upto] - lthr).min(initial=1.0) < 0.01:
            continue
        checked += 1
        assert ctr[q, 0] == want, f"query {q}: kernel stopped at {ctr[q, 0]}, oracle network at {want}"
    print(f"stop decisions checked against the oracle network: {checked} of {nq}", flush=True)
    assert checked >= nq * 0.9
    m.free()


def test_learned_extension_matches_oracle(hg, orc):
    """The extended learned search (base ef 40, extension to ef 160): a query is answered by the
    search at ef 40 stopped where the network fired, or -- when at that search's own end the
    network is not confident its top-10 has converged -- by the search at ef 160 (stopped where
    the network fired there).  Every row equals the oracle's total-order search of the path the
    oracle network's decisions give, hop-capped at the GPU's expansion count."""
    from hnsw_rabitq_amd import DeviceArray, _lib, datasets, learned_stop
    p, space, data, queries = built_index("c2s")
    gpu = p.gpu()
    nq, k, ef, ext, period, thr, thr_ext = queries.shape[0], 10, 40, 160, 8, 0.95, 0.6
    print(f"case=c2s learned extension ef={ef} ext={ext}", flush=True)
    tq = datasets.c2_gamma(n=4000, seed=7)
    X, Y = learned_stop.extension_records(gpu, DeviceArray.from_numpy(tq), 4000, ef, ext, k, period)
    m, mean, inv_std, _ = learned_stop.train_model(X, Y, steps=400, batch=2048)
    m.config(period=period, threshold=thr, min_hops=period)
    m.config_extend(ext, thr_ext)
    dq = DeviceArray.from_numpy(queries)
    dl, dd, dc = DeviceArray(nq * k * 8), DeviceArray(nq * k * 4), DeviceArray(nq * 16)
    gpu.search_device(dq, nq, k, ef, dl, dd, d_counters=dc, stop_model=m.h)
    lab = dl.to_numpy(np.uint64, (nq, k))
    dis = dd.to_numpy(np.float32, (nq, k))
    h0 = dc.to_numpy(np.int32, (nq, 4))[:, 0]
    st, n_ext = ctypes.c_uint32(), ctypes.c_uint32()
    _lib.call("hg_gpu_search_status", gpu.handle, ctypes.byref(st), ctypes.byref(n_ext))
    print(f"extended {n_ext.value} of {nq}; mean hops {h0.mean():.1f}", flush=True)
    assert 0 < n_ext.value < nq
    # the oracle network's decisions on both searches' traces
    eff = effective(m.get_params(), 12, 64)
    lthr, lext = np.log(thr / (1 - thr)), np.log(thr_ext / (1 - thr_ext))

    def decisions(e, end):
        f, _, r, _ = learned_stop.collect_traces(gpu, dq, nq, e, k, period=period, max_records=512, end_record=end)
        fn = ((f - mean[None, None, :]) * inv_std[None, None, :]).astype(np.float32)
        lg = orc.mlp_forward(eff, 12, 64, fn.reshape(-1, 12)).reshape(nq, 512).astype(np.float64)
        return lg, r

    lb, rb = decisions(ef, True)   # periodic records, then the end record
    le, re = decisions(ext, False)
    o = orc.OracleHnsw(blob=p.serialize())
    base = o.search_hopcap(0, queries, k, ef, h0, mode=orc.MODE_TOTAL, counters=True)
    extd = o.search_hopcap(0, queries, k, ext, h0, mode=orc.MODE_TOTAL, counters=True)
    checked = predicted_ext = 0
    for q in range(nq):
        per = lb[q, :rb[q] - 1]  # periodic records of the base search
        fire = np.nonzero(per >= lthr)[0]
        near = np.abs(per[:fire[0] + 1 if fire.size else per.size] - lthr).min(initial=1.0) < 0.01
        if fire.size:
            path, cap = base, (fire[0] + 1) * period
        else:
            end = lb[q, rb[q] - 1]
            near |= abs(end - lext) < 0.01
            if end >= lext:
                path, cap = base, None
            else:
                predicted_ext += 1
                pe = le[q, :re[q]]
                fe = np.nonzero(pe >= lthr)[0]
                near |= np.abs(pe[:fe[0] + 1 if fe.size else pe.size] - lthr).min(initial=1.0) < 0.01
                path, cap = extd, ((fe[0] + 1) * period if fe.size else None)
        if near:
            continue
        checked += 1
        if cap is not None:
            assert h0[q] == cap, (q, h0[q], cap)
        np.testing.assert_array_equal(lab[q], path[0][q])
        np.testing.assert_array_equal(dis[q].view(np.uint32), path[1][q].view(np.uint32))
        assert h0[q] == path[2][q, 0]
    print(f"checked {checked} of {nq}; extended by the oracle network {predicted_ext}", flush=True)
    assert checked >= nq * 0.85
    assert abs(predicted_ext - n_ext.value) <= nq - checked
    m.free()


def test_learned_stop_gain_over_plain(hg, orc):
    """What the trained network gains over the method without it: on the reduced 1M x 128
    workload (20,000 x 128 integer Gamma), the learned search at ef=96 (confidence 0.9) against
    the plain fixed-ef search's recall-vs-expansions curve (ef 16..96): the level-0 expansions
    the plain search needs for the learned search's recall, over the learned search's own.
    Expansions are the search's unit of work (one link row and up to 2M vector reads each); the
    time per expansion is measured by the benchmark."""
    from hnsw_rabitq_amd import DeviceArray, datasets, learned_stop
    p, space, data, _ = built_index("c2s")
    gpu = p.gpu()
    q = datasets.c2_gamma(n=2000, seed=9)
    nq, k, ef, period = q.shape[0], 10, 96, 8
    gt, _ = orc.bf_search(0, data, q, k)
    dq = DeviceArray.from_numpy(q)
    dl, dd, dc = DeviceArray(nq * k * 8), DeviceArray(nq * k * 4), DeviceArray(nq * 16)

    def run(e, model=None):
        gpu.search_device(dq, nq, k, e, dl, dd, d_counters=dc, stop_model=model)
        lab = dl.to_numpy(np.uint64, (nq, k))
        rec = np.mean([len(np.intersect1d(a, b)) for a, b in zip(lab, gt)]) / k
        return float(rec), float(dc.to_numpy(np.int32, (nq, 4))[:, 0].mean())

    curve = [run(e) for e in range(16, ef + 1, 4)]
    m, _, _, _ = _trained_model(gpu, ef, k, period, n_train=6000)
    m.config(period=period, threshold=0.9, min_hops=period)
    rec_l, hops_l = run(ef, m.h)
    assert hops_l < 0.95 * ef, "the network should stop a share of the search"
    recs = np.array([c[0] for c in curve])
    hops = np.array([c[1] for c in curve])
    assert rec_l <= recs.max(), "the learned search stops a fixed-ef search early: it cannot exceed it"
    i = int(np.searchsorted(recs, rec_l))
    i = min(max(i, 1), len(recs) - 1)
    w = (rec_l - recs[i - 1]) / max(recs[i] - recs[i - 1], 1e-12)
    hops_plain = hops[i - 1] + w * (hops[i] - hops[i - 1])
    gain = hops_plain / hops_l
    print(f"learned ef={ef}: recall {rec_l:.4f} at {hops_l:.1f} expansions; plain needs {hops_plain:.1f} for that "
          f"recall: gain {gain:.3f}x in expansions", flush=True)
    assert gain > 1.0
    m.free()
