"""Pins of the oracle functions that have no reference counterpart, against the mathematics.

* or_mlp_forward (the learned stop's 12 -> 64 -> 64 -> 1 tanh-GELU network): a vectorised numpy
  float64 forward pass; a regime where every pre-activation is large, so GELU is the identity and
  the network is a closed-form affine map (pins the wiring: layer order, transposes, biases); the
  tanh form against the exact GELU x * Phi(x) (within its known 4.8e-4 approximation error) and
  its exact odd-part identity gelu(x) - gelu(-x) = x.
* or_mlp_loss_grad: every gradient entry against central finite differences of the function's own
  FP64 loss, and the loss against BCE-with-logits computed from the forward logits.
* or_distance_fp64 / or_fp64_pairs: exact sums (math.fsum) of the FP64 terms, integer data where
  every sum is exact, and the pairing of query and row indices.
* or_hnsw_search_hopcap (the oracle of the learned stop): a cap at or above a query's natural hop
  count changes nothing; a cap reports min(cap, natural) hops; the capped result's distances at
  every rank never increase with the cap (the evaluated sets are nested); heap and total order
  agree on tie-free queries.
"""
import math

import numpy as np
import pytest

from helpers import built_index


def _unpack(p, i, h):
    o = 0
    W1 = p[o:o + h * i].reshape(h, i); o += h * i
    b1 = p[o:o + h]; o += h
    W2 = p[o:o + h * h].reshape(h, h); o += h * h
    b2 = p[o:o + h]; o += h
    w3 = p[o:o + h]; o += h
    b3 = p[o]
    return W1, b1, W2, b2, w3, b3


def _pack(W1, b1, W2, b2, w3, b3):
    return np.concatenate([W1.ravel(), b1, W2.ravel(), b2, w3, [b3]]).astype(np.float32)


def _gelu_tanh(x):
    return 0.5 * x * (1.0 + np.tanh(np.sqrt(2.0 / np.pi) * (x + 0.044715 * x ** 3)))


def _numpy_forward(p, i, h, x):
    W1, b1, W2, b2, w3, b3 = (np.asarray(a, np.float64) for a in _unpack(p.astype(np.float64), i, h))
    a1 = x.astype(np.float64) @ W1.T + b1
    a2 = _gelu_tanh(a1) @ W2.T + b2
    return _gelu_tanh(a2) @ w3 + b3


def test_mlp_param_count(orc):
    assert orc.mlp_param_count(12, 64) == 12 * 64 + 64 + 64 * 64 + 64 + 64 + 1


def test_mlp_forward_matches_numpy(orc):
    rng = np.random.default_rng(5)
    i, h = 12, 64
    p = (rng.standard_normal(orc.mlp_param_count(i, h)) * 0.3).astype(np.float32)
    x = rng.standard_normal((257, i)).astype(np.float32)
    got = orc.mlp_forward(p, i, h, x)
    want = _numpy_forward(p, i, h, x)
    np.testing.assert_allclose(got, want, rtol=2e-7, atol=1e-6)


def test_mlp_forward_affine_regime(orc):
    """Pre-activations >= 20: GELU(a) == a in FP64, so the logit is w3.(W2 (W1 x + b1) + b2) + b3."""
    rng = np.random.default_rng(6)
    i, h = 5, 7
    W1 = rng.random((h, i)).astype(np.float32)
    b1 = np.full(h, 20.0, np.float32)
    W2 = (rng.random((h, h)) * 0.1).astype(np.float32)
    b2 = np.full(h, 20.0, np.float32)
    w3 = rng.standard_normal(h).astype(np.float32)
    b3 = np.float32(-3.5)
    p = _pack(W1, b1, W2, b2, w3, b3)
    x = rng.random((31, i)).astype(np.float32)
    f = lambda a: a.astype(np.float64)  # noqa: E731
    want = (f(x) @ f(W1).T + f(b1)) @ f(W2).T + f(b2)
    want = want @ f(w3) + float(b3)
    # (non-square layers: a transposed weight read would give another value)
    np.testing.assert_allclose(orc.mlp_forward(p, i, h, x), want, rtol=1e-6)


def test_mlp_gelu_against_exact(orc):
    """in=1, hidden=1, second layer in its affine regime: logit = gelu_tanh(x)."""
    p = _pack(np.array([[1.0]]), np.array([0.0]), np.array([[1.0]]), np.array([20.0]), np.array([1.0]), -20.0)
    x = np.linspace(-6, 6, 1201, dtype=np.float32).reshape(-1, 1)
    g = orc.mlp_forward(p, 1, 1, x).astype(np.float64)
    xs = x[:, 0].astype(np.float64)
    exact = 0.5 * xs * (1.0 + np.array([math.erf(v / math.sqrt(2.0)) for v in xs]))
    assert np.abs(g - exact).max() < 5e-4
    assert abs(g[600]) < 1e-6  # gelu(0) = 0
    # odd part: gelu(x) - gelu(-x) = x exactly for the tanh form
    np.testing.assert_allclose(g - g[::-1], xs, atol=2e-5)


def test_mlp_grad_central_differences(orc):
    rng = np.random.default_rng(7)
    i, h, n = 4, 6, 9
    P = orc.mlp_param_count(i, h)
    p = (rng.standard_normal(P) * 0.5).astype(np.float32)
    x = rng.standard_normal((n, i)).astype(np.float32)
    y = (rng.random(n) < 0.5).astype(np.float32)
    loss, g = orc.mlp_loss_grad(p, i, h, x, y)
    step = np.float32(2.0 ** -10)  # p +- step is exact in FP32 for |p| < 16
    assert np.abs(p).max() < 16
    fd = np.empty(P)
    for j in range(P):
        pp, pm = p.copy(), p.copy()
        pp[j] += step
        pm[j] -= step
        lp, _ = orc.mlp_loss_grad(pp, i, h, x, y)
        lm, _ = orc.mlp_loss_grad(pm, i, h, x, y)
        fd[j] = (lp - lm) / (2.0 * float(step))
    np.testing.assert_allclose(g, fd, rtol=1e-4, atol=1e-7)
    # the loss itself: mean BCE-with-logits of the forward logits
    o = orc.mlp_forward(p, i, h, x).astype(np.float64)
    bce = np.maximum(o, 0) - o * y + np.log1p(np.exp(-np.abs(o)))
    assert abs(loss - bce.mean()) < 1e-6


def test_distance_fp64(orc):
    rng = np.random.default_rng(8)
    for dim in (1, 7, 16, 128, 768):
        a = rng.standard_normal(dim).astype(np.float32)
        b = rng.standard_normal(dim).astype(np.float32)
        fa, fb = a.astype(np.float64), b.astype(np.float64)
        l2 = math.fsum((fa - fb) ** 2)
        ip = 1.0 - math.fsum(fa * fb)
        assert orc.lib().or_distance_fp64(0, dim, orc._p(a), orc._p(b)) == pytest.approx(l2, rel=1e-13)
        assert orc.lib().or_distance_fp64(1, dim, orc._p(a), orc._p(b)) == pytest.approx(ip, rel=1e-12, abs=1e-13)
    # integer data: every FP64 sum is exact
    a = rng.integers(0, 256, 128).astype(np.float32)
    b = rng.integers(0, 256, 128).astype(np.float32)
    assert orc.lib().or_distance_fp64(0, 128, orc._p(a), orc._p(b)) == float(((a.astype(np.int64) - b) ** 2).sum())
    assert orc.lib().or_distance_fp64(1, 128, orc._p(a), orc._p(b)) == 1.0 - float((a.astype(np.int64) * b).sum())


def test_fp64_pairs(orc):
    rng = np.random.default_rng(9)
    q = rng.standard_normal((11, 24)).astype(np.float32)
    d = rng.standard_normal((37, 24)).astype(np.float32)
    qi = rng.integers(0, 11, 200)
    ri = rng.integers(0, 37, 200)
    for space in (0, 1):
        got = orc.fp64_pairs(space, q, d, qi, ri)
        fq, fd = q.astype(np.float64)[qi], d.astype(np.float64)[ri]
        want = ((fq - fd) ** 2).sum(1) if space == 0 else 1.0 - (fq * fd).sum(1)
        np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12)


def test_hopcap_oracle(hg, orc):
    p, space, data, queries = built_index("c2s")
    o = orc.OracleHnsw(blob=p.serialize())
    q = queries[:100]
    k, ef = 10, 40
    base_l, base_d, base_c, _, _ = o.search(0, q, k, ef, mode=orc.MODE_TOTAL, counters=True)
    natural = base_c[:, 0].astype(np.int32)
    # no cap, a cap at the natural count and a cap above it: the plain search
    for caps in (np.zeros(100, np.int32), natural, natural + 5):
        l, d, c, _ = o.search_hopcap(0, q, k, ef, caps, counters=True)
        np.testing.assert_array_equal(l, base_l)
        np.testing.assert_array_equal(c, base_c)
    prev = None
    for cap in (1, 2, 5, 10, 20, 40, 80):
        caps = np.full(100, cap, np.int32)
        l, d, c, short = o.search_hopcap(0, q, k, ef, caps, counters=True)
        np.testing.assert_array_equal(c[:, 0], np.minimum(cap, natural))
        # (after one or two hops a query may hold fewer than k results: padded with +inf)
        if prev is not None:
            assert (d <= prev).all(), "a larger cap evaluates a superset: no rank can get farther"
        prev = d
        # heap order agrees with total order where the capped search met no equal distances
        hl, hd, _, tf, _ = o.search(0, q, k, ef, mode=orc.MODE_TOTAL, ties=True)
        lh, dh, _, _ = o.search_hopcap(0, q, k, ef, caps, mode=orc.MODE_HEAP)
        free = tf == 0
        np.testing.assert_array_equal(lh[free], l[free])
