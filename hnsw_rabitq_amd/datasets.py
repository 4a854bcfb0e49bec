"""Synthetic workloads of the five benchmark configurations (numpy PCG64 generators).

  c1_uniform      N=10,000 x 16 f32 uniform[0,1) (seed 0); queries 1,000 (seed 1); L2
  c2_gamma        1,000,000 x 128 f32, clip(round(Gamma(0.5, 40)), 0, 255) per coordinate
                  (seed 0); queries 10,000 (seed 1); L2
  c3_gmm / c4_gmm Gaussian mixture of unit-normalised points: centres N(0,1)^dim, points
                  centre + N(0, sigma^2), L2-normalised; inner-product space
  c4_ef_k         per-query ef log-uniform in [16, 512] and k uniform over {1, 10, 100}
Generation is chunked so the million-row sets never hold more than one float64 chunk.
"""
import numpy as np

CHUNK = 1 << 16


def c1_uniform(n=10_000, dim=16, seed=0):
    return np.random.default_rng(seed).random((n, dim), dtype=np.float32)


def c2_gamma(n=1_000_000, dim=128, seed=0):
    rng = np.random.default_rng(seed)
    out = np.empty((n, dim), dtype=np.float32)
    for s in range(0, n, CHUNK):
        e = min(n, s + CHUNK)
        g = rng.gamma(0.5, 40.0, size=(e - s, dim))
        out[s:e] = np.clip(np.round(g), 0, 255).astype(np.float32)
    return out


def gmm_centres(n_centres, dim, seed):
    return np.random.default_rng(seed).standard_normal((n_centres, dim)).astype(np.float32)


def gmm_points(centres, n, sigma, seed):
    rng = np.random.default_rng(seed)
    dim = centres.shape[1]
    out = np.empty((n, dim), dtype=np.float32)
    for s in range(0, n, CHUNK):
        e = min(n, s + CHUNK)
        idx = rng.integers(0, centres.shape[0], size=e - s)
        x = centres[idx].astype(np.float64) + sigma * rng.standard_normal((e - s, dim))
        x /= np.linalg.norm(x, axis=1, keepdims=True)
        out[s:e] = x.astype(np.float32)
    return out


def c3_gmm(n=1_000_000, nq=10_000, dim=768, n_centres=1024, sigma=0.35, seed=0):
    c = gmm_centres(n_centres, dim, seed)
    return gmm_points(c, n, sigma, seed + 1), gmm_points(c, nq, sigma, seed + 2)


def c4_gmm(n=10_000_000, nq=100_000, dim=96, n_centres=4096, sigma=0.5, seed=0):
    c = gmm_centres(n_centres, dim, seed)
    return gmm_points(c, n, sigma, seed + 1), gmm_points(c, nq, sigma, seed + 2)


def c4_ef_k(nq, seed=3, ef_lo=16, ef_hi=512, ks=(1, 10, 100)):
    rng = np.random.default_rng(seed)
    ef = np.round(np.exp(rng.uniform(np.log(ef_lo), np.log(ef_hi), size=nq))).astype(np.int32)
    k = np.asarray(ks, dtype=np.int32)[rng.integers(0, len(ks), size=nq)]
    return ef, k
