"""Shared test helpers: workloads of the golden fixtures, rebuilt with the product's CPU builder."""
import hashlib
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from hnsw_rabitq_amd import datasets  # noqa: E402

_cache = {}


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def workload(name):
    """(space, data, queries) of a golden case."""
    if name == "c1":
        return "l2", datasets.c1_uniform(), datasets.c1_uniform(n=1000, seed=1)
    if name == "c2s":
        return "l2", datasets.c2_gamma(n=20_000), datasets.c2_gamma(n=200, seed=1)
    if name == "c4s":
        d, q = datasets.c4_gmm(n=20_000, nq=200, n_centres=64)
        return "ip", d, q
    if name == "cos21":
        z = np.load(os.path.join(ROOT, "tests", "golden", "cos21_inputs.npz"))
        return "cosine", z["data"], z["queries"]
    raise KeyError(name)


def built_index(name, num_threads=1):
    """Product Index built single-threaded (deterministic, byte-equal to the reference)."""
    key = (name, num_threads)
    if key not in _cache:
        import hnsw_rabitq_amd as H
        space, data, queries = workload(name)
        p = H.Index(space, data.shape[1])
        p.init_index(data.shape[0], M=16, ef_construction=200, random_seed=100)
        p.add_items(data, num_threads=num_threads)
        _cache[key] = (p, space, data, queries)
    return _cache[key]
