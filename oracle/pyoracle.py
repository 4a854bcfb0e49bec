"""ctypes wrapper of the CPU oracle (liboracle.so) and, where it was built, of the reference
compiled from its own sources (_ref/).  Test infrastructure: imported only by tests/, by the
smoke check and by bench.py's CPU-baseline leg.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")

MODE_HEAP, MODE_TOTAL = 0, 1
_vp, _sz, _i32 = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int

_lib = None


def _p(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_LIB):
            raise ImportError(f"{ORACLE_LIB} missing: run `make -C oracle liboracle.so`")
        L = ctypes.CDLL(ORACLE_LIB)
        L.or_last_error.restype = ctypes.c_char_p
        L.or_distance.restype = ctypes.c_float
        L.or_distance.argtypes = [_i32, _sz, _vp, _vp]
        L.or_distance_batch.argtypes = [_i32, _sz, _vp, _vp, _sz, _vp]
        L.or_distance_fp64.restype = ctypes.c_double
        L.or_distance_fp64.argtypes = [_i32, _sz, _vp, _vp]
        L.or_hnsw_load.restype = _vp
        L.or_hnsw_load.argtypes = [ctypes.c_char_p]
        L.or_hnsw_load_bytes.restype = _vp
        L.or_hnsw_load_bytes.argtypes = [_vp, _sz]
        L.or_hnsw_free.argtypes = [_vp]
        L.or_hnsw_count.restype = _sz
        L.or_hnsw_count.argtypes = [_vp]
        L.or_hnsw_maxlevel.argtypes = [_vp]
        L.or_hnsw_entry.restype = ctypes.c_uint32
        L.or_hnsw_entry.argtypes = [_vp]
        L.or_hnsw_search.argtypes = [_vp, _i32, _vp, _sz, _sz, _sz, _vp, _vp, _sz, _i32, _i32, _vp, _vp, _vp, _vp]
        L.or_hnsw_search_allowed.argtypes = [_vp, _i32, _vp, _sz, _sz, _sz, _vp, _vp, _sz, _i32, _i32, _vp, _vp,
                                             _vp, _vp, _vp]
        L.or_hnsw_labels.argtypes = [_vp, _vp, _vp]
        L.or_hnsw_search_hopcap.argtypes = [_vp, _i32, _vp, _sz, _sz, _sz, _vp, _sz, _i32, _i32, _vp, _vp, _vp, _vp]
        L.or_hnsw_search_multivector.argtypes = [_vp, _i32, _vp, _sz, _vp, _sz, _sz, _i32, _i32, _vp, _sz, _vp, _vp,
                                                 _vp, _vp, _vp]
        L.or_hnsw_search_epsilon.argtypes = [_vp, _i32, _vp, _sz, ctypes.c_float, _sz, _sz, _i32, _i32, _vp, _sz,
                                             _vp, _vp, _vp, _vp, _vp]
        L.or_bf_search.argtypes = [_i32, _sz, _vp, _vp, _sz, _vp, _sz, _sz, _i32, _vp, _vp]
        L.or_fp64_pairs.argtypes = [_i32, _sz, _vp, _vp, _vp, _vp, _sz, _vp]
        L.or_mlp_forward.argtypes = [_vp, _i32, _i32, _vp, _sz, _vp]
        L.or_mlp_loss_grad.restype = ctypes.c_double
        L.or_mlp_loss_grad.argtypes = [_vp, _i32, _i32, _vp, _vp, _sz, _vp]
        L.or_mlp_param_count.argtypes = [_i32, _i32]
        _lib = L
    return _lib


def distance(space, a, b):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    return lib().or_distance(space, a.shape[-1], _p(a), _p(b))


def distance_batch(space, a, b):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    out = np.empty(a.shape[0], np.float32)
    lib().or_distance_batch(space, a.shape[1], _p(a), _p(b), a.shape[0], _p(out))
    return out


class OracleHnsw:
    """Reference-format HNSW index held by the oracle (loaded from saveIndex bytes)."""

    def __init__(self, blob=None, path=None):
        if blob is not None:
            blob = np.ascontiguousarray(blob, dtype=np.uint8)
            self.h = lib().or_hnsw_load_bytes(_p(blob), blob.nbytes)
        else:
            self.h = lib().or_hnsw_load(str(path).encode())
        if not self.h:
            raise RuntimeError(lib().or_last_error().decode())

    def count(self):
        return lib().or_hnsw_count(self.h)

    def labels(self):
        """(label of each internal id, delete mark of each internal id)."""
        n = self.count()
        lab = np.empty(n, np.uint64)
        dele = np.empty(n, np.uint8)
        lib().or_hnsw_labels(self.h, _p(lab), _p(dele))
        return lab, dele

    def search(self, space, queries, k, ef, mode=MODE_HEAP, nthreads=0, ef_per_query=None, k_per_query=None,
               counters=False, ties=False, allowed=None):
        """allowed: optional uint8 mask over internal ids (the filter of searchKnn)."""
        q = np.ascontiguousarray(queries, np.float32)
        nq = q.shape[0]
        kmax = int(k if k_per_query is None else max(k, int(np.max(k_per_query))))
        labels = np.empty((nq, kmax), np.uint64)
        dists = np.empty((nq, kmax), np.float32)
        ctr = np.zeros((nq, 4), np.int64) if counters else None
        tf = np.zeros(nq, np.uint8) if ties else None
        efp = None if ef_per_query is None else np.ascontiguousarray(ef_per_query, np.int32)
        kp = None if k_per_query is None else np.ascontiguousarray(k_per_query, np.int32)
        al = None if allowed is None else np.ascontiguousarray(allowed, np.uint8)
        r = lib().or_hnsw_search_allowed(self.h, space, _p(q), nq, k, ef, _p(efp), _p(kp), kmax, mode, nthreads,
                                         _p(al), _p(labels), _p(dists), _p(ctr), _p(tf))
        if r < 0:
            raise RuntimeError(lib().or_last_error().decode())
        return labels, dists, ctr, tf, r

    def search_hopcap(self, space, queries, k, ef, hop_caps, mode=MODE_TOTAL, nthreads=0, counters=False):
        """The bare-bone search stopped after hop_caps[q] level-0 expansions (<= 0: no cap):
        (labels, dists, counters, short rows)."""
        q = np.ascontiguousarray(queries, np.float32)
        nq = q.shape[0]
        caps = np.ascontiguousarray(hop_caps, np.int32)
        assert caps.shape == (nq,)
        labels = np.empty((nq, k), np.uint64)
        dists = np.empty((nq, k), np.float32)
        ctr = np.zeros((nq, 4), np.int64) if counters else None
        r = lib().or_hnsw_search_hopcap(self.h, space, _p(q), nq, k, ef, _p(caps), k, mode, nthreads, _p(labels),
                                        _p(dists), _p(ctr), None)
        if r < 0:
            raise RuntimeError(lib().or_last_error().decode())
        return labels, dists, ctr, r

    def search_epsilon(self, space, queries, eps, min_c, max_c, out_stride, mode=MODE_HEAP, nthreads=0,
                       allowed=None, counters=False, ties=False):
        """searchStopConditionClosest + EpsilonSearchStopCondition: (labels, dists, counts,
        counters, tie flags); row q holds its first min(counts[q], out_stride) results."""
        q = np.ascontiguousarray(queries, np.float32)
        nq = q.shape[0]
        labels = np.empty((nq, out_stride), np.uint64)
        dists = np.empty((nq, out_stride), np.float32)
        counts = np.empty(nq, np.uint32)
        ctr = np.zeros((nq, 4), np.int64) if counters else None
        tf = np.zeros(nq, np.uint8) if ties else None
        al = None if allowed is None else np.ascontiguousarray(allowed, np.uint8)
        r = lib().or_hnsw_search_epsilon(self.h, space, _p(q), nq, eps, min_c, max_c, mode, nthreads, _p(al),
                                         out_stride, _p(labels), _p(dists), _p(counts), _p(ctr), _p(tf))
        if r < 0:
            raise RuntimeError(lib().or_last_error().decode())
        return labels, dists, counts, ctr, tf

    def search_multivector(self, space, queries, docs, num_docs, ef_collection, out_stride, mode=MODE_HEAP,
                           nthreads=0, allowed=None, counters=False, ties=False):
        """searchStopConditionClosest + MultiVectorSearchStopCondition; docs: uint32 per internal
        id.  Returns (labels, dists, counts, counters, tie flags)."""
        q = np.ascontiguousarray(queries, np.float32)
        nq = q.shape[0]
        dc = np.ascontiguousarray(docs, np.uint32)
        labels = np.empty((nq, out_stride), np.uint64)
        dists = np.empty((nq, out_stride), np.float32)
        counts = np.empty(nq, np.uint32)
        ctr = np.zeros((nq, 4), np.int64) if counters else None
        tf = np.zeros(nq, np.uint8) if ties else None
        al = None if allowed is None else np.ascontiguousarray(allowed, np.uint8)
        r = lib().or_hnsw_search_multivector(self.h, space, _p(q), nq, _p(dc), num_docs, ef_collection, mode,
                                             nthreads, _p(al), out_stride, _p(labels), _p(dists), _p(counts),
                                             _p(ctr), _p(tf))
        if r < 0:
            raise RuntimeError(lib().or_last_error().decode())
        return labels, dists, counts, ctr, tf

    def __del__(self):
        try:
            lib().or_hnsw_free(self.h)
        except Exception:
            pass


def bf_search(space, data, queries, k, labels=None, nthreads=0):
    data = np.ascontiguousarray(data, np.float32)
    q = np.ascontiguousarray(queries, np.float32)
    nq = q.shape[0]
    lab = np.empty((nq, k), np.uint64)
    dis = np.empty((nq, k), np.float32)
    li = None if labels is None else np.ascontiguousarray(labels, np.uint64)
    lib().or_bf_search(space, data.shape[1], _p(data), _p(li), data.shape[0], _p(q), nq, k, nthreads, _p(lab),
                       _p(dis))
    return lab, dis


def fp64_pairs(space, queries, data, qi, ri):
    q = np.ascontiguousarray(queries, np.float32)
    d = np.ascontiguousarray(data, np.float32)
    qi = np.ascontiguousarray(qi, np.int64)
    ri = np.ascontiguousarray(ri, np.int64)
    out = np.empty(qi.shape[0], np.float64)
    lib().or_fp64_pairs(space, q.shape[1], _p(q), _p(d), _p(qi), _p(ri), qi.shape[0], _p(out))
    return out


def mlp_param_count(in_dim, hidden):
    return lib().or_mlp_param_count(in_dim, hidden)


def mlp_forward(params, in_dim, hidden, x):
    params = np.ascontiguousarray(params, np.float32)
    x = np.ascontiguousarray(x, np.float32)
    out = np.empty(x.shape[0], np.float32)
    lib().or_mlp_forward(_p(params), in_dim, hidden, _p(x), x.shape[0], _p(out))
    return out


def mlp_loss_grad(params, in_dim, hidden, x, y):
    params = np.ascontiguousarray(params, np.float32)
    x = np.ascontiguousarray(x, np.float32)
    y = np.ascontiguousarray(y, np.float32)
    g = np.empty(params.shape[0], np.float64)
    loss = lib().or_mlp_loss_grad(_p(params), in_dim, hidden, _p(x), _p(y), x.shape[0], _p(g))
    return loss, g


def ref_module():
    """The reference's own Python module compiled from its sources, or None where absent."""
    import importlib
    import sys
    if not os.path.isdir(REF_DIR):
        return None
    if REF_DIR not in sys.path:
        sys.path.insert(0, REF_DIR)
    try:
        return importlib.import_module("hnswlib")
    except ImportError:
        return None


def ref_shim():
    """The reference headers compiled into _ref/libhnswref.so, or None where absent."""
    path = os.path.join(REF_DIR, "libhnswref.so")
    if not os.path.exists(path):
        return None
    try:
        L = ctypes.CDLL(path)
    except OSError:
        return None
    L.ref_distance_batch.argtypes = [_i32, _sz, _vp, _vp, _sz, _vp]
    L.ref_random_levels.argtypes = [_sz, _sz, _sz, _vp]
    L.ref_last_error.restype = ctypes.c_char_p
    L.ref_hnsw_build.restype = _vp
    L.ref_hnsw_build.argtypes = [_i32, _sz, _vp, _sz, _sz, _sz, _sz]
    L.ref_hnsw_free.argtypes = [_vp]
    L.ref_hnsw_mark_deleted.argtypes = [_vp, ctypes.c_uint64]
    L.ref_hnsw_search_epsilon.argtypes = [_vp, _vp, _sz, ctypes.c_float, _sz, _sz, _vp, _sz, _sz, _vp, _vp, _vp]
    L.ref_mv_build.restype = _vp
    L.ref_mv_build.argtypes = [_sz, _vp, _vp, _sz, _sz, _sz, _sz]
    L.ref_mv_free.argtypes = [_vp]
    L.ref_mv_search.argtypes = [_vp, _vp, _sz, _sz, _sz, _sz, _vp, _vp, _vp]
    return L
