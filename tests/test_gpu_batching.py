"""Uneven batches: per-query ef and k in one persistent launch, captured in a hipGraph and
replayed; every replay equals the eager launch and the oracle's total-order mode."""
import ctypes

import numpy as np
import pytest

from helpers import built_index

pytestmark = pytest.mark.gpu


def _arrays(nq, seed):
    from hnsw_rabitq_amd import datasets
    ef, k = datasets.c4_ef_k(nq, seed=seed, ef_lo=16, ef_hi=512, ks=(1, 10, 100))
    return ef, k


def _rows(lab, dis, k):
    return [(lab[i, :k[i]], dis[i, :k[i]]) for i in range(lab.shape[0])]


def test_per_query_ef_k_match_oracle(hg, orc):
    from hnsw_rabitq_amd import DeviceArray
    p, space, data, queries = built_index("c4s")
    nq = queries.shape[0]
    ef, k = _arrays(nq, 5)
    kmax = int(k.max())
    print(f"uneven batch nq={nq} ef in [{ef.min()},{ef.max()}] k in {sorted(set(k.tolist()))}", flush=True)
    gpu = p.gpu()
    dq = DeviceArray.from_numpy(queries)
    def_, dk = DeviceArray.from_numpy(ef), DeviceArray.from_numpy(k)
    dl, dd = DeviceArray(nq * kmax * 8), DeviceArray(nq * kmax * 4)
    gpu.search_device(dq, nq, kmax, int(ef.max()), dl, dd, d_ef=def_, d_k=dk, k_stride=kmax)
    lab = dl.to_numpy(np.uint64, (nq, kmax))
    dis = dd.to_numpy(np.float32, (nq, kmax))
    o = orc.OracleHnsw(blob=p.serialize())
    ol, od, _, _, short = o.search(1, queries, kmax, int(ef.max()), mode=orc.MODE_TOTAL, ef_per_query=ef,
                                   k_per_query=k)
    assert short == 0
    for i in range(nq):
        np.testing.assert_array_equal(lab[i, :k[i]], ol[i, :k[i]])
        np.testing.assert_array_equal(dis[i, :k[i]].view(np.uint32), od[i, :k[i]].view(np.uint32))


@pytest.mark.parametrize("case", ["c4s", "c2s"])  # FP32 rows; the 8-bit rows of integer data
def test_graph_capture_replays_equal_eager(hg, case):
    from hnsw_rabitq_amd import DeviceArray, _lib
    print(f"graph capture: {case}", flush=True)
    p, space, data, queries = built_index(case)
    nq = queries.shape[0]
    ef, k = _arrays(nq, 6)
    kmax = int(k.max())
    gpu = p.gpu()
    stream = ctypes.c_void_p()
    _lib.call("hg_stream_create", ctypes.byref(stream))
    dq = DeviceArray.from_numpy(queries)
    def_, dk = DeviceArray.from_numpy(ef), DeviceArray.from_numpy(k)
    el, ed = DeviceArray(nq * kmax * 8), DeviceArray(nq * kmax * 4)
    gpu.search_device(dq, nq, kmax, int(ef.max()), el, ed, d_ef=def_, d_k=dk, k_stride=kmax, stream=stream)
    ref_l = el.to_numpy(np.uint64, (nq, kmax))
    ref_d = ed.to_numpy(np.float32, (nq, kmax))
    gl, gd = DeviceArray(nq * kmax * 8), DeviceArray(nq * kmax * 4)
    graph = ctypes.c_void_p()
    _lib.call("hg_graph_capture_search", gpu.handle, dq.ptr, nq, def_.ptr, dk.ptr, int(ef.max()), kmax, kmax,
              gl.ptr, gd.ptr, None, None, stream, ctypes.byref(graph))
    for rep in range(4):
        _lib.call("hg_memset", gl.ptr, 0, nq * kmax * 8, stream)
        _lib.call("hg_graph_launch", graph, stream)
        _lib.call("hg_stream_synchronize", stream)
        # eager work between replays
        gpu.search_device(dq, nq, kmax, 32, el, ed, stream=stream)
        l = gl.to_numpy(np.uint64, (nq, kmax))
        d = gd.to_numpy(np.float32, (nq, kmax))
        for i in range(nq):
            np.testing.assert_array_equal(l[i, :k[i]], ref_l[i, :k[i]], err_msg=f"replay {rep}")
            np.testing.assert_array_equal(d[i, :k[i]], ref_d[i, :k[i]], err_msg=f"replay {rep}")
    _lib.call("hg_graph_free", graph)
    _lib.call("hg_stream_destroy", stream)


def test_autotune_keeps_results(hg):
    from hnsw_rabitq_amd import DeviceArray
    p, space, data, queries = built_index("c1")
    gpu = p.gpu()
    nq, k, ef = queries.shape[0], 10, 50
    dq = DeviceArray.from_numpy(queries)
    l0, d0 = DeviceArray(nq * k * 8), DeviceArray(nq * k * 4)
    gpu.search_device(dq, nq, k, ef, l0, d0)
    a = l0.to_numpy(np.uint64, (nq, k))
    gpu.autotune(dq, nq, ef, k)
    plan = gpu.plan(ef, k, nq)
    assert plan["hash_slots"] >= 2 * k and plan["beam_cap"] >= ef
    gpu.search_device(dq, nq, k, ef, l0, d0)
    np.testing.assert_array_equal(l0.to_numpy(np.uint64, (nq, k)), a)


def test_graph_replay_refuses_stale_index(hg):
    """A captured search keeps the kernels chosen at capture: after the delete marks are re-read
    (mark_deleted, then the index refreshes them) or the GPU index is freed, a replay fails
    instead of returning deleted labels or reading freed buffers."""
    from hnsw_rabitq_amd import DeviceArray, HnswGpuError, _lib
    rng = np.random.default_rng(2)
    data = rng.random((3000, 16), dtype=np.float32)
    p = hg.Index("l2", 16)
    p.init_index(3000, M=16, ef_construction=100)
    p.add_items(data)
    q = data[:64] + 0.001
    gpu = p.gpu()
    stream = ctypes.c_void_p()
    _lib.call("hg_stream_create", ctypes.byref(stream))
    dq = DeviceArray.from_numpy(q)
    gl, gd = DeviceArray(64 * 5 * 8), DeviceArray(64 * 5 * 4)
    graph = ctypes.c_void_p()
    _lib.call("hg_graph_capture_search", gpu.handle, dq.ptr, 64, None, None, 40, 5, 5, gl.ptr, gd.ptr, None, None,
              stream, ctypes.byref(graph))
    _lib.call("hg_graph_launch", graph, stream)
    _lib.call("hg_stream_synchronize", stream)
    assert gl.to_numpy(np.uint64, (64, 5))[0, 0] == 0
    p.mark_deleted(0)
    assert p.gpu() is gpu  # delete marks re-read in place
    with pytest.raises(HnswGpuError, match="delete marks changed"):
        _lib.call("hg_graph_launch", graph, stream)
    _lib.call("hg_graph_free", graph)
    graph = ctypes.c_void_p()
    _lib.call("hg_graph_capture_search", gpu.handle, dq.ptr, 64, None, None, 40, 5, 5, gl.ptr, gd.ptr, None, None,
              stream, ctypes.byref(graph))
    _lib.call("hg_graph_launch", graph, stream)
    _lib.call("hg_stream_synchronize", stream)
    assert 0 not in gl.to_numpy(np.uint64, (64, 5))  # captured with the mark: the filtered kernel
    p._gpu.free()
    p._gpu = None
    with pytest.raises(HnswGpuError, match="was freed"):
        _lib.call("hg_graph_launch", graph, stream)
    _lib.call("hg_graph_free", graph)
    _lib.call("hg_stream_destroy", stream)
