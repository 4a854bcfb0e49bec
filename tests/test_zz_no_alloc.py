"""The allocation-free device search (hg_gpu_search_device_noalloc) against the allocating entry
points it mirrors (hg_gpu_search_device, hg_gpu_search_device_filtered): bit-identical rows,
counters and status on two calls in a row, the second with fewer queries, for the bare-bone
search and for the non-bare-bone search with a filter; the query limit of the fixed scratch.
"""
import ctypes

import numpy as np
import pytest

from helpers import built_index

pytestmark = pytest.mark.gpu


def _search(gpu, dq, nq, k, ef, noalloc, filt=None):
    from hnsw_rabitq_amd import DeviceArray, _lib
    dl, dd, dc, ds = DeviceArray(nq * k * 8), DeviceArray(nq * k * 4), DeviceArray(nq * 16), DeviceArray(4)
    _lib.call("hg_memset", ds.ptr, 0xFF, 4, None)
    fh = filt.handle if filt is not None else None
    if noalloc:
        _lib.call("hg_gpu_search_device_noalloc", gpu.handle, dq.ptr, nq, None, None, ef, k, k, dl.ptr, dd.ptr,
                  dc.ptr, fh, ds.ptr, None)
    elif filt is None:
        _lib.call("hg_gpu_search_device", gpu.handle, dq.ptr, nq, None, None, ef, k, k, dl.ptr, dd.ptr, dc.ptr, None,
                  None)
        _lib.call("hg_memset", ds.ptr, 0, 4, None)
    else:
        _lib.call("hg_gpu_search_device_filtered", gpu.handle, dq.ptr, nq, None, None, ef, k, k, dl.ptr, dd.ptr,
                  dc.ptr, fh, ds.ptr, None)
    _lib.call("hg_device_synchronize")
    return (dl.to_numpy(np.uint64, (nq, k)), dd.to_numpy(np.float32, (nq, k)).view(np.uint32),
            dc.to_numpy(np.int32, (nq, 4)), int(ds.to_numpy(np.uint32, (1,))[0]))


@pytest.mark.parametrize("filtered", [False, True])
def test_noalloc_equals_allocating_search(hg, filtered):
    from hnsw_rabitq_amd import DeviceArray
    p, space, data, queries = built_index("c2s")
    gpu = p.gpu()
    filt = None
    if filtered:
        labs = gpu.labels()
        filt = gpu.make_filter(labs[labs % 3 != 0])
    k, ef = 10, 64
    for nq in (200, 57):  # the second call is smaller: nothing of the first may leak into it
        print(f"case=c2s noalloc filtered={filtered} nq={nq}", flush=True)
        dq = DeviceArray.from_numpy(queries[:nq])
        want = _search(gpu, dq, nq, k, ef, False, filt)
        got = _search(gpu, dq, nq, k, ef, True, filt)
        for a, b in zip(want[:3], got[:3]):
            np.testing.assert_array_equal(a, b)
        assert got[3] == want[3] == 0
        if filtered:
            assert not np.isin(want[0], np.asarray(labs[labs % 3 == 0])).any()
    if filt is not None:
        filt.free()


def test_noalloc_query_limit(hg):
    from hnsw_rabitq_amd import DeviceArray, HnswGpuError, _lib
    p, space, data, queries = built_index("c2s")
    gpu = p.gpu()
    labs = gpu.labels()
    filt = gpu.make_filter(labs[::2])
    nq = 262_145
    dq = DeviceArray(nq * 128 * 4)
    _lib.call("hg_memset", dq.ptr, 0, nq * 128 * 4, None)
    dl, dd = DeviceArray(nq * 8), DeviceArray(nq * 4)
    with pytest.raises(HnswGpuError, match="at most 262,144 queries"):
        _lib.call("hg_gpu_search_device_noalloc", gpu.handle, dq.ptr, nq, None, None, 16, 1, 1, dl.ptr, dd.ptr, None,
                  filt.handle, None, None)
    filt.free()
