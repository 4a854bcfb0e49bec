"""MI355X query engine for hnswlib indexes.

Drop-in mirror of the reference Python API for the query path (python_bindings/bindings.cpp):
  Index(space, dim)     init_index / add_items (CPU graph construction, restating addPoint)
                        knn_query -> batched HNSW search on the GPU
                        set_ef / ef, save_index / load_index (hnswlib on-disk format)
  BFIndex(space, dim)   init_index / add_items / knn_query -> GPU brute force
Results have the reference's shapes, dtypes and order: labels uint64 [rows, k], distances
float32 [rows, k], closest first (bindings.cpp:612-697, 850-929).  Errors raise RuntimeError
with the reference's messages where the condition is the same.

The query path runs only on the GPU: importing works without one, but searching without the
built library or a device raises.
"""
import ctypes
import os
import sys

import numpy as np

from . import _lib
from ._lib import HnswGpuError, PREC_BF16, PREC_FP32, PREC_FP8

__all__ = ["Index", "BFIndex", "DeviceArray", "GpuHnsw", "HnswGpuError", "normalize_rows"]

_SPACES = {"l2": (_lib.SPACE_L2, False), "ip": (_lib.SPACE_IP, False), "cosine": (_lib.SPACE_IP, True)}
_PRECISIONS = {"fp32": PREC_FP32, "bf16": PREC_BF16, "fp8": PREC_FP8}


def _space(name):
    if name not in _SPACES:
        raise RuntimeError("Space name must be one of l2, ip, or cosine.")
    return _SPACES[name]


def _as_rows(data, dim):
    arr = np.ascontiguousarray(np.asarray(data, dtype=np.float32))
    if arr.ndim == 1:
        arr = arr.reshape(1, -1)
    if arr.ndim != 2:
        raise RuntimeError(
            f"Input vector data wrong shape. Number of dimensions {arr.ndim}. Data must be a 1D or 2D array.")
    if arr.shape[1] != dim:
        raise RuntimeError("Wrong dimensionality of the vectors")
    return arr


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


def normalize_rows(arr):
    """Index.normalize_vector (bindings.cpp:241-248) row by row, in the reference's FP32 order."""
    arr = np.ascontiguousarray(arr, dtype=np.float32)
    out = np.empty_like(arr)
    _lib.call("hg_normalize_rows", _ptr(arr), _ptr(out), arr.shape[0], arr.shape[1])
    return out


class DeviceArray:
    """A device allocation owned by Python (hipMalloc through the C ABI)."""

    def __init__(self, nbytes):
        p = ctypes.c_void_p()
        _lib.call("hg_malloc", ctypes.byref(p), max(int(nbytes), 1))
        self.ptr = p
        self.nbytes = int(nbytes)

    @classmethod
    def from_numpy(cls, a, stream=None):
        a = np.ascontiguousarray(a)
        d = cls(a.nbytes)
        _lib.call("hg_memcpy_h2d", d.ptr, _ptr(a), a.nbytes, stream)
        return d

    def to_numpy(self, dtype, shape, stream=None):
        out = np.empty(shape, dtype=dtype)
        _lib.call("hg_memcpy_d2h", _ptr(out), self.ptr, out.nbytes, stream)
        return out

    def free(self):
        if self.ptr is not None and self.ptr.value:
            _lib.call("hg_free", self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class GpuHnsw:
    """Device copy of a CPU-built HNSW index (the exported GPU layout)."""

    def __init__(self, cpu_handle, device=0, path=None):
        h = ctypes.c_void_p()
        if path is not None:
            _lib.call("hg_gpu_index_load", str(path).encode(), device, ctypes.byref(h))
        else:
            _lib.call("hg_gpu_index_create", cpu_handle, device, ctypes.byref(h))
        self.handle = h

    @classmethod
    def load(cls, path, device=0):
        """A GPU index from a layout cache file written by save()."""
        return cls(None, device, path=path)

    def save(self, path):
        """Write the versioned GPU-layout cache of this index."""
        _lib.call("hg_gpu_index_save", self.handle, str(path).encode())

    @property
    def vector_code(self):
        """"u8" when the search reads the lossless 8-bit copy of the vectors, else "f32"."""
        c = ctypes.c_int()
        _lib.call("hg_gpu_vector_code", self.handle, ctypes.byref(c))
        return "u8" if c.value == _lib.VEC_U8 else "f32"

    def set_vector_code(self, code):
        """"auto": read the 8-bit copy where the export found it lossless (default); "f32": read
        the FP32 rows.  Distances are bit-identical either way."""
        if code not in ("auto", "f32"):
            raise ValueError("code must be 'auto' or 'f32'")
        _lib.call("hg_gpu_set_vector_code", self.handle, _lib.VEC_AUTO if code == "auto" else _lib.VEC_F32)

    def search(self, queries, k, ef, counters=False, filter=None):
        """filter: None or a GpuFilter of this index (non-bare-bone search)."""
        q = np.ascontiguousarray(queries, dtype=np.float32)
        nq = q.shape[0]
        labels = np.empty((nq, k), dtype=np.uint64)
        dists = np.empty((nq, k), dtype=np.float32)
        ctr = np.empty((nq, 4), dtype=np.int32) if counters else None
        lib = _lib.load()
        if filter is None:
            code = lib.hg_gpu_search(self.handle, _ptr(q), nq, k, ef, _ptr(labels), _ptr(dists), _ptr(ctr))
        else:
            if filter.owner is not self:
                raise ValueError("the filter was built for another GPU index")
            code = lib.hg_gpu_search_filtered(self.handle, _ptr(q), nq, k, ef, filter.handle, _ptr(labels),
                                              _ptr(dists), _ptr(ctr))
        if code not in (_lib.HG_OK, _lib.HG_ERR_SHORT):
            _lib.check(code)
        return labels, dists, ctr, code == _lib.HG_ERR_SHORT

    def search_epsilon(self, queries, epsilon, min_candidates, max_candidates, out_stride=None, filter=None):
        """Range search (searchStopConditionClosest + EpsilonSearchStopCondition): (labels, dists,
        counts); row q holds the closest min(counts[q], out_stride) results."""
        q = np.ascontiguousarray(queries, dtype=np.float32)
        nq = q.shape[0]
        stride = int(out_stride or min(max_candidates, 8192))
        if filter is not None and filter.owner is not self:
            raise ValueError("the filter was built for another GPU index")
        labels = np.empty((nq, stride), dtype=np.uint64)
        dists = np.empty((nq, stride), dtype=np.float32)
        counts = np.empty(nq, dtype=np.uint32)
        _lib.call("hg_gpu_search_epsilon", self.handle, _ptr(q), nq, float(epsilon), int(min_candidates),
                  int(max_candidates), filter.handle if filter is not None else None, stride, _ptr(labels),
                  _ptr(dists), _ptr(counts))
        return labels, dists, counts

    def make_docs(self, labels, docs):
        """Documents of the elements (for search_multivector): docs[i] is the document of labels[i]."""
        return GpuDocs(self, labels, docs)

    def search_multivector(self, queries, docs, num_docs, ef_collection, out_stride=None, filter=None):
        """MultiVectorSearchStopCondition search: (labels, dists, counts)."""
        q = np.ascontiguousarray(queries, dtype=np.float32)
        nq = q.shape[0]
        stride = int(out_stride or 1024)
        if docs.owner is not self or (filter is not None and filter.owner is not self):
            raise ValueError("the documents or filter were built for another GPU index")
        labels = np.empty((nq, stride), dtype=np.uint64)
        dists = np.empty((nq, stride), dtype=np.float32)
        counts = np.empty(nq, dtype=np.uint32)
        _lib.call("hg_gpu_search_multivector", self.handle, _ptr(q), nq, docs.handle, int(num_docs),
                  int(ef_collection), filter.handle if filter is not None else None, stride, _ptr(labels),
                  _ptr(dists), _ptr(counts))
        return labels, dists, counts

    def labels(self):
        """External label of every internal id (getExternalLabel, hnswalg.h:185-189)."""
        n = ctypes.c_size_t()
        _lib.call("hg_gpu_index_labels", self.handle, None, 0, ctypes.byref(n))
        out = np.empty(n.value, dtype=np.uint64)
        _lib.call("hg_gpu_index_labels", self.handle, _ptr(out), n.value, ctypes.byref(n))
        return out

    def make_filter(self, allowed_labels):
        return GpuFilter(self, allowed_labels)

    def refresh_deleted(self, cpu_handle):
        _lib.call("hg_gpu_refresh_deleted", self.handle, cpu_handle)

    def status(self):
        """(status word, queries that needed the second pass) of the last launch on this index."""
        st, second = ctypes.c_uint32(), ctypes.c_uint32()
        _lib.call("hg_gpu_search_status", self.handle, ctypes.byref(st), ctypes.byref(second))
        return st.value, second.value

    def search_device(self, d_queries, nq, k, ef, d_labels, d_dists, d_counters=None, d_ef=None, d_k=None,
                      k_stride=None, stop_model=None, stream=None):
        _lib.call("hg_gpu_search_device", self.handle, d_queries.ptr, nq, d_ef.ptr if d_ef else None,
                  d_k.ptr if d_k else None, ef, k, k_stride or k, d_labels.ptr, d_dists.ptr,
                  d_counters.ptr if d_counters else None, stop_model, stream)

    def search_device_filtered(self, d_queries, nq, k, ef, d_labels, d_dists, filter=None, d_status=None,
                               d_counters=None, d_ef=None, d_k=None, k_stride=None, stream=None):
        """Device-resident non-bare-bone search (deleted elements, optional GpuFilter)."""
        if filter is not None and filter.owner is not self:
            raise ValueError("the filter was built for another GPU index")
        _lib.call("hg_gpu_search_device_filtered", self.handle, d_queries.ptr, nq, d_ef.ptr if d_ef else None,
                  d_k.ptr if d_k else None, ef, k, k_stride or k, d_labels.ptr, d_dists.ptr,
                  d_counters.ptr if d_counters else None, filter.handle if filter is not None else None,
                  d_status.ptr if d_status else None, stream)

    def plan(self, max_ef, max_k, nq):
        out = np.zeros(5, dtype=np.int32)
        _lib.call("hg_gpu_search_plan", self.handle, max_ef, max_k, nq, _ptr(out))
        return dict(waves_per_group=int(out[0]), hash_slots=int(out[1]), beam_cap=int(out[2]),
                    lds_bytes=int(out[3]), workgroups=int(out[4]))

    def autotune(self, d_queries, nq, ef, k, stream=None):
        _lib.call("hg_gpu_autotune", self.handle, d_queries.ptr, nq, ef, k, stream)

    def set_plan(self, ef, k, hash_slots, workgroups_per_cu):
        """Fix the launch plan for (ef, k): hash_slots must be a power of two."""
        bits = int(hash_slots).bit_length() - 1
        if hash_slots != 1 << bits:
            raise ValueError("hash_slots must be a power of two")
        _lib.call("hg_gpu_set_plan", self.handle, ef, k, bits, workgroups_per_cu)

    def nbytes(self):
        b = ctypes.c_size_t()
        _lib.call("hg_gpu_index_bytes", self.handle, ctypes.byref(b))
        return b.value

    def free(self):
        if self.handle is not None and self.handle.value:
            _lib.load().hg_gpu_index_free(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class GpuFilter:
    """The allowed set of a filtered search over one GPU index, as a bitmap over its elements.

    hnswlib evaluates its filter functor (hnswlib.h:16-20) on each element the search reaches;
    the GPU search reads the same answers from this bitmap, built once from the allowed labels.
    """

    def __init__(self, gpu_index, allowed_labels):
        lab = np.ascontiguousarray(np.asarray(allowed_labels, dtype=np.uint64).reshape(-1))
        h = ctypes.c_void_p()
        _lib.call("hg_gpu_filter_create", gpu_index.handle, _ptr(lab), lab.shape[0], ctypes.byref(h))
        self.handle = h
        self.owner = gpu_index

    @property
    def allowed(self):
        c = ctypes.c_size_t()
        _lib.call("hg_gpu_filter_count", self.handle, ctypes.byref(c))
        return c.value

    def free(self):
        if self.handle is not None and self.handle.value:
            _lib.load().hg_gpu_filter_free(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class GpuDocs:
    """The document of every element of one GPU index (multi-vector search)."""

    def __init__(self, gpu_index, labels, docs):
        lab = np.ascontiguousarray(np.asarray(labels, dtype=np.uint64).reshape(-1))
        dc = np.ascontiguousarray(np.asarray(docs, dtype=np.uint32).reshape(-1))
        if lab.shape != dc.shape:
            raise ValueError("labels and docs differ in length")
        h = ctypes.c_void_p()
        _lib.call("hg_gpu_docs_create", gpu_index.handle, _ptr(lab), _ptr(dc), lab.shape[0], ctypes.byref(h))
        self.handle = h
        self.owner = gpu_index

    def free(self):
        if self.handle is not None and self.handle.value:
            _lib.load().hg_gpu_docs_free(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Index:
    """hnswlib.Index mirror (bindings.cpp:143-722, 933-1010) with the GPU query path."""

    def __init__(self, space, dim):
        self.space_id, self.normalize = _space(space)
        self.space = space
        self.dim = int(dim)
        self._h = None
        self._gpu = None
        self._dirty = True
        self._marks_dirty = False
        self.default_ef = 10
        self.num_threads = os.cpu_count() or 1
        self.device = 0
        self.ef_construction_ = None
        self.M_ = None

    # -- construction (CPU) --
    def init_index(self, max_elements, M=16, ef_construction=200, random_seed=100, allow_replace_deleted=False):
        if self._h is not None:
            raise RuntimeError("The index is already initiated.")
        if allow_replace_deleted:
            raise RuntimeError("allow_replace_deleted: deletions are outside the GPU query path")
        h = ctypes.c_void_p()
        _lib.call("hg_hnsw_create", self.space_id, self.dim, int(max_elements), int(M), int(ef_construction),
                  int(random_seed), ctypes.byref(h))
        self._h = h
        _lib.call("hg_hnsw_set_ef", self._h, self.default_ef)
        self._dirty = True

    def add_items(self, data, ids=None, num_threads=-1, replace_deleted=False):
        if self._h is None:
            raise RuntimeError("Index not initialized")
        if replace_deleted:
            raise RuntimeError("Replacement of deleted elements is disabled in constructor")
        arr = _as_rows(data, self.dim)
        if self.normalize:
            arr = normalize_rows(arr)
        lab = None
        if ids is not None:
            lab = np.ascontiguousarray(np.asarray(ids, dtype=np.uint64).reshape(-1))
            if lab.shape[0] != arr.shape[0]:
                raise RuntimeError("The input label shape %d does not match the input data vector shape %d"
                                   % (lab.ndim, arr.shape[0]))
        nt = self.num_threads if num_threads <= 0 else num_threads
        _lib.call("hg_hnsw_add_items", self._h, _ptr(arr), arr.shape[0], _ptr(lab), int(nt))
        self._dirty = True

    def add_items_gpu(self, data, ids=None, num_threads=-1):
        """add_items with the level-0 graph built on the GPU (levels above 0 on the CPU)."""
        if self._h is None:
            raise RuntimeError("Index not initialized")
        arr = _as_rows(data, self.dim)
        if self.normalize:
            arr = normalize_rows(arr)
        lab = None
        if ids is not None:
            lab = np.ascontiguousarray(np.asarray(ids, dtype=np.uint64).reshape(-1))
            if lab.shape[0] != arr.shape[0]:
                raise RuntimeError("The input label shape %d does not match the input data vector shape %d"
                                   % (lab.ndim, arr.shape[0]))
        nt = self.num_threads if num_threads <= 0 else num_threads
        _lib.call("hg_hnsw_build_gpu", self._h, _ptr(arr), arr.shape[0], _ptr(lab), int(nt), self.device)
        self._dirty = True

    # -- parameters --
    def set_ef(self, ef):
        self.default_ef = int(ef)
        if self._h is not None:
            _lib.call("hg_hnsw_set_ef", self._h, int(ef))

    @property
    def ef(self):
        return self.default_ef

    @ef.setter
    def ef(self, v):
        self.set_ef(v)

    def set_num_threads(self, n):
        self.num_threads = int(n)

    def _info(self):
        info = np.zeros(8, dtype=np.int64)
        _lib.call("hg_hnsw_info", self._h, _ptr(info))
        return info

    def get_current_count(self):
        return int(self._info()[0]) if self._h is not None else 0

    @property
    def element_count(self):
        return self.get_current_count()

    def get_max_elements(self):
        return int(self._info()[1])

    @property
    def max_elements(self):
        return self.get_max_elements() if self._h is not None else 0

    @property
    def M(self):
        return int(self._info()[3])

    @property
    def ef_construction(self):
        return int(self._info()[4])

    # -- persistence (hnswlib on-disk format) --
    def save_index(self, path):
        _lib.call("hg_hnsw_save", self._h, str(path).encode())

    def serialize(self):
        size = ctypes.c_size_t()
        _lib.call("hg_hnsw_serialize", self._h, None, 0, ctypes.byref(size))
        buf = np.empty(size.value, dtype=np.uint8)
        _lib.call("hg_hnsw_serialize", self._h, _ptr(buf), size.value, ctypes.byref(size))
        return buf

    def load_index(self, path, max_elements=0, allow_replace_deleted=False):
        if self._h is not None:
            _lib.load().hg_hnsw_free(self._h)
            self._h = None
        h = ctypes.c_void_p()
        _lib.call("hg_hnsw_load", self.space_id, self.dim, str(path).encode(), int(max_elements), ctypes.byref(h))
        self._h = h
        _lib.call("hg_hnsw_set_ef", self._h, self.default_ef)
        self._dirty = True

    # -- query (GPU) --
    def gpu(self):
        """The device copy of the index, exported on first use after a change (delete marks
        alone are re-read without a new export)."""
        if self._gpu is None or self._dirty:
            if self._gpu is not None:
                self._gpu.free()
            self._gpu = GpuHnsw(self._h, self.device)
            self._dirty = False
            self._marks_dirty = False
        elif self._marks_dirty:
            self._gpu.refresh_deleted(self._h)
            self._marks_dirty = False
        return self._gpu

    def mark_deleted(self, label):
        """HierarchicalNSW::markDelete: the element stays in the graph and is never returned."""
        if self._h is None:
            raise RuntimeError("Index not initialized")
        _lib.call("hg_hnsw_mark_deleted", self._h, int(label))
        self._marks_dirty = True

    def unmark_deleted(self, label):
        if self._h is None:
            raise RuntimeError("Index not initialized")
        _lib.call("hg_hnsw_unmark_deleted", self._h, int(label))
        self._marks_dirty = True

    def make_filter(self, allowed_labels):
        """A reusable GpuFilter of the current export that allows the given labels."""
        return self.gpu().make_filter(allowed_labels)

    def multivector_query(self, data, docs, num_docs, ef_collection=10, filter=None):
        """The C++ API's searchStopConditionClosest with MultiVectorSearchStopCondition (no Python
        binding in hnswlib): per query, the closest elements covering num_docs documents.  docs:
        a mapping label -> document, or a sequence indexed by label.  Returns a list of
        (labels, distances) pairs."""
        if self._h is None:
            raise RuntimeError("Index not initialized")
        arr = _as_rows(data, self.dim)
        if self.normalize:
            arr = normalize_rows(arr)
        if arr.shape[0] == 0:
            return []
        gpu = self.gpu()
        labs = gpu.labels()
        if isinstance(docs, dict):
            dc = np.array([docs[int(x)] for x in labs], dtype=np.uint32)
        else:
            dc = np.asarray(docs, dtype=np.uint32)[labs.astype(np.int64)]
        gd = gpu.make_docs(labs, dc)
        gf = filter if isinstance(filter, GpuFilter) or filter is None else None
        if filter is not None and gf is None:
            keep = np.fromiter((bool(filter(int(x))) for x in labs), dtype=bool, count=labs.shape[0])
            gf = gpu.make_filter(labs[keep])
        try:
            lab, dis, cnt = gpu.search_multivector(arr, gd, num_docs, ef_collection, filter=gf)
        finally:
            gd.free()
            if gf is not None and gf is not filter:
                gf.free()
        if (cnt > lab.shape[1]).any():
            raise RuntimeError("more results than the row stride; lower num_docs")
        return [(lab[i, :cnt[i]].copy(), dis[i, :cnt[i]].copy()) for i in range(arr.shape[0])]

    def range_query(self, data, epsilon, min_candidates, max_candidates, filter=None):
        """The C++ API's searchStopConditionClosest with EpsilonSearchStopCondition (no Python
        binding in hnswlib): per query, the elements within `epsilon` (distance as the space
        reports it) found by the search, closest first, at most max_candidates.  Returns a list
        of (labels, distances) pairs.  filter: a callable label -> bool or a GpuFilter."""
        if self._h is None:
            raise RuntimeError("Index not initialized")
        arr = _as_rows(data, self.dim)
        if self.normalize:
            arr = normalize_rows(arr)
        if arr.shape[0] == 0:
            return []
        gpu = self.gpu()
        gf = filter if isinstance(filter, GpuFilter) or filter is None else None
        if filter is not None and gf is None:
            labs = gpu.labels()
            keep = np.fromiter((bool(filter(int(x))) for x in labs), dtype=bool, count=labs.shape[0])
            gf = gpu.make_filter(labs[keep])
        try:
            lab, dis, cnt = gpu.search_epsilon(arr, epsilon, min_candidates, max_candidates, filter=gf)
        finally:
            if gf is not None and gf is not filter:
                gf.free()
        if (cnt > lab.shape[1]).any():
            raise RuntimeError("more than 8192 results in range; lower max_candidates")
        return [(lab[i, :cnt[i]].copy(), dis[i, :cnt[i]].copy()) for i in range(arr.shape[0])]

    def knn_query(self, data, k=1, num_threads=-1, filter=None):
        """filter: a callable label -> bool, as in hnswlib (evaluated here once per element),
        or a GpuFilter from make_filter."""
        if self._h is None:
            raise RuntimeError("Index not initialized")
        if filter is not None and not (isinstance(filter, GpuFilter) or callable(filter)):
            raise TypeError("filter must be a callable taking a label, or a GpuFilter")
        arr = _as_rows(data, self.dim)
        if self.normalize:
            arr = normalize_rows(arr)
        k = int(k)
        if arr.shape[0] == 0:
            return np.empty((0, k), dtype=np.uint64), np.empty((0, k), dtype=np.float32)
        gpu = self.gpu()
        gf = None
        if isinstance(filter, GpuFilter):
            gf = filter
        elif filter is not None:
            labs = gpu.labels()
            keep = np.fromiter((bool(filter(int(x))) for x in labs), dtype=bool, count=labs.shape[0])
            gf = gpu.make_filter(labs[keep])
        try:
            labels, dists, _, short = gpu.search(arr, k, self.default_ef, filter=gf)
        finally:
            if gf is not None and gf is not filter:
                gf.free()
        if short:
            raise RuntimeError("Cannot return the results in a contiguous 2D array. Probably ef or M is too small")
        return labels, dists

    def cpu_handle(self):
        return self._h

    def __del__(self):
        try:
            if self._gpu is not None:
                self._gpu.free()
            if self._h is not None:
                _lib.load().hg_hnsw_free(self._h)
        except Exception:
            pass

    def __repr__(self):
        return f"<hnsw_rabitq_amd.Index(space='{self.space}', dim={self.dim})>"


class BFIndex:
    """hnswlib.BFIndex mirror (bindings.cpp:725-929, 1011-1023) on the GPU."""

    def __init__(self, space, dim):
        self.space_id, self.normalize = _space(space)
        self.space = space
        self.dim = int(dim)
        self._h = None
        self._labels = set()
        self._cur_l = 0
        self.num_threads = os.cpu_count() or 1
        self.device = 0
        self.precision = "fp32"
        self.rerank_c = 4096

    def set_precision(self, precision, rerank_c=4096):
        """Distance precision of knn_query: "fp32" (exact scan), "bf16" or "fp8" (MFMA distance
        GEMM, rerank_c candidates re-ranked in exact FP32, every query the error bound cannot
        certify answered by the exact scan: the lists equal the "fp32" ones).  No counterpart in
        the reference; the reduced precisions need dim % 64 == 0."""
        if precision not in _PRECISIONS:
            raise ValueError("precision must be one of " + ", ".join(_PRECISIONS))
        if precision != "fp32" and self.dim % 64 != 0:
            raise ValueError("the bf16 / fp8 distance GEMM needs dim to be a multiple of 64")
        self.precision = precision
        self.rerank_c = int(rerank_c)

    def lowp_report(self):
        """What the last bf16 / fp8 search did besides the re-rank (hg_bf_lowp_report)."""
        out = np.zeros(9, np.float64)
        _lib.call("hg_bf_lowp_report", self._h, _ptr(out))
        keys = ("exact_overflow", "exact_uncertified", "near_ties", "near_tie_disagree", "boundary_near_ties",
                "boundary_disagree", "max_bound", "certified", "reranked_rows")
        return {k: (float(v) if k == "max_bound" else int(v)) for k, v in zip(keys, out)}

    def init_index(self, max_elements):
        if self._h is not None:
            raise RuntimeError("The index is already initiated.")
        h = ctypes.c_void_p()
        _lib.call("hg_bf_create", self.space_id, self.dim, int(max_elements), self.device, ctypes.byref(h))
        self._h = h
        self._labels = set()

    def add_items(self, data, ids=None):
        arr = _as_rows(data, self.dim)
        if self.normalize:
            arr = normalize_rows(arr)
        if ids is not None:
            lab = np.ascontiguousarray(np.asarray(ids, dtype=np.uint64).reshape(-1))
            if lab.shape[0] != arr.shape[0]:
                raise RuntimeError("The input label shape %d does not match the input data vector shape %d"
                                   % (lab.ndim, arr.shape[0]))
        else:
            # default labels continue a counter of every row ever passed in (bindings.cpp:812-822)
            lab = np.arange(self._cur_l, self._cur_l + arr.shape[0], dtype=np.uint64)
        try:
            _lib.call("hg_bf_add", self._h, _ptr(arr), arr.shape[0], _ptr(lab))
        except RuntimeError:
            # the rows placed before the index filled up stay added (bindings.cpp:811-822);
            # the default-label counter does not advance, as in the reference
            n = self.get_current_count()
            labs = np.empty(n, dtype=np.uint64)
            if n:
                _lib.call("hg_bf_labels", self._h, _ptr(labs), n)
            self._labels = set(int(x) for x in labs)
            raise
        self._cur_l += arr.shape[0]
        self._labels.update(int(x) for x in lab)

    def get_current_count(self):
        c = ctypes.c_size_t()
        _lib.call("hg_bf_count", self._h, ctypes.byref(c))
        return c.value

    def set_num_threads(self, n):
        self.num_threads = int(n)

    def knn_query(self, data, k=1, num_threads=-1, filter=None):
        """filter: a callable label -> bool, as in hnswlib (evaluated here once per element).
        Runs at the precision set by set_precision; a filtered query always takes the exact
        FP32 scan."""
        if filter is not None and not callable(filter):
            raise TypeError("filter must be a callable taking a label")
        arr = _as_rows(data, self.dim)
        if self.normalize:
            arr = normalize_rows(arr)
        k = int(k)
        nq = arr.shape[0]
        labels = np.empty((nq, k), dtype=np.uint64)
        dists = np.empty((nq, k), dtype=np.float32)
        if nq == 0:
            return labels, dists
        lib = _lib.load()
        if filter is None and self.precision != "fp32":
            code = lib.hg_bf_search_precision(self._h, _ptr(arr), nq, k, _PRECISIONS[self.precision], self.rerank_c,
                                              _ptr(labels), _ptr(dists))
        elif filter is None:
            code = lib.hg_bf_search(self._h, _ptr(arr), nq, k, _ptr(labels), _ptr(dists))
        else:
            labs = np.array(list(self._labels), dtype=np.uint64)
            keep = np.fromiter((bool(filter(int(x))) for x in labs), dtype=bool, count=labs.shape[0])
            allowed = np.ascontiguousarray(labs[keep])
            f = ctypes.c_void_p()
            _lib.call("hg_bf_filter_create", self._h, _ptr(allowed), allowed.shape[0], ctypes.byref(f))
            try:
                code = lib.hg_bf_search_filtered(self._h, _ptr(arr), nq, k, f, _ptr(labels), _ptr(dists))
            finally:
                lib.hg_bf_filter_free(f)
        if code == _lib.HG_ERR_SHORT:
            raise RuntimeError("Cannot return the results in a contiguous 2D array. There are not enough elements.")
        _lib.check(code)
        return labels, dists

    def delete_vector(self, label):
        """BruteforceSearch::removePoint (bruteforce.h:85-104); an unknown label is ignored."""
        _lib.call("hg_bf_delete", self._h, int(label))
        self._labels.discard(int(label))

    def save_index(self, path_to_index):
        _lib.call("hg_bf_save", self._h, str(path_to_index).encode())

    def load_index(self, path_to_index, max_elements=0):
        """BruteforceSearch::loadIndex: the capacity comes from the file (max_elements is ignored,
        as in the reference)."""
        if self._h is not None:
            print("Warning: Calling load_index for an already inited index. Old index is being deallocated.",
                  file=sys.stderr)
            _lib.load().hg_bf_free(self._h)
            self._h = None
        h = ctypes.c_void_p()
        _lib.call("hg_bf_load", self.space_id, self.dim, str(path_to_index).encode(), self.device, ctypes.byref(h))
        self._h = h
        n = self.get_current_count()
        labs = np.empty(n, dtype=np.uint64)
        if n:
            _lib.call("hg_bf_labels", self._h, _ptr(labs), n)
        self._labels = set(int(x) for x in labs)
        self._cur_l = n

    def __del__(self):
        try:
            if self._h is not None:
                _lib.load().hg_bf_free(self._h)
        except Exception:
            pass

    def __repr__(self):
        return f"<hnsw_rabitq_amd.BFIndex(space='{self.space}', dim={self.dim})>"
