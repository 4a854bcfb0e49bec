"""ctypes binding of libhnswgpu.so (the C ABI declared in include/hnswgpu.h).

The library is built in-tree by hnsw_rabitq_amd/build.py.  Loading fails loudly when it is
missing: there is no CPU fallback for the query path.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhnswgpu.so")

HG_OK, HG_ERR_INVALID, HG_ERR_RUNTIME, HG_ERR_HIP, HG_ERR_SHORT = 0, 1, 2, 3, 4
SPACE_L2, SPACE_IP = 0, 1
PREC_FP32, PREC_BF16, PREC_FP8 = 0, 1, 2
GEMM_VARIANTS = 5  # HG_GEMM_VARIANTS
VEC_F32, VEC_U8, VEC_AUTO = 0, 1, 1

_vp = ctypes.c_void_p
_sz = ctypes.c_size_t
_i32 = ctypes.c_int
_pvp = ctypes.POINTER(ctypes.c_void_p)

# name -> argument types (all functions return int except hg_last_error)
SIGNATURES = {
    "hg_hnsw_create": [_i32, _sz, _sz, _sz, _sz, _sz, _pvp],
    "hg_hnsw_load": [_i32, _sz, ctypes.c_char_p, _sz, _pvp],
    "hg_hnsw_add_items": [_vp, _vp, _sz, _vp, _i32],
    "hg_hnsw_save": [_vp, ctypes.c_char_p],
    "hg_hnsw_serialize": [_vp, _vp, _sz, ctypes.POINTER(_sz)],
    "hg_hnsw_info": [_vp, _vp],
    "hg_hnsw_set_ef": [_vp, _sz],
    "hg_hnsw_get_ef": [_vp, ctypes.POINTER(_sz)],
    "hg_hnsw_free": [_vp],
    "hg_hnsw_mark_deleted": [_vp, ctypes.c_uint64],
    "hg_hnsw_unmark_deleted": [_vp, ctypes.c_uint64],
    "hg_normalize_rows": [_vp, _vp, _sz, _sz],
    "hg_gpu_index_create": [_vp, _i32, _pvp],
    "hg_gpu_index_free": [_vp],
    "hg_gpu_index_bytes": [_vp, ctypes.POINTER(_sz)],
    "hg_gpu_set_vector_code": [_vp, _i32],
    "hg_gpu_vector_code": [_vp, ctypes.POINTER(_i32)],
    "hg_gpu_search": [_vp, _vp, _sz, _sz, _sz, _vp, _vp, _vp],
    "hg_gpu_search_device": [_vp, _vp, _sz, _vp, _vp, _sz, _sz, _sz, _vp, _vp, _vp, _vp, _vp],
    "hg_gpu_search_plan": [_vp, _sz, _sz, _sz, _vp],
    "hg_gpu_autotune": [_vp, _vp, _sz, _sz, _sz, _vp],
    "hg_gpu_set_plan": [_vp, _sz, _sz, ctypes.c_uint32, _i32],
    "hg_gpu_id_labels": [_vp, _vp, _sz, _vp, _vp],
    "hg_hnsw_build_gpu": [_vp, _vp, _sz, _vp, _i32, _i32],
    "hg_hnsw_build_gpu_upper_min_rows": [_sz],
    "hg_gpu_index_save": [_vp, ctypes.c_char_p],
    "hg_gpu_index_load": [ctypes.c_char_p, _i32, _vp],
    "hg_hnsw_add_items_upper": [_vp, _vp, _sz, _vp, _i32],
    "hg_bf_delete": [_vp, ctypes.c_uint64],
    "hg_bf_labels": [_vp, _vp, _sz],
    "hg_bf_save": [_vp, ctypes.c_char_p],
    "hg_bf_load": [_i32, _sz, ctypes.c_char_p, _i32, _vp],
    "hg_bf_filter_create": [_vp, _vp, _sz, _vp],
    "hg_bf_filter_free": [_vp],
    "hg_bf_search_filtered": [_vp, _vp, _sz, _sz, _vp, _vp, _vp],
    "hg_bf_search_device_filtered": [_vp, _vp, _sz, _sz, _vp, _vp, _vp, _vp],
    "hg_gpu_filter_create": [_vp, _vp, _sz, _vp],
    "hg_gpu_filter_free": [_vp],
    "hg_gpu_filter_count": [_vp, _vp],
    "hg_gpu_search_filtered": [_vp, _vp, _sz, _sz, _sz, _vp, _vp, _vp, _vp],
    "hg_gpu_search_device_filtered": [_vp, _vp, _sz, _vp, _vp, _sz, _sz, _sz, _vp, _vp, _vp, _vp, _vp, _vp],
    "hg_gpu_search_status": [_vp, _vp, _vp],
    "hg_gpu_search_device_hopcap": [_vp, _vp, _sz, _vp, _sz, _sz, _vp, _vp, _vp, _vp],
    "hg_gpu_search_device_noalloc": [_vp, _vp, _sz, _vp, _vp, _sz, _sz, _sz, _vp, _vp, _vp, _vp, _vp, _vp],
    "hg_gpu_docs_create": [_vp, _vp, _vp, _sz, _vp],
    "hg_gpu_docs_free": [_vp],
    "hg_gpu_search_multivector": [_vp, _vp, _sz, _vp, _sz, _sz, _vp, _sz, _vp, _vp, _vp],
    "hg_gpu_search_multivector_device": [_vp, _vp, _sz, _vp, _sz, _sz, _vp, _sz, _vp, _vp, _vp, _vp, _vp,
                                         ctypes.c_uint32, _vp],
    "hg_gpu_search_epsilon": [_vp, _vp, _sz, ctypes.c_float, _sz, _sz, _vp, _sz, _vp, _vp, _vp],
    "hg_gpu_search_epsilon_device": [_vp, _vp, _sz, ctypes.c_float, _sz, _sz, _vp, _sz, _vp, _vp, _vp, _vp, _vp,
                                     _vp],
    "hg_gpu_refresh_deleted": [_vp, _vp],
    "hg_gpu_index_labels": [_vp, _vp, _sz, _vp],
    "hg_bf_create": [_i32, _sz, _sz, _i32, _pvp],
    "hg_bf_add": [_vp, _vp, _sz, _vp],
    "hg_bf_count": [_vp, ctypes.POINTER(_sz)],
    "hg_bf_free": [_vp],
    "hg_bf_search": [_vp, _vp, _sz, _sz, _vp, _vp],
    "hg_bf_search_device": [_vp, _vp, _sz, _sz, _i32, _sz, _vp, _vp, _vp, _vp],
    "hg_bf_autotune": [_vp, _vp, _sz, _sz, _i32, _sz, _vp, _vp, _vp, _vp, _vp],
    "hg_bf_gemm_variant": [_i32],
    "hg_bf_fp64_check": [_vp, _vp, _sz, _sz, _vp, _vp, ctypes.c_double, _vp, _vp],
    "hg_bf_search_precision": [_vp, _vp, _sz, _sz, _i32, _sz, _vp, _vp],
    "hg_bf_lowp_report": [_vp, _vp],
    "hg_stop_model_create": [_i32, _i32, _i32, _pvp],
    "hg_stop_model_free": [_vp],
    "hg_stop_model_param_count": [_vp, ctypes.POINTER(_sz)],
    "hg_stop_model_set_params": [_vp, _vp, _sz],
    "hg_stop_model_get_params": [_vp, _vp, _sz],
    "hg_stop_model_set_norm": [_vp, _vp, _vp, _sz],
    "hg_stop_model_config": [_vp, _i32, ctypes.c_float, _i32],
    "hg_gpu_collect_traces": [_vp, _vp, _sz, _sz, _sz, _i32, _i32, _vp, _vp, _vp, _vp, _i32, _vp],
    "hg_stop_model_config_extend": [_vp, _i32, ctypes.c_float],
    "hg_stop_model_forward": [_vp, _vp, _sz, _vp, _vp],
    "hg_stop_model_train_step": [_vp, _vp, _vp, _sz, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                 ctypes.c_float, ctypes.POINTER(ctypes.c_float), _vp, _vp],
    "hg_device_count": [ctypes.POINTER(_i32)],
    "hg_set_device": [_i32],
    "hg_malloc": [_pvp, _sz],
    "hg_free": [_vp],
    "hg_memcpy_h2d": [_vp, _vp, _sz, _vp],
    "hg_memcpy_d2h": [_vp, _vp, _sz, _vp],
    "hg_memset": [_vp, _i32, _sz, _vp],
    "hg_copy_kernel": [_vp, _vp, _sz, _vp],
    "hg_stream_create": [_pvp],
    "hg_stream_destroy": [_vp],
    "hg_stream_synchronize": [_vp],
    "hg_device_synchronize": [],
    "hg_timer_create": [_pvp],
    "hg_timer_destroy": [_vp],
    "hg_timer_start": [_vp, _vp],
    "hg_timer_stop": [_vp, _vp, ctypes.POINTER(ctypes.c_float)],
    "hg_graph_capture_search": [_vp, _vp, _sz, _vp, _vp, _sz, _sz, _sz, _vp, _vp, _vp, _vp, _vp, _pvp],
    "hg_graph_launch": [_vp, _vp],
    "hg_graph_free": [_vp],
}

_lib = None


class HnswGpuError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


def load():
    """Load the library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: run hnsw_rabitq_amd/build.py (there is no CPU fallback)")
    lib = ctypes.CDLL(LIB_PATH)
    lib.hg_last_error.restype = ctypes.c_char_p
    lib.hg_last_error.argtypes = []
    for name, args in SIGNATURES.items():
        f = getattr(lib, name, None)
        if f is None:
            continue
        f.restype = ctypes.c_int
        f.argtypes = args
    _lib = lib
    return lib


def check(code):
    if code != HG_OK:
        msg = load().hg_last_error().decode(errors="replace")
        raise HnswGpuError(code, msg)
    return code


def call(name, *args):
    return check(getattr(load(), name)(*args))
