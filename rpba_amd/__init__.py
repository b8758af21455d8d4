"""MI355X-native robust LM Schur-complement bundle adjustment (the global
"rbacore" solve of helmayer/RPBA) behind the C ABI of include/rpba.h.

``lib()`` loads the in-tree HIP library and fails loudly when it is missing:
there is no CPU fallback in the product path.
"""
import ctypes as C
import os

from . import capi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librpba_hip.so")

# every symbol include/rpba.h declares
EXPORTS = [
    "rpba_calc_plsa", "rpba_ctx_create", "rpba_ctx_run", "rpba_ctx_reset", "rpba_ctx_download",
    "rpba_ctx_set_timing", "rpba_ctx_kernel_time", "rpba_ctx_progress", "rpba_ctx_sizes", "rpba_ctx_destroy",
    "rpba_ctx_tuning", "rpba_dense_solve_batch", "rpba_ctx_set_graph",
    "rpba_batch_create", "rpba_batch_run", "rpba_batch_reset", "rpba_batch_download", "rpba_batch_set_graph",
    "rpba_batch_set_timing", "rpba_batch_kernel_time", "rpba_batch_spmv_bytes", "rpba_batch_tuning", "rpba_batch_sizes", "rpba_batch_destroy",
    "rpba_calc_plsa_batch",
    "rpba_lp_create", "rpba_lp_destroy", "rpba_lp_n_params", "rpba_lp_get_params", "rpba_lp_predict",
    "rpba_lp_gradient", "rpba_lp_train", "rpba_ctx_set_lp", "rpba_ctx_last_omega", "rpba_ctx_lp_sample",
    "rpba_read_bal", "rpba_write_bal", "rpba_free_block", "rpba_read_parameters", "rpba_version",
    "rpba_exteval_device", "rpba_exteval_device_max_points", "rpba_ctx_set_pcg_budget", "rpba_partition_images",
]

_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"HIP library {LIB_PATH} is not built; run `python -m rpba_amd.build`")
    L = C.CDLL(LIB_PATH)
    P = C.POINTER
    L.rpba_calc_plsa.argtypes = [P(capi.RpbaBlock), P(capi.RpbaFlags), capi.LOG_FN, C.c_void_p, P(capi.RpbaStats)]
    L.rpba_calc_plsa.restype = C.c_int
    L.rpba_ctx_create.argtypes = [P(C.c_void_p), P(capi.RpbaBlock), P(capi.RpbaFlags), capi.LOG_FN, C.c_void_p]
    L.rpba_ctx_create.restype = C.c_int
    L.rpba_ctx_run.argtypes = [C.c_void_p, C.c_int32, P(C.c_int32), P(C.c_int32)]
    L.rpba_ctx_run.restype = C.c_int
    L.rpba_ctx_reset.argtypes = [C.c_void_p]
    L.rpba_ctx_reset.restype = C.c_int
    L.rpba_ctx_download.argtypes = [C.c_void_p, P(capi.RpbaStats)]
    L.rpba_ctx_download.restype = C.c_int
    L.rpba_ctx_set_timing.argtypes = [C.c_void_p, C.c_int32]
    L.rpba_ctx_kernel_time.argtypes = [C.c_void_p, C.c_char_p, P(C.c_double), P(C.c_int64)]
    L.rpba_ctx_progress.argtypes = [C.c_void_p, P(C.c_int32), P(C.c_int32), P(C.c_double)]
    L.rpba_ctx_progress.restype = C.c_int
    L.rpba_ctx_sizes.argtypes = [C.c_void_p, P(C.c_int64)]
    L.rpba_ctx_tuning.argtypes = [C.c_void_p, C.c_char_p, C.c_int32]
    L.rpba_ctx_set_graph.argtypes = [C.c_void_p, C.c_int32]
    L.rpba_ctx_set_pcg_budget.argtypes = [C.c_void_p, C.c_int32]
    L.rpba_ctx_set_pcg_budget.restype = C.c_int
    L.rpba_ctx_tuning.restype = C.c_int
    L.rpba_dense_solve_batch.argtypes = [C.c_int32, P(C.c_int32), P(C.c_double), P(C.c_double), P(C.c_double),
                                         C.c_int32, C.c_int32]
    L.rpba_dense_solve_batch.restype = C.c_int
    L.rpba_batch_create.argtypes = [P(C.c_void_p), C.c_int32, P(capi.RpbaBlock), P(capi.RpbaFlags),
                                    capi.BATCH_LOG_FN, C.c_void_p]
    L.rpba_batch_create.restype = C.c_int
    L.rpba_batch_run.argtypes = [C.c_void_p, C.c_int32, P(C.c_int32), P(C.c_int32)]
    L.rpba_batch_run.restype = C.c_int
    L.rpba_batch_reset.argtypes = [C.c_void_p]
    L.rpba_batch_reset.restype = C.c_int
    L.rpba_batch_download.argtypes = [C.c_void_p, P(capi.RpbaStats), P(C.c_int32)]
    L.rpba_batch_download.restype = C.c_int
    L.rpba_batch_set_graph.argtypes = [C.c_void_p, C.c_int32]
    L.rpba_batch_set_timing.argtypes = [C.c_void_p, C.c_int32]
    L.rpba_batch_kernel_time.argtypes = [C.c_void_p, C.c_char_p, P(C.c_double), P(C.c_int64)]
    L.rpba_batch_spmv_bytes.argtypes = [C.c_void_p, P(C.c_double)]
    L.rpba_batch_tuning.argtypes = [C.c_void_p, C.c_char_p, C.c_int32]
    L.rpba_batch_sizes.argtypes = [C.c_void_p, P(C.c_int64)]
    L.rpba_batch_destroy.argtypes = [C.c_void_p]
    L.rpba_batch_destroy.restype = None
    L.rpba_calc_plsa_batch.argtypes = [C.c_int32, P(capi.RpbaBlock), P(capi.RpbaFlags), capi.BATCH_LOG_FN,
                                       C.c_void_p, P(capi.RpbaStats), P(C.c_int32)]
    L.rpba_calc_plsa_batch.restype = C.c_int
    G = P(capi.RpbaLpGraphs)
    L.rpba_lp_create.argtypes = [P(C.c_void_p), P(C.c_float), C.c_uint64]
    L.rpba_lp_create.restype = C.c_int
    L.rpba_lp_destroy.argtypes = [C.c_void_p]
    L.rpba_lp_destroy.restype = None
    L.rpba_lp_n_params.restype = C.c_int
    L.rpba_lp_get_params.argtypes = [C.c_void_p, P(C.c_float)]
    L.rpba_lp_predict.argtypes = [C.c_void_p, G, P(C.c_float)]
    L.rpba_lp_gradient.argtypes = [C.c_void_p, G, P(C.c_float), P(C.c_float), P(C.c_float)]
    L.rpba_lp_train.argtypes = [C.c_void_p, G, C.c_int32, C.c_float, P(C.c_float)]
    L.rpba_ctx_set_lp.argtypes = [C.c_void_p, C.c_void_p]
    L.rpba_ctx_last_omega.argtypes = [C.c_void_p, P(C.c_float)]
    L.rpba_ctx_lp_sample.argtypes = [C.c_void_p, C.c_int32, P(C.c_float), P(C.c_int32), P(C.c_float), P(C.c_int32),
                                     P(C.c_int32), P(C.c_float), P(C.c_int64)]
    L.rpba_ctx_destroy.argtypes = [C.c_void_p]
    L.rpba_ctx_destroy.restype = None
    L.rpba_read_bal.argtypes = [C.c_char_p, C.c_int32, P(capi.RpbaBlock)]
    L.rpba_read_bal.restype = C.c_int
    L.rpba_write_bal.argtypes = [C.c_char_p, P(capi.RpbaBlock), C.c_int32]
    L.rpba_write_bal.restype = C.c_int
    L.rpba_free_block.argtypes = [P(capi.RpbaBlock)]
    L.rpba_free_block.restype = None
    L.rpba_read_parameters.argtypes = [C.c_char_p, P(capi.RpbaFlags), C.c_char_p, C.c_int32, C.c_char_p, C.c_int32]
    L.rpba_read_parameters.restype = C.c_int
    L.rpba_version.restype = C.c_int
    vp = C.c_void_p
    L.rpba_exteval_device.argtypes = [C.c_int32, C.c_int32] + [vp] * 10 + [C.c_int32] * 3 + [vp, vp]
    L.rpba_exteval_device.restype = C.c_int
    L.rpba_exteval_device_max_points.restype = C.c_int
    L.rpba_partition_images.argtypes = [C.c_int32, C.c_int32, P(C.c_int32), P(C.c_int32), C.c_int32, C.c_int32,
                                        P(C.c_int32)]
    L.rpba_partition_images.restype = C.c_int
    _lib = L
    return L
