"""ctypes binding of lib/libasz_gpu.so (declarations in include/asz_gpu.h).

The library is the product: every compute call goes to its HIP kernels.  If it
is not built, importing this module raises; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libasz_gpu.so")
BIN_DIR = os.path.join(HERE, "bin")

ASZ_OK = 0
ASZ_TSEL_GPU, ASZ_TSEL_EXACT, ASZ_TSEL_INJECT, ASZ_TSEL_GRAD = 0, 1, 2, 3
MAX_LAYERS = 63

_P = ctypes.c_void_p
_U64 = ctypes.c_uint64
_U64P = ctypes.POINTER(ctypes.c_uint64)


class Params(ctypes.Structure):
    _fields_ = [("rel_eb", ctypes.c_float), ("pivot_eb_ratio", ctypes.c_float), ("eb_ratio", ctypes.c_float),
                ("slope_ratio", ctypes.c_float), ("block_pivot_ratio", ctypes.c_uint64),
                ("t_precision", ctypes.c_int32), ("stat_eb_ratio", ctypes.c_float),
                ("h1_block_sampling_num", ctypes.c_uint64), ("t_min", ctypes.c_float), ("t_max", ctypes.c_float),
                ("default_slope_eb_span", ctypes.c_float), ("default_slope_invalid_ratio_threshold", ctypes.c_float),
                ("default_slope_fallback", ctypes.c_float), ("default_slope", ctypes.c_float),
                ("quintic", ctypes.c_int32)]


class Plan(ctypes.Structure):
    _fields_ = [("layers", ctypes.c_int32), ("axis", ctypes.c_uint8 * MAX_LAYERS),
                ("order", ctypes.c_uint8 * MAX_LAYERS)]


class Stats(ctypes.Structure):
    _fields_ = [("range_ms", ctypes.c_double), ("preprocess_ms", ctypes.c_double), ("tsel_ms", ctypes.c_double),
                ("rounds_ms", ctypes.c_double), ("huffman_ms", ctypes.c_double),
                ("huffman_host_ms", ctypes.c_double), ("total_gpu_ms", ctypes.c_double),
                ("wall_ms", ctypes.c_double), ("max_err", ctypes.c_double), ("sq_err_sum", ctypes.c_double),
                ("stream_bytes", ctypes.c_uint64), ("t_count", ctypes.c_uint64), ("t_fallbacks", ctypes.c_uint64),
                ("layers", ctypes.c_uint64), ("qb_counts", ctypes.c_uint64 * (MAX_LAYERS + 1)),
                ("pivot_ms", ctypes.c_double), ("tsel_round_ms", ctypes.c_double * MAX_LAYERS),
                ("round_ms", ctypes.c_double * MAX_LAYERS), ("hist_ms", ctypes.c_double),
                ("chunk_ms", ctypes.c_double), ("encode_ms", ctypes.c_double), ("decode_ms", ctypes.c_double),
                ("codebook_wait_ms", ctypes.c_double), ("codebook_max_ms", ctypes.c_double),
                ("layout_ms", ctypes.c_double), ("plan_cache_hit", ctypes.c_uint64),
                ("coarse_rounds", ctypes.c_uint64), ("coarse_ms", ctypes.c_double),
                ("graph_launched", ctypes.c_uint64), ("graph_captured", ctypes.c_uint64)]

class BatchField(ctypes.Structure):
    _fields_ = [("min_val", ctypes.c_float), ("max_val", ctypes.c_float), ("default_slope", ctypes.c_float),
                ("pad", ctypes.c_float), ("max_err", ctypes.c_double), ("sq_err_sum", ctypes.c_double),
                ("t_fallbacks", ctypes.c_uint64)]


class PlanReport(ctypes.Structure):
    _fields_ = [("layers", ctypes.c_uint32), ("searched", ctypes.c_uint32),
                ("best_dist", ctypes.c_float * (MAX_LAYERS + 1)), ("best_axis", ctypes.c_int8 * (MAX_LAYERS + 1)),
                ("best_order", ctypes.c_int8 * (MAX_LAYERS + 1)), ("statistics_s", ctypes.c_double),
                ("sample_prediction_s", ctypes.c_double), ("lossless_estimation_s", ctypes.c_double),
                ("total_s", ctypes.c_double)]


class FamilyOpts(ctypes.Structure):
    _fields_ = [("suffix_layers", ctypes.c_uint64), ("sample_blocks", ctypes.c_uint64),
                ("sample_point_limit", ctypes.c_uint64), ("sample_fraction_limit", ctypes.c_float),
                ("fine_eb_threshold", ctypes.c_float), ("coarse_probe_bpr", ctypes.c_uint64),
                ("fine_probe_bpr", ctypes.c_uint64)]


class FamilyReport(ctypes.Structure):
    _fields_ = [("suffix_layers", ctypes.c_uint32), ("blocks", ctypes.c_uint32),
                ("sampled_block_points", ctypes.c_uint64), ("sample_point_budget", ctypes.c_uint64),
                ("probe_block_pivot_ratio", ctypes.c_uint64), ("points", ctypes.c_uint64),
                ("cubic_mse", ctypes.c_longdouble), ("quintic_mse", ctypes.c_longdouble),
                ("cubic_psnr", ctypes.c_longdouble), ("quintic_psnr", ctypes.c_longdouble),
                ("cubic_h1", ctypes.c_longdouble), ("quintic_h1", ctypes.c_longdouble),
                ("score_gain", ctypes.c_longdouble), ("quintic", ctypes.c_int32), ("time_s", ctypes.c_double)]


ASZ_OPT_KERNEL_EVENTS = 1
ASZ_OPT_STREAM_ON_DEVICE = 2
ASZ_OPT_TIMED_ROUND = 3
ASZ_OPT_BATCH_FIELDS = 4
ASZ_OPT_WIDE_CODES = 5
ASZ_OPT_COARSE_ROUNDS = 6
ASZ_OPT_GRAPHS = 7


class RoundTune(ctypes.Structure):
    _fields_ = [("seg", ctypes.c_uint8), ("xsh", ctypes.c_uint8), ("rpw", ctypes.c_uint16)]


class TuneReport(ctypes.Structure):
    _fields_ = [("rounds_tuned", ctypes.c_uint32), ("trials", ctypes.c_uint32), ("default_ms", ctypes.c_double),
                ("tuned_ms", ctypes.c_double)]


# exported entry points (must match include/asz_gpu.h)
EXPORTS = ("asz_ctx_create", "asz_ctx_destroy", "asz_last_error", "asz_params_default", "asz_get_stats",
           "asz_device_alloc", "asz_device_free", "asz_copy_to_device", "asz_copy_to_host", "asz_geometry",
           "asz_range_f32", "asz_default_slope", "asz_compress", "asz_compress_host", "asz_decompress",
           "asz_decompress_host", "asz_decompress_device", "asz_huffman_codebook", "asz_zstd_compress", "asz_zstd_compress_mt", "asz_zstd_decompress",
           "asz_free",
           "asz_round_layout", "asz_set_option", "asz_stream_wait", "asz_compress_field", "asz_prepare_round", "asz_round_energy",
           "asz_round_energy_adjoint", "asz_compress_batch", "asz_generate_plan",
           "asz_select_family", "asz_range_f32_noalloc", "asz_autotune", "asz_tune_lookup", "asz_tune_save",
           "asz_tune_load")

_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
    L = ctypes.CDLL(LIB_PATH)
    L.asz_ctx_create.argtypes = [ctypes.c_int, ctypes.POINTER(_P)]
    L.asz_ctx_destroy.argtypes = [_P]
    L.asz_ctx_destroy.restype = None
    L.asz_last_error.argtypes = [_P]
    L.asz_last_error.restype = ctypes.c_char_p
    L.asz_params_default.argtypes = [ctypes.POINTER(Params)]
    L.asz_params_default.restype = None
    L.asz_get_stats.argtypes = [_P, ctypes.POINTER(Stats)]
    L.asz_set_option.argtypes = [_P, ctypes.c_int, ctypes.c_int64]
    L.asz_stream_wait.argtypes = [_P, ctypes.c_void_p]
    L.asz_compress_field.argtypes = [_P, _P, _P, ctypes.POINTER(Plan), ctypes.POINTER(Params), ctypes.c_int,
                                     ctypes.POINTER(_P), _U64P, ctypes.POINTER(ctypes.c_float),
                                     ctypes.POINTER(ctypes.c_float)]
    L.asz_device_alloc.argtypes = [_P, _U64, ctypes.POINTER(_P)]
    L.asz_device_free.argtypes = [_P, _P]
    L.asz_copy_to_device.argtypes = [_P, _P, _P, _U64]
    L.asz_copy_to_host.argtypes = [_P, _P, _P, _U64]
    L.asz_geometry.argtypes = [_P, ctypes.POINTER(Plan), _U64, _P, _P, _P]
    L.asz_range_f32.argtypes = [_P, _P, _U64, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)]
    L.asz_range_f32_noalloc.argtypes = [_P, _P, _U64, ctypes.POINTER(ctypes.c_float),
                                        ctypes.POINTER(ctypes.c_float)]
    L.asz_default_slope.argtypes = [_P, _P, _P, ctypes.POINTER(Params), ctypes.c_float, ctypes.c_float,
                                    ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float),
                                    ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
    L.asz_compress.argtypes = [_P, _P, _P, ctypes.POINTER(Plan), ctypes.POINTER(Params), ctypes.c_float,
                               ctypes.c_float, ctypes.c_int, _P, ctypes.POINTER(_P), _U64P, _P, _P]
    L.asz_compress_host.argtypes = [_P, _P, _P, ctypes.POINTER(Plan), ctypes.POINTER(Params), ctypes.c_int, _P,
                                    ctypes.POINTER(_P), _U64P, _P, _P, _P]
    L.asz_decompress.argtypes = [_P, _P, _U64, _P, ctypes.POINTER(Params), _P, _P, _P]
    L.asz_decompress_host.argtypes = [_P, _P, _U64, _P, ctypes.POINTER(Params), _P]
    L.asz_decompress_device.argtypes = [_P, _P, _U64, _P, ctypes.POINTER(Params), _P]
    L.asz_huffman_codebook.argtypes = [_P, _U64, ctypes.c_int64, _U64, ctypes.c_int, _P, _U64, _U64P, _P, _P,
                                       _U64P]
    L.asz_zstd_compress.argtypes = [_P, _U64, ctypes.POINTER(_P), _U64P]
    L.asz_zstd_compress_mt.argtypes = [_P, _U64, ctypes.c_int, ctypes.c_uint32, ctypes.POINTER(_P), _U64P,
                                       ctypes.POINTER(ctypes.c_uint32)]
    L.asz_zstd_decompress.argtypes = [_P, _U64, ctypes.POINTER(_P), _U64P]
    L.asz_prepare_round.argtypes = [_P, _P, _P, ctypes.POINTER(Plan), ctypes.POINTER(Params), ctypes.c_float,
                                    ctypes.c_float, ctypes.c_int, _P]
    L.asz_round_energy.argtypes = [_P, _P, _P, ctypes.POINTER(Plan), ctypes.POINTER(Params), ctypes.c_int,
                                   ctypes.c_int, _P, _P]
    L.asz_round_energy_adjoint.argtypes = [_P, _P, _P, ctypes.POINTER(Plan), ctypes.POINTER(Params), ctypes.c_int,
                                           ctypes.c_int, _P, _P, _P]
    L.asz_compress_batch.argtypes = [_P, ctypes.c_uint32, _P, _P, _P, ctypes.POINTER(Params), ctypes.c_int,
                                     ctypes.c_uint32, _P, _P, _P]
    L.asz_generate_plan.argtypes = [_P, _P, _P, ctypes.POINTER(Params), ctypes.c_float, ctypes.c_float,
                                    ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(Plan),
                                    ctypes.POINTER(PlanReport)]
    L.asz_autotune.argtypes = [_P, _P, _P, ctypes.POINTER(Plan), ctypes.POINTER(Params), ctypes.c_uint32,
                               ctypes.POINTER(TuneReport)]
    L.asz_tune_lookup.argtypes = [_P, _P, ctypes.POINTER(Plan), _U64, _U64, ctypes.POINTER(RoundTune)]
    L.asz_tune_save.argtypes = [_P, ctypes.c_char_p]
    L.asz_tune_load.argtypes = [_P, ctypes.c_char_p]
    L.asz_select_family.argtypes = [_P, _P, _P, ctypes.POINTER(Plan), ctypes.POINTER(Params), ctypes.c_float,
                                    ctypes.c_float, ctypes.POINTER(FamilyOpts), ctypes.POINTER(FamilyReport)]
    L.asz_free.argtypes = [_P]
    L.asz_free.restype = None
    _lib = L
    return L
