# coding: utf-8
"""ctypes loader of the beamforming libraries.

Mirrors beampower/core.py:7-169 of the reference: the same LIBRARIES table
layout, the same load-once cache and the same errors (NameError for an
unknown device, RuntimeError when the GPU library is missing, OSError when
the CPU library is missing).  The "gpu" entry is this package's gfx950
library, built from beampower_amd/csrc; it additionally exports the GPU
restatement of prestack_waveform_features and the bp_* extension API, and
every entry point returns an int status (0 = success).  The "cpu" entry
points at lib/beamform_cpu.so, which this package does not build: it is the
reference's own CPU engine, loaded only if the user places it there.
"""
import ctypes as ct
import os

DIRPATH_LIBRARIES = os.path.join(os.path.dirname(__file__), "lib")

_P_F = ct.POINTER(ct.c_float)
_P_I = ct.POINTER(ct.c_int)
_SZ = ct.c_size_t

LIBRARIES = {
    "cpu": {
        "filepath_library": os.path.join(DIRPATH_LIBRARIES, "beamform_cpu.so"),
        "is_loaded": False,
        "lib": None,
        # core.py:14-49 of the reference
        "beamform_argtypes": [_P_F, _P_I, _P_F, _SZ, _SZ, _SZ, _SZ, ct.c_int, ct.c_int, _P_F],
        "beamform_differential_argtypes": [_P_F, _P_I, _P_F, _SZ, _SZ, _SZ, _SZ, ct.c_int,
                                           _P_F],
        "beamform_max_argtypes": [_P_F, _P_I, _P_F, _SZ, _SZ, _SZ, _SZ, ct.c_int, ct.c_int,
                                  _P_F, _P_I],
        "prestack_waveform_features_argtypes": [_P_F, _P_F, _SZ, _SZ, _SZ, _SZ, ct.c_int, _P_F],
    },
    "gpu": {
        "filepath_library": os.path.join(DIRPATH_LIBRARIES, "beamform_gpu.so"),
        "is_loaded": False,
        "lib": None,
        # core.py:65-87 of the reference
        "beamform_argtypes": [_P_F, _P_I, _P_F, _SZ, _SZ, _SZ, _SZ, ct.c_int, _P_F],
        "beamform_max_argtypes": [_P_F, _P_I, _P_F, _SZ, _SZ, _SZ, _SZ, ct.c_int, _P_F, _P_I],
        # the CPU library's prestack and differential signatures
        # (core.py:26-49), on the GPU
        "prestack_waveform_features_argtypes": [_P_F, _P_F, _SZ, _SZ, _SZ, _SZ, ct.c_int, _P_F],
        "beamform_differential_argtypes": [_P_F, _P_I, _P_F, _SZ, _SZ, _SZ, _SZ, ct.c_int,
                                           _P_F],
    },
}

# Symbols the gpu library declares in include/beamform_gpu.h.
GPU_SYMBOLS = (
    "beamform", "beamform_max", "prestack_waveform_features", "beamform_differential",
    "bp_dropin_cache_clear", "bp_beamform_max_fused", "bp_beamform_fused",
    "bp_last_error", "bp_version",
    "bp_network_create", "bp_network_destroy", "bp_network_info", "bp_network_classes",
    "bp_network_create_ordered", "bp_network_source_order", "bp_kernel_shape",
    "bp_network_beamform_max", "bp_network_beamform",
    "bp_interleaved_to_planar", "bp_prestack_planar",
    "bp_parity_stride", "bp_quantize_planar",
    "bp_profiling", "bp_profiling_read",
    "bp_network_beamform_max_view", "bp_network_beamform_view", "bp_network_padding",
    "bp_network_beamform_max_noalloc", "bp_plan",
    "bp_network_beamform_differential", "bp_network_diff_table_bytes",
    "bp_pack_features",
    "bp_batch_create", "bp_batch_launch", "bp_batch_graph_launch", "bp_batch_tune",
    "bp_batch_get_config", "bp_batch_set_config", "bp_batch_destroy",
    "bp_network_prepare", "bp_detect_threshold", "bp_detect_threshold_type", "bp_detect_peaks",
    "bp_server_create", "bp_server_add_network", "bp_server_submit", "bp_server_wait",
    "bp_server_stats", "bp_server_destroy", "bp_server_last_error",
)


class ServerConfig(ct.Structure):
    """bp_server_config of include/beamform_gpu.h."""
    _fields_ = [("hbm_budget", ct.c_size_t), ("max_batch", ct.c_int), ("fmt", ct.c_int),
                ("in_flight", ct.c_int)]


class RequestInfo(ct.Structure):
    """bp_request_info of include/beamform_gpu.h."""
    _fields_ = [("status", ct.c_int), ("batch_windows", ct.c_int), ("submit_s", ct.c_double),
                ("admit_s", ct.c_double), ("ready_s", ct.c_double), ("launch_s", ct.c_double),
                ("gpu_done_s", ct.c_double), ("done_s", ct.c_double)]


class ServerCounters(ct.Structure):
    """bp_server_counters of include/beamform_gpu.h."""
    _fields_ = [("pending", ct.c_longlong), ("ready", ct.c_longlong),
                ("in_flight_batches", ct.c_longlong), ("completed", ct.c_longlong),
                ("batches", ct.c_longlong), ("batched_windows", ct.c_longlong),
                ("arena_bytes", ct.c_size_t), ("arena_used", ct.c_size_t),
                ("arena_peak", ct.c_size_t)]


class PlanRequest(ct.Structure):
    """bp_plan_request of include/beamform_gpu.h."""
    _fields_ = [("n_sources", ct.c_size_t), ("n_stations", ct.c_size_t),
                ("n_phases", ct.c_size_t), ("n_samples", ct.c_size_t), ("fmt", ct.c_int),
                ("reduce", ct.c_int), ("hbm_budget", ct.c_size_t),
                ("time_delays", ct.c_void_p), ("weights_sources", ct.c_void_p),
                ("order", ct.c_int)]


class PlanResult(ct.Structure):
    """bp_plan_result of include/beamform_gpu.h."""
    _fields_ = [("source_tile", ct.c_int), ("time_tile", ct.c_int),
                ("lds_window_limit", ct.c_int), ("tiles_per_class", ct.c_int * 3),
                ("window_per_class", ct.c_int * 2), ("similarity_order", ct.c_int),
                ("pad_before", ct.c_longlong), ("pad_after", ct.c_longlong),
                ("lds_bytes", ct.c_int * 2), ("workgroups_per_cu", ct.c_int * 2),
                ("network_bytes", ct.c_size_t), ("feature_bytes", ct.c_size_t),
                ("output_bytes", ct.c_size_t), ("call_bytes", ct.c_size_t), ("fits", ct.c_int)]


class FeatureViewC(ct.Structure):
    """bp_feature_view of include/beamform_gpu.h."""
    _fields_ = [("data", ct.c_void_p), ("fmt", ct.c_int), ("n_samples", ct.c_size_t),
                ("row_stride", ct.c_size_t), ("pad_before", ct.c_size_t),
                ("station_scale", ct.c_void_p)]


class BatchConfig(ct.Structure):
    """bp_batch_config of include/beamform_gpu.h."""
    _fields_ = [("time_group", ct.c_int), ("stage_buffers", ct.c_int),
                ("blocks_per_cu", ct.c_int), ("ms", ct.c_float)]


def _declare_gpu_extension(lib):
    vp = ct.c_void_p
    lib.bp_last_error.restype = ct.c_char_p
    lib.bp_last_error.argtypes = []
    lib.bp_version.restype = ct.c_char_p
    lib.bp_version.argtypes = []
    lib.bp_network_create.argtypes = [vp, vp, _SZ, _SZ, _SZ, ct.POINTER(vp)]
    lib.bp_network_destroy.argtypes = [vp]
    lib.bp_dropin_cache_clear.argtypes = []
    lib.bp_beamform_max_fused.argtypes = [_P_F, _P_F, _P_I, _P_F, _SZ, _SZ, _SZ, _SZ, _SZ,
                                          ct.c_int, _P_F, _P_I]
    lib.bp_beamform_fused.argtypes = [_P_F, _P_F, _P_I, _P_F, _SZ, _SZ, _SZ, _SZ, _SZ,
                                      ct.c_int, _P_F]
    lib.bp_network_create_ordered.argtypes = [vp, vp, _SZ, _SZ, _SZ, ct.c_int,
                                              ct.POINTER(vp)]
    lib.bp_network_source_order.argtypes = [vp, _P_I, vp]
    lib.bp_kernel_shape.argtypes = [_P_I, _P_I, _P_I]
    lib.bp_network_info.argtypes = [vp, _P_I, _P_I, _P_I]
    lib.bp_network_classes.argtypes = [vp, _P_I, _P_I]
    lib.bp_network_beamform_max.argtypes = [vp, vp, ct.c_int, vp, _SZ, ct.c_int, vp, vp, vp,
                                            vp]
    lib.bp_network_beamform.argtypes = [vp, vp, _SZ, ct.c_int, _SZ, _SZ, vp, vp]
    lib.bp_interleaved_to_planar.argtypes = [vp, _SZ, _SZ, _SZ, vp, vp]
    lib.bp_prestack_planar.argtypes = [vp, vp, _SZ, _SZ, _SZ, _SZ, vp, vp]
    lib.bp_parity_stride.argtypes = [_SZ]
    lib.bp_quantize_planar.argtypes = [vp, _SZ, _SZ, _SZ, ct.c_int, vp, vp, vp]
    lib.bp_profiling.argtypes = [ct.c_int]
    lib.bp_profiling_read.argtypes = [ct.POINTER(ct.c_double), ct.POINTER(ct.c_longlong)]
    lib.bp_network_beamform_max_view.argtypes = [vp, ct.POINTER(FeatureViewC), ct.c_int, vp,
                                                 vp, vp, vp]
    lib.bp_network_padding.argtypes = [vp, ct.POINTER(_SZ), ct.POINTER(_SZ)]
    lib.bp_plan.argtypes = [ct.POINTER(PlanRequest), ct.POINTER(PlanResult)]
    lib.bp_network_beamform_max_noalloc.argtypes = [vp, ct.POINTER(FeatureViewC), ct.c_int, vp,
                                                    vp, vp]
    lib.bp_network_beamform_view.argtypes = [vp, ct.POINTER(FeatureViewC), ct.c_int, _SZ, _SZ,
                                             vp, vp]
    lib.bp_network_beamform_differential.argtypes = [vp, ct.POINTER(FeatureViewC), vp, vp]
    lib.bp_network_diff_table_bytes.argtypes = [vp, ct.POINTER(_SZ), ct.POINTER(_SZ), _P_I]
    lib.bp_pack_features.argtypes = [vp, _SZ, _SZ, _SZ, ct.c_int, _SZ, _SZ, vp, vp, vp]
    lib.bp_batch_create.argtypes = [ct.c_int, vp, ct.POINTER(FeatureViewC), vp, vp, vp, vp,
                                    ct.POINTER(vp)]
    lib.bp_batch_launch.argtypes = [vp, vp]
    lib.bp_batch_graph_launch.argtypes = [vp, vp]
    lib.bp_batch_tune.argtypes = [vp, ct.c_int, ct.POINTER(BatchConfig), ct.c_int, _P_I,
                                  ct.POINTER(BatchConfig), vp]
    lib.bp_batch_get_config.argtypes = [vp, ct.POINTER(BatchConfig)]
    lib.bp_batch_set_config.argtypes = [vp, ct.POINTER(BatchConfig)]
    lib.bp_batch_destroy.argtypes = [vp]
    lib.bp_network_prepare.argtypes = [vp, ct.c_int, vp]
    lib.bp_detect_threshold.argtypes = [vp, _SZ, _SZ, ct.c_double, ct.c_double, vp, vp]
    lib.bp_detect_threshold_type.argtypes = [vp, _SZ, _SZ, ct.c_double, ct.c_double, ct.c_int,
                                             vp, vp]
    lib.bp_detect_peaks.argtypes = [vp, vp, _SZ, _SZ, vp, _SZ, ct.POINTER(_SZ), vp]
    lib.bp_server_create.argtypes = [ct.POINTER(ServerConfig), ct.POINTER(vp)]
    lib.bp_server_add_network.argtypes = [vp, vp, vp, vp, _SZ, _SZ, _SZ, _SZ, _P_I]
    lib.bp_server_submit.argtypes = [vp, ct.c_int, vp, _SZ, ct.c_int, vp, vp,
                                     ct.POINTER(ct.c_longlong)]
    lib.bp_server_wait.argtypes = [vp, ct.c_longlong, ct.c_double, ct.POINTER(RequestInfo)]
    lib.bp_server_stats.argtypes = [vp, ct.POINTER(ServerCounters)]
    lib.bp_server_destroy.argtypes = [vp]
    lib.bp_server_last_error.restype = ct.c_char_p
    lib.bp_server_last_error.argtypes = []
    for name in GPU_SYMBOLS:
        if name not in ("bp_last_error", "bp_version", "bp_parity_stride",
                        "bp_server_last_error"):
            getattr(lib, name).restype = ct.c_int
    lib.bp_parity_stride.restype = _SZ


def load_library(device="cpu"):
    """Load the library for `device` ("cpu" or "gpu") once and cache it.

    Raises NameError for an unknown device, RuntimeError if the GPU library
    is missing and OSError if the CPU library is missing
    (beampower/core.py:92-169)."""
    device_name = device.lower()
    if device_name not in LIBRARIES:
        raise NameError(f"Device should be cpu or gpu, not {device_name}")
    info = LIBRARIES[device_name]
    path = info["filepath_library"]
    if os.path.exists(path):
        if info["is_loaded"] is True:
            return info["lib"]
        lib = ct.cdll.LoadLibrary(path)
        lib.beamform.argtypes = info["beamform_argtypes"]
        lib.beamform_max.argtypes = info["beamform_max_argtypes"]
        lib.prestack_waveform_features.argtypes = info["prestack_waveform_features_argtypes"]
        lib.beamform_differential.argtypes = info["beamform_differential_argtypes"]
        if device_name == "gpu":
            for name in ("beamform", "beamform_max", "prestack_waveform_features",
                         "beamform_differential"):
                getattr(lib, name).restype = ct.c_int
            _declare_gpu_extension(lib)
        info["is_loaded"] = True
        info["lib"] = lib
        return lib
    if device_name == "gpu":
        raise RuntimeError(
            f"GPU library not available. The shared object {path} does not exist.\n"
            "Build it with `python -m beampower_amd.build` (hipcc, gfx950).")
    raise OSError(
        f"CPU library is required but not found at {path}.\n"
        "This package ships the gfx950 engine only; place the reference's "
        "beamform_cpu.so there to use device='cpu'.")


def last_error(lib=None):
    lib = lib or load_library("gpu")
    return lib.bp_last_error().decode(errors="replace")


def check(status, what, lib=None):
    """Raise RuntimeError for a non-zero status of a gpu-library call."""
    if status != 0:
        raise RuntimeError(f"{what} failed with status {status}: {last_error(lib)}")
