"""ctypes binding of libgreenpeas_amd.so (the C-ABI in include/greenpeas_amd.h).

The library is the product: every compile runs on the GPU through it.  There
is no CPU fallback -- if the library is missing the import of this module
raises, loudly.
"""

from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgreenpeas_amd.so")

GP_OK = 0
GP_INVALID_ARGUMENT = 1
GP_RUNTIME_ERROR = 2
# components gp_node_rates_backward_noalloc can hold (GP_NOALLOC_MAX_COMPONENTS)
NOALLOC_MAX_COMPONENTS = 1 << 20

# every entry point declared in include/greenpeas_amd.h
EXPORTS = (
    "gp_last_error", "gp_version",
    "gp_get_driver", "gp_get_driver_on_stream", "gp_driver_set_stream", "gp_driver_load_stepg",
    "gp_driver_destroy", "gp_compile_detector_error_model", "gp_driver_parse",
    "gp_dem_shape", "gp_dem_copy", "gp_dem_acquire", "gp_dem_release", "gp_dem_text", "gp_driver_params",
    "gp_driver_timings",
    "gp_driver_bench_device", "gp_debug_pnodes", "gp_debug_sensitivities", "gp_debug_stepg", "gp_debug_sweep_plan",
    "gp_build_stepg_host", "gp_debug_image_digest",
    "gp_batch_create", "gp_batch_destroy", "gp_batch_set_circuits", "gp_batch_compile",
    "gp_batch_stage_us", "gp_batch_device_us", "gp_batch_dem_shape", "gp_batch_dem_copy", "gp_batch_device_bytes", "gp_batch_layout_stats",
    "gp_estimate_device_bytes", "gp_prepare_circuit", "gp_prepared_destroy", "gp_prepared_device_bytes",
    "gp_batch_set_prepared", "gp_batch_bytes_after", "gp_device_memory", "gp_circuit_summary", "gp_set_tuning", "gp_get_tuning",
    "gp_set_tuning_class", "gp_get_tuning_class", "gp_clear_tuning_classes",
    "gp_grad_plan_create", "gp_grad_plan_destroy", "gp_grad_plan_sizes", "gp_grad_plan_dem_structure",
    "gp_grad_plan_device_arrays", "gp_grad_plan_copy_arrays", "gp_xor_merge_forward",
    "gp_xor_merge_backward", "gp_node_rates_forward", "gp_node_rates_backward",
    "gp_node_rates_backward_noalloc",
)

_lib = None


class DemView(ctypes.Structure):
    """gp_dem_view: a DEM's arrays in the driver's pinned result memory."""

    _fields_ = [
        ("probs", ctypes.c_void_p),
        ("offsets", ctypes.c_void_p),
        ("targets", ctypes.c_void_p),
        ("hashes", ctypes.c_void_p),
        ("num_errors", ctypes.c_uint32),
        ("num_detectors", ctypes.c_uint32),
        ("num_observables", ctypes.c_uint32),
        ("target_bytes", ctypes.c_uint32),
        ("num_targets", ctypes.c_uint64),
        ("token", ctypes.c_void_p),
    ]


class Stepg(ctypes.Structure):
    """gp_stepg: a caller's STEPG in the reference's layout."""

    _fields_ = [
        ("correlation_level", ctypes.c_int),
        ("num_qubits", ctypes.c_uint32),
        ("num_layers", ctypes.c_uint32),
        ("num_measurements", ctypes.c_uint32),
        ("num_detectors", ctypes.c_uint32),
        ("num_observables", ctypes.c_uint32),
        ("adjacency", ctypes.c_void_p),
        ("probs", ctypes.c_void_p),
        ("keys", ctypes.c_void_p),
        ("sensitivity", ctypes.c_void_p),
    ]


def load() -> ctypes.CDLL:
    """Loads the native library (once) and declares its signatures."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"greenpeas_amd native library not built ({LIB_PATH}); run "
            "`python -m greenpeas_amd.build` (hipcc, gfx950)"
        )
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    PP = ctypes.POINTER(ctypes.c_void_p)
    u32, u64, i32, f64 = ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int, ctypes.c_double
    pu32, pu64, pf64 = ctypes.POINTER(u32), ctypes.POINTER(u64), ctypes.POINTER(f64)
    cs = ctypes.c_char_p

    def sig(name, args, res=i32):
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res

    sig("gp_last_error", [], cs)
    sig("gp_version", [], cs)
    sig("gp_get_driver", [cs, i32, i32, PP])
    sig("gp_get_driver_on_stream", [cs, i32, i32, P, PP])
    sig("gp_driver_set_stream", [P, P])
    sig("gp_driver_load_stepg", [P, ctypes.POINTER(Stepg)])
    sig("gp_driver_destroy", [P])
    sig("gp_compile_detector_error_model", [P, cs])
    sig("gp_driver_parse", [P, cs])
    sig("gp_dem_shape", [P, pu32, pu64, pu32, pu32])
    sig("gp_dem_copy", [P, P, P, P, P])
    sig("gp_dem_acquire", [P, ctypes.POINTER(DemView)])
    sig("gp_dem_release", [P])
    sig("gp_dem_text", [P, P, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)])
    sig("gp_driver_params", [P, P])
    sig("gp_driver_timings", [P, P])
    sig("gp_driver_bench_device", [P, u32, pf64])
    sig("gp_debug_pnodes", [P, pu32, pu32])
    sig("gp_debug_sensitivities", [P, P, P])
    sig("gp_debug_stepg", [P, P, P])
    sig("gp_debug_sweep_plan", [P, P])
    sig("gp_build_stepg_host", [cs, i32, P, P, P])
    sig("gp_debug_image_digest", [cs, i32, i32, P])
    sig("gp_batch_create", [i32, i32, PP])
    sig("gp_batch_destroy", [P])
    sig("gp_batch_set_circuits", [P, ctypes.POINTER(cs), u32, u32])
    sig("gp_batch_compile", [P, i32])
    sig("gp_batch_stage_us", [P, P, pu64])
    sig("gp_batch_device_us", [P, pf64])
    sig("gp_batch_dem_shape", [P, u32, pu32, pu64, pu32, pu32])
    sig("gp_batch_dem_copy", [P, u32, P, P, P, P])
    sig("gp_batch_device_bytes", [P, pu64])
    sig("gp_batch_layout_stats", [P, P])
    sig("gp_estimate_device_bytes", [cs, i32, pu64])
    sig("gp_prepare_circuit", [cs, i32, PP])
    sig("gp_prepared_destroy", [P])
    sig("gp_prepared_device_bytes", [P, pu64])
    sig("gp_batch_set_prepared", [P, P, u32])
    sig("gp_batch_bytes_after", [P, P, u32, pu64])
    sig("gp_device_memory", [i32, pu64, pu64])
    sig("gp_circuit_summary", [cs, P])
    sig("gp_set_tuning", [cs, u32])
    sig("gp_get_tuning", [cs, pu32])
    sig("gp_set_tuning_class", [i32, u32, cs, u32])
    sig("gp_get_tuning_class", [i32, u32, cs, pu32])
    sig("gp_clear_tuning_classes", [])
    sig("gp_grad_plan_create", [cs, i32, i32, PP])
    sig("gp_grad_plan_destroy", [P])
    sig("gp_grad_plan_sizes", [P, P])
    sig("gp_grad_plan_dem_structure", [P, P, P])
    sig("gp_grad_plan_device_arrays", [P, PP, PP, PP, PP, PP, PP])
    sig("gp_grad_plan_copy_arrays", [P, P, P, P, P, P, P])
    sig("gp_xor_merge_forward", [P, P, P, u32, P, P])
    sig("gp_xor_merge_backward", [P, P, P, u32, P, P, P])
    sig("gp_node_rates_forward", [P, P, P, P, P, u32, P, P])
    sig("gp_node_rates_backward", [P, P, P, P, P, u32, P, P, P, P, u32, P, P])
    sig("gp_node_rates_backward_noalloc", [P, P, P, P, P, u32, u32, P, P, P, u32, P, P])
    _lib = L
    return L


def check(rc: int) -> None:
    """Raises the Python counterpart of a C-ABI status code."""
    if rc == GP_OK:
        return
    msg = load().gp_last_error().decode(errors="replace")
    if rc == GP_INVALID_ARGUMENT:
        raise ValueError(msg)
    raise RuntimeError(msg)
