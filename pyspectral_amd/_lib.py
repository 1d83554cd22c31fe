"""ctypes binding of libpsx.so (the C ABI declared in include/psx.h).

The library is loaded from this package directory; there is no fallback: if it
is missing or cannot be loaded every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("PSX_LIBRARY") or os.path.join(HERE, "libpsx.so")

PSX_OK = 0
PSX_E_NONFINITE = -5
MODEL_IDS = {
    0: 0, "dns": 0,
    1: 1, "smagorinsky": 1,
    2: 2, "tmodel": 2,
    "Static tau-model": 3, "static_tau": 3,
    "Dynamic tau-model": 4, "dynamic_tau": 4,
    7: 5, "dsmag": 5,
    "learned": 6,
}

# every symbol include/psx.h declares
EXPORTS = (
    "psx_create", "psx_destroy", "psx_set_stream", "psx_set_params", "psx_last_error",
    "psx_scratch_bytes", "psx_rhs", "psx_rk_stage", "psx_compute_dt", "psx_step", "psx_step_graph", "psx_stats",
    "psx_spectrum_bins", "psx_spectrum", "psx_tau", "psx_timing_begin", "psx_timing_end",
    "psx_filter", "psx_sgs_dns", "psx_create_batch",
    "psx_ensemble_create", "psx_ensemble_destroy", "psx_ensemble_last_error", "psx_ensemble_size",
    "psx_ensemble_offset", "psx_ensemble_slot", "psx_ensemble_set_stream", "psx_ensemble_step",
    "psx_ensemble_step_graph", "psx_ensemble_stats", "psx_set_tuning", "psx_get_tuning", "psx_autotune",
    "psx_mlp_create", "psx_mlp_destroy", "psx_mlp_last_error", "psx_mlp_set_stream", "psx_mlp_get_params",
    "psx_mlp_set_params", "psx_mlp_set_norm", "psx_mlp_forward", "psx_mlp_grad", "psx_mlp_gradient",
    "psx_mlp_train_step", "psx_set_closure", "psx_physical", "psx_gradients", "psx_restrict",
    "psx_ensemble_set_closure", "psx_mlp_closure", "psx_status", "psx_clear_status", "psx_ensemble_status",
    "psx_stats_noalloc", "psx_spectral", "psx_stream_probe", "psx_ensemble_classes", "psx_ensemble_autotune",
)

MLP_NPARAM = 9350

PSX_FILTER_SHARP, PSX_FILTER_BOX = 0, 1
PSX_CLOSURE_CLIP_BACKSCATTER = 1

KERNEL_NAMES = ("x_inverse", "y_inverse", "z_pass", "y_forward", "x_forward_epilogue",
                "cfl_x", "cfl_y", "cfl_z", "other", "closure_network")


class psx_params(C.Structure):
    _fields_ = [
        ("nu", C.c_double),
        ("L1", C.c_double), ("L2", C.c_double), ("L3", C.c_double),
        ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
        ("cfl", C.c_double),
        ("Cs", C.c_double),
        ("Ct", C.c_double),
        ("tau_a", C.c_double), ("tau_b", C.c_double),
        ("kc", C.c_int),
        ("rotate", C.c_int),
        ("Om1", C.c_double), ("Om2", C.c_double), ("Om3", C.c_double),
        ("force", C.c_double),
    ]


class PsxError(RuntimeError):
    pass


_lib = None


def load() -> C.CDLL:
    """Load libpsx.so and declare its signatures (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise PsxError(f"{LIB_PATH} is not built; run pyspectral_amd/build_lib.py (or __graft_entry__.build())")
    lib = C.CDLL(LIB_PATH)
    vp, dp, i, d = C.c_void_p, C.c_void_p, C.c_int, C.c_double
    sig = {
        "psx_create": (i, [C.POINTER(vp), i, i, i, i, C.POINTER(psx_params), i, vp]),
        "psx_destroy": (i, [vp]),
        "psx_set_stream": (i, [vp, vp]),
        "psx_set_params": (i, [vp, C.POINTER(psx_params)]),
        "psx_last_error": (C.c_char_p, [vp]),
        "psx_scratch_bytes": (C.c_longlong, [vp]),
        "psx_rhs": (i, [vp, dp, dp, dp, dp, d]),
        "psx_rk_stage": (i, [vp, dp, dp, dp, dp, d, d, i]),
        "psx_compute_dt": (i, [vp, dp, dp]),
        "psx_step": (i, [vp, dp, dp, dp, d]),
        "psx_step_graph": (i, [vp, dp, dp, dp, dp]),
        "psx_stats": (i, [vp, dp, dp]),
        "psx_spectrum_bins": (i, [vp]),
        "psx_spectrum": (i, [vp, dp, dp, i]),
        "psx_tau": (i, [vp, C.POINTER(C.c_double)]),
        "psx_timing_begin": (i, [vp]),
        "psx_timing_end": (i, [vp, C.POINTER(C.c_double), C.POINTER(C.c_int)]),
        "psx_filter": (i, [vp, dp, dp, i, d]),
        "psx_sgs_dns": (i, [vp, dp, dp, dp, i, d]),
        "psx_create_batch": (i, [C.POINTER(vp), i, i, i, i, C.POINTER(psx_params), i, C.POINTER(C.c_double), i,
                                 vp]),
        "psx_ensemble_create": (i, [C.POINTER(vp), i, C.POINTER(C.c_int), C.POINTER(C.c_double), i,
                                    C.POINTER(psx_params), i, vp]),
        "psx_ensemble_destroy": (i, [vp]),
        "psx_ensemble_last_error": (C.c_char_p, [vp]),
        "psx_ensemble_size": (C.c_longlong, [vp]),
        "psx_ensemble_offset": (C.c_longlong, [vp, i]),
        "psx_ensemble_slot": (i, [vp, i]),
        "psx_ensemble_set_stream": (i, [vp, vp]),
        "psx_ensemble_step": (i, [vp, dp, dp, dp, d]),
        "psx_ensemble_step_graph": (i, [vp, dp, dp, dp]),
        "psx_ensemble_stats": (i, [vp, dp, dp]),
        "psx_set_tuning": (i, [vp, C.POINTER(C.c_int)]),
        "psx_get_tuning": (i, [vp, C.POINTER(C.c_int)]),
        "psx_autotune": (i, [vp, dp, i, C.POINTER(C.c_int)]),
        "psx_mlp_create": (i, [C.POINTER(vp), i, vp, C.c_ulonglong]),
        "psx_mlp_destroy": (i, [vp]),
        "psx_mlp_last_error": (C.c_char_p, [vp]),
        "psx_mlp_set_stream": (i, [vp, vp]),
        "psx_mlp_get_params": (i, [vp, C.POINTER(C.c_float)]),
        "psx_mlp_set_params": (i, [vp, C.POINTER(C.c_float)]),
        "psx_mlp_set_norm": (i, [vp] + [C.POINTER(C.c_float)] * 4),
        "psx_mlp_forward": (i, [vp, dp, dp, C.c_longlong]),
        "psx_mlp_grad": (i, [vp, dp, dp, C.c_longlong, dp]),
        "psx_mlp_gradient": (i, [vp, dp]),
        "psx_mlp_train_step": (i, [vp, dp, dp, C.c_longlong, C.c_float, dp]),
        "psx_set_closure": (i, [vp, vp, d, i]),
        "psx_physical": (i, [vp, dp, i, dp]),
        "psx_gradients": (i, [vp, dp, dp]),
        "psx_restrict": (i, [vp, dp, i, i, i, i, dp]),
        "psx_ensemble_set_closure": (i, [vp, vp, d, i]),
        "psx_mlp_closure": (i, [vp, dp, dp, C.c_longlong, d, d, i]),
        "psx_status": (i, [vp]),
        "psx_clear_status": (i, [vp]),
        "psx_ensemble_status": (i, [vp]),
        "psx_stats_noalloc": (i, [dp, i, i, i, i, d, d, d, i, dp, vp]),
        "psx_spectral": (i, [vp, dp, i, dp]),
        "psx_stream_probe": (i, [dp, dp, C.c_longlong, i, i, vp]),
        "psx_ensemble_classes": (i, [vp]),
        "psx_ensemble_autotune": (i, [vp, dp, i, C.POINTER(C.c_int)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(ctx, rc: int, what: str) -> int:
    """Raise on a failure code.  PSX_E_NONFINITE is a status (the work was enqueued, a
    reduction met a NaN or an infinite value) and is returned, not raised: the
    reference keeps stepping a NaN state."""
    if rc != PSX_OK and rc != PSX_E_NONFINITE:
        msg = load().psx_last_error(ctx).decode() if ctx else ""
        raise PsxError(f"{what} failed ({rc}): {msg}")
    return rc
