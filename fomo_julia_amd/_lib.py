"""ctypes binding of libfomo_mi355x.so (include/fomo_mi355x.h).

The library is the only compute path: if it is missing or cannot load, every
entry point raises; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# FOMO_MI355X_LIB selects another build of the library (the Julia binding
# reads the same variable)
LIB_PATH = os.environ.get("FOMO_MI355X_LIB") or os.path.join(HERE, "libfomo_mi355x.so")

c_int32 = ctypes.c_int32
c_int64 = ctypes.c_int64
c_float = ctypes.c_float
c_void_p = ctypes.c_void_p
P_f32 = ctypes.POINTER(ctypes.c_float)
P_i32 = ctypes.POINTER(ctypes.c_int32)

FOMO_OK, FOMO_ERR_ARG, FOMO_ERR_HIP, FOMO_ERR_UNSUPPORTED, FOMO_ERR_STATE = 0, -1, -2, -3, -4
SRC_TYPES = {"pressure": 0, "force_z": 1, "force_x": 2}
REC_TYPES = {"vz": 0, "vx": 1, "p": 2}
STORAGE = {"f32": 0, "bf16": 1, "f16": 2, "f16c": 3, "f24": 4}
FIELDS = {"vx": 0, "vz": 1, "txx": 2, "tzz": 3, "txz": 4,
          "vx_old": 5, "vz_old": 6, "txx_old": 7, "tzz_old": 8, "txz_old": 9}


class ModelDesc(ctypes.Structure):
    _fields_ = [("nx", c_int32), ("nz", c_int32), ("pad", c_int32), ("nbc", c_int32),
                ("fd_order", c_int32), ("free_surface", c_int32),
                ("dx", c_float), ("dz", c_float), ("dt", c_float), ("dtx", c_float),
                ("dtz", c_float), ("qx", c_float), ("qz", c_float), ("qt_x", c_float),
                ("qt_z", c_float), ("qxt", c_float), ("storage", c_int32)]


# name -> (restype, argtypes); every symbol include/fomo_mi355x.h declares.
SIGNATURES = {
    "fomo_ctx_create": (c_int32, [c_int32, ctypes.POINTER(c_void_p)]),
    "fomo_ctx_destroy": (c_int32, [c_void_p]),
    "fomo_ctx_last_error": (ctypes.c_char_p, [c_void_p]),
    "fomo_ctx_synchronize": (c_int32, [c_void_p]),
    "fomo_ctx_stream": (c_void_p, [c_void_p]),
    "fomo_ctx_set_tiling": (c_int32, [c_void_p, c_int32, c_int32, c_int32]),
    "fomo_ctx_set_step_mode": (c_int32, [c_void_p, c_int32]),
    "fomo_ctx_set_split": (c_int32, [c_void_p, c_int32, c_int32]),
    "fomo_model_create": (c_int32, [c_void_p, ctypes.POINTER(ModelDesc), P_f32, P_f32, P_f32,
                                    P_f32, P_f32, P_f32, P_f32, P_f32, P_f32,
                                    ctypes.POINTER(c_void_p)]),
    "fomo_model_create_from_velocity": (c_int32, [c_void_p, ctypes.POINTER(ModelDesc), P_f32, P_f32,
                                                   P_f32, P_f32, c_int32, c_int32,
                                                   ctypes.POINTER(c_void_p)]),
    "fomo_model_destroy": (c_int32, [c_void_p]),
    "fomo_model_device_bytes": (c_int64, [c_void_p]),
    "fomo_wavefield_create": (c_int32, [c_void_p, ctypes.POINTER(c_void_p)]),
    "fomo_wavefield_destroy": (c_int32, [c_void_p]),
    "fomo_wavefield_reset": (c_int32, [c_void_p]),
    "fomo_wavefield_download": (c_int32, [c_void_p, c_int32, P_f32]),
    "fomo_wavefield_upload": (c_int32, [c_void_p, c_int32, P_f32]),
    "fomo_wavefield_device_bytes": (c_int64, [c_void_p]),
    "fomo_shot_create": (c_int32, [c_void_p, c_int32, c_int32, c_int32, P_f32, c_int32, c_int32,
                                   P_i32, P_i32, P_i32, c_int32, ctypes.POINTER(c_void_p)]),
    "fomo_shot_destroy": (c_int32, [c_void_p]),
    "fomo_shot_field_scales": (c_int32, [c_void_p, P_f32]),
    "fomo_shot_reset": (c_int32, [c_void_p]),
    "fomo_shot_gather": (c_int32, [c_void_p, P_f32]),
    "fomo_shot_gather_noalloc": (c_int32, [c_void_p, P_f32]),
    "fomo_time_step": (c_int32, [c_void_p, c_void_p, c_void_p, c_int32]),
    "fomo_run_time_loop": (c_int32, [c_void_p, c_void_p, c_void_p, c_int32, c_int32]),
    "fomo_run_time_loop_prepare": (c_int32, [c_void_p, c_void_p, c_void_p]),
    "fomo_frames_create": (c_int32, [c_void_p, c_int32, c_int32, ctypes.POINTER(c_int32), c_int32,
                                     c_int32, c_int32, c_int32, ctypes.POINTER(c_void_p)]),
    "fomo_frames_destroy": (c_int32, [c_void_p]),
    "fomo_frames_shape": (c_int32, [c_void_p, ctypes.POINTER(c_int32), ctypes.POINTER(c_int32)]),
    "fomo_frames_device_bytes": (ctypes.c_int64, [c_void_p]),
    "fomo_frames_download": (c_int32, [c_void_p, c_int32, c_int32, P_f32]),
    "fomo_run_time_loop_frames": (c_int32, [c_void_p, c_void_p, c_void_p, c_int32, c_int32, c_void_p]),
    "fomo_wavefield_frame": (c_int32, [c_void_p, c_void_p, c_int32]),
    "fomo_backup_boundary": (c_int32, [c_void_p, c_void_p]),
    "fomo_inject_source": (c_int32, [c_void_p, c_void_p, c_void_p, c_int32]),
    "fomo_update_velocity": (c_int32, [c_void_p, c_void_p]),
    "fomo_apply_habc_velocity": (c_int32, [c_void_p, c_void_p]),
    "fomo_update_stress": (c_int32, [c_void_p, c_void_p]),
    "fomo_apply_habc_stress": (c_int32, [c_void_p, c_void_p]),
    "fomo_apply_free_surface": (c_int32, [c_void_p, c_void_p]),
    "fomo_record_receivers": (c_int32, [c_void_p, c_void_p, c_void_p, c_int32]),
    "fomo_batch_create": (c_int32, [c_void_p, c_int32, ctypes.POINTER(c_void_p),
                                    ctypes.POINTER(c_void_p), ctypes.POINTER(c_void_p),
                                    ctypes.POINTER(c_void_p)]),
    "fomo_batch_destroy": (c_int32, [c_void_p]),
    "fomo_batch_run": (c_int32, [c_void_p, c_int32, c_int32]),
    "fomo_launch_stage": (c_int32, [c_void_p, c_void_p, c_void_p, c_int32, c_int32, c_int32]),
    "fomo_pool_create": (c_int32, [c_void_p, c_int32, c_int32, c_int32, ctypes.POINTER(c_void_p)]),
    "fomo_pool_destroy": (c_int32, [c_void_p]),
    "fomo_pool_add": (c_int32, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "fomo_pool_remove": (c_int32, [c_void_p, c_void_p]),
    "fomo_pool_run": (c_int32, [c_void_p, c_int32]),
    "fomo_pool_size": (c_int32, [c_void_p]),
    "fomo_pool_synchronize": (c_int32, [c_void_p]),
    "fomo_pool_next_step": (c_int32, [c_void_p, c_void_p, P_i32]),
}

_lib = None


class FomoError(RuntimeError):
    pass


def load(path: str | None = None):
    """Loads the HIP library (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or LIB_PATH
    if not os.path.exists(path):
        raise FomoError(f"{path} not found: build it with `python -m fomo_julia_amd.build` "
                        "(the MI355X path has no CPU fallback)")
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc: int, ctx=None, what: str = "") -> None:
    if rc != 0:
        msg = load().fomo_ctx_last_error(ctx)
        msg = msg.decode() if msg else ""
        raise FomoError(f"{what} failed ({rc}): {msg}")
