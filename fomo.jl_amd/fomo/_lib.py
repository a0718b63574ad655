"""ctypes binding of libfomo_hip.so (include/fomo_hip.h).

The library is built in-tree (fomo.jl_amd/libfomo_hip.so).  Loading fails loudly when it is
missing: there is no CPU fallback anywhere in this package.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.normpath(os.path.join(_HERE, "..", "libfomo_hip.so"))

FOMO_OK, FOMO_EINVAL, FOMO_EHIP, FOMO_ENOMEM, FOMO_EUNSUP, FOMO_EOVERFLOW = 0, 1, 2, 3, 4, 5
ACOUSTIC2D, ELASTIC2D, ACOUSTIC3D, ELASTIC3D, SCALAR2D, COUPLED2D = 1, 2, 3, 4, 5, 6
SRC_EXPLOSIVE, SRC_FZ = 0, 1
STORE_F32, STORE_F16, STORE_BF16, STORE_F32_MAT16 = 0, 1, 2, 3
BOUNDARY_HABC, BOUNDARY_SPONGE = 0, 1


class ArgumentError(ValueError):
    """Raised where the reference raises Julia's ArgumentError."""


class FomoError(RuntimeError):
    """HIP runtime / allocation failure inside the engine."""


class StorageOverflowError(FomoError):
    """16-bit storage: a wavefield value overflowed the storage format (FOMO_EOVERFLOW)."""


class Desc(C.Structure):
    """struct fomo_desc."""
    _fields_ = [
        ("eq", C.c_int32), ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("nbc", C.c_int32), ("fd_order", C.c_int32), ("nt", C.c_int32),
        ("dh", C.c_float), ("dt", C.c_float), ("v_ref", C.c_float),
        ("n_src", C.c_int32), ("n_rec", C.c_int32), ("source_type", C.c_int32), ("storage", C.c_int32),
        ("snap_interval", C.c_int32), ("snap_plane", C.c_int32), ("snap_index", C.c_int32),
        ("use_graph", C.c_int32), ("record_p", C.c_int32), ("tile_cfg", C.c_int32),
        ("boundary", C.c_int32), ("sponge_factor", C.c_float), ("smooth_sigma", C.c_double),
        ("v_ref_s", C.c_float), ("reserved", C.c_int32 * 1),
    ]


class Shot(C.Structure):
    """struct fomo_shot."""
    _fields_ = [
        ("x0", C.c_int32), ("nx", C.c_int32), ("n_src", C.c_int32),
        ("sx", C.c_void_p), ("sz", C.c_void_p), ("wavelet", C.c_void_p),
        ("n_rec", C.c_int32), ("rx", C.c_void_p), ("rz", C.c_void_p), ("v_ref", C.c_float),
    ]


# Every symbol declared in include/fomo_hip.h, with (restype, argtypes).
_p = C.c_void_p
_i = C.c_int32
_f = C.c_float
SIGNATURES = {
    "fomo_last_error": (C.c_char_p, []),
    "fomo_version": (C.c_char_p, []),
    "fomo_device_count": (C.c_int, []),
    "fomo_stream_copy": (C.c_int, [C.c_int64, _i, C.POINTER(C.c_double)]),
    "fomo_stream_probe": (C.c_int, [_i, _i, C.c_int64, _i, C.POINTER(C.c_double)]),
    "fomo_fd_coefficients": (C.c_int, [_i, _p]),
    "fomo_centered_d2": (C.c_int, [_i, _p, _p]),
    "fomo_centered_d1": (C.c_int, [_i, _p]),
    "fomo_coupled2d": (C.c_int, [C.POINTER(Desc), _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p,
                                 C.POINTER(C.c_double)]),
    "fomo_scalar2d": (C.c_int, [C.POINTER(Desc), _p, _p, _p, _p, _p, _p, _p, C.POINTER(C.c_double)]),
    "fomo_ricker_wavelet": (C.c_int, [_f, _f, _i, _p]),
    "fomo_habc_coefficients": (C.c_int, [_f, _f, _f, _p]),
    "fomo_acoustic2d": (C.c_int, [C.POINTER(Desc), _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p,
                                  C.POINTER(C.c_double)]),
    "fomo_elastic2d": (C.c_int, [C.POINTER(Desc), _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p,
                                 C.POINTER(C.c_double)]),
    "fomo_acoustic3d": (C.c_int, [C.POINTER(Desc), _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p,
                                  C.POINTER(C.c_double)]),
    "fomo_elastic3d": (C.c_int, [C.POINTER(Desc), _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p,
                                 C.POINTER(C.c_double)]),
    "fomo_acoustic2d_batch": (C.c_int, [C.POINTER(Desc), _p, _p, _i, _p, _p, _p, _p, _p, _p, _p, _p,
                                        C.POINTER(C.c_double)]),
    "fomo_elastic2d_batch": (C.c_int, [C.POINTER(Desc), _p, _p, _p, _i, _p, _p, _p, _p, _p, _p, _p,
                                       C.POINTER(C.c_double)]),
    "fomo_plan_create": (C.c_int, [C.POINTER(Desc), _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, C.POINTER(_p)]),
    "fomo_batch_create": (C.c_int, [C.POINTER(Desc), _p, _p, _p, _i, C.POINTER(Shot), C.POINTER(_p)]),
    "fomo_plan_n_shots": (C.c_int, [_p, C.POINTER(_i)]),
    "fomo_plan_read_shot_seis": (C.c_int, [_p, _i, _p, _p, _p, _p]),
    "fomo_trace_misfit": (C.c_int, [_p, _i, _p, _i, _i, _i, _i, _p, _p]),
    "fomo_trace_misfit_noalloc": (C.c_int, [_p, _i, _p, _i, _i, _i, _i, _p, _p]),
    "fomo_plan_read_shot_field": (C.c_int, [_p, _i, _i, _p]),
    "fomo_plan_shot_dims": (C.c_int, [_p, _i, C.POINTER(_i), C.POINTER(_i), C.POINTER(_i), C.POINTER(_i)]),
    "fomo_plan_reset": (C.c_int, [_p]),
    "fomo_plan_advance": (C.c_int, [_p, _i]),
    "fomo_plan_timed_advance": (C.c_int, [_p, _i, C.POINTER(C.c_double)]),
    "fomo_plan_sync": (C.c_int, [_p]),
    "fomo_plan_kernel_times": (C.c_int, [_p, _i, C.POINTER(C.c_double)]),
    "fomo_plan_steps_done": (C.c_int, [_p, C.POINTER(_i)]),
    "fomo_plan_tuning": (C.c_int, [_p, _p]),
    "fomo_plan_read_snaps": (C.c_int, [_p, _i, C.POINTER(_i), _p]),
    "fomo_plan_launch_count": (C.c_int, [_p, C.POINTER(_i)]),
    "fomo_plan_launch_info": (C.c_int, [_p, _i, C.POINTER(C.c_char_p), C.POINTER(C.c_int64), C.POINTER(_i)]),
    "fomo_plan_launch_times": (C.c_int, [_p, _i, _p]),
    "fomo_plan_read_seis": (C.c_int, [_p, _p, _p, _p, _p]),
    "fomo_plan_read_field": (C.c_int, [_p, _i, _p]),
    "fomo_plan_dims": (C.c_int, [_p, C.POINTER(_i), C.POINTER(_i), C.POINTER(_i), C.POINTER(_i)]),
    "fomo_plan_destroy": (C.c_int, [_p]),
    "fomo_learned_n_params": (C.c_int, []),
    "fomo_learned_enable": (C.c_int, [_p, _p, _i]),
    "fomo_learned_set_weights": (C.c_int, [_p, _p]),
    "fomo_learned_train_step": (C.c_int, [_p, _p, _i, _i, _f, _f]),
    "fomo_learned_grad": (C.c_int, [_p, _p, C.POINTER(C.c_double)]),
    "fomo_learned_alpha": (C.c_int, [_p, C.POINTER(_i), _p, _p, _p]),
    "fomo_mlp_eval": (C.c_int, [_p, _i, _p, _p, _p, _p]),
}

_lib = None
_lock = threading.Lock()


def lib() -> C.CDLL:
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise ImportError(
                    f"{LIB_PATH} is missing: build it with `make -C fomo.jl_amd` (or __graft_entry__.build()). "
                    "The engine has no CPU fallback.")
            L = C.CDLL(LIB_PATH)
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(L, name)
                fn.restype = res
                fn.argtypes = args
            _lib = L
    return _lib


def check(rc: int) -> None:
    if rc == FOMO_OK:
        return
    msg = lib().fomo_last_error().decode(errors="replace")
    if rc == FOMO_EINVAL:
        raise ArgumentError(msg)
    if rc == FOMO_EOVERFLOW:
        raise StorageOverflowError(msg)
    raise FomoError(f"fomo error {rc}: {msg}")
