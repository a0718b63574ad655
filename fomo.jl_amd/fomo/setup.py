"""Host-side helpers exported by Fomo.jl, computed by the native library."""
from __future__ import annotations

import numpy as np

from . import _lib


def get_fd_coefficients(order: int) -> np.ndarray:
    """get_fd_coefficients, src/utils/FD_utils.jl:16-19 (raises for unsupported orders)."""
    out = np.zeros(5, np.float32)
    rc = _lib.lib().fomo_fd_coefficients(int(order), out.ctypes.data)
    if rc != 0:
        raise ValueError(f"Unsupported FD order: {order}")
    return out[: order // 2].copy()


def get_centered_d2(order: int):
    """get_centered_d2, src/utils/FD_centered.jl:59-62: (c0, d2) as Float32 (raises for unsupported
    orders)."""
    c0 = np.zeros(1, np.float32)
    d2 = np.zeros(5, np.float32)
    rc = _lib.lib().fomo_centered_d2(int(order), c0.ctypes.data, d2.ctypes.data)
    if rc != 0:
        raise ValueError(f"Unsupported FD order: {order} (centered d2)")
    return np.float32(c0[0]), d2[: order // 2].copy()


def get_centered_d1(order: int) -> np.ndarray:
    """get_centered_d1, src/utils/FD_centered.jl:48-51 (Float32; raises for unsupported orders)."""
    d1 = np.zeros(5, np.float32)
    rc = _lib.lib().fomo_centered_d1(int(order), d1.ctypes.data)
    if rc != 0:
        raise ValueError(f"Unsupported FD order: {order} (centered d1)")
    return d1[: order // 2].copy()


def ricker_wavelet(f0: float, dt: float, nt: int) -> np.ndarray:
    """ricker_wavelet, src/utils/ricker_wavelet.jl:6-12 (Float32, length nt)."""
    out = np.zeros(max(int(nt), 1), np.float32)
    _lib.check(_lib.lib().fomo_ricker_wavelet(float(f0), float(dt), int(nt), out.ctypes.data))
    return out[: int(nt)].copy()


def habc_coefficients(v_ref: float, dt: float, dh: float) -> tuple[float, float, float]:
    """(q, qt, qxt) of init_habc, src/boundary/habc/habc.jl:21-28."""
    out = np.zeros(3, np.float32)
    _lib.check(_lib.lib().fomo_habc_coefficients(float(v_ref), float(dt), float(dh), out.ctypes.data))
    return float(out[0]), float(out[1]), float(out[2])
