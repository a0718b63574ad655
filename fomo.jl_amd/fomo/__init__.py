"""fomo -- MI355X (gfx950) forward-modeling engine with Fomo.jl's API.

Python mirror of the reference's public API over the native library libfomo_hip.so.  There is
no CPU fallback: every solver call runs on the GPU through the C-ABI in include/fomo_hip.h.
"""
from ._lib import ArgumentError, FomoError  # noqa: F401
from .api import (Plan, acoustic2d, acoustic2d_batch, acoustic3d, coupled2d, elastic2d,  # noqa: F401
                  elastic2d_batch, elastic3d, scalar2d)
from .setup import (get_centered_d1, get_centered_d2, get_fd_coefficients, habc_coefficients,  # noqa: F401
                    ricker_wavelet)
from .plots import plot_shot, plot_wavefield_video  # noqa: F401
from .traces import trace_norm  # noqa: F401

__all__ = ["acoustic2d", "elastic2d", "acoustic2d_batch", "elastic2d_batch", "acoustic3d", "elastic3d",
           "scalar2d", "coupled2d", "ricker_wavelet", "trace_norm", "plot_shot", "plot_wavefield_video",
           "get_fd_coefficients", "get_centered_d1", "get_centered_d2", "habc_coefficients", "Plan", "ArgumentError", "FomoError"]
