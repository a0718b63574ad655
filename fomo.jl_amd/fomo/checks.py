"""Input validation mirroring src/utils/checks.jl (same exception type and conditions)."""
from __future__ import annotations

import warnings

import numpy as np

from ._lib import ArgumentError
from .setup import get_centered_d2, get_fd_coefficients


def check_geometry(nx: int, nz: int, sx, sz, rx, rz) -> None:
    """_check_geometry, src/utils/checks.jl:8-14."""
    sx, sz, rx, rz = (np.asarray(a) for a in (sx, sz, rx, rz))
    if not (np.all((1 <= sx) & (sx <= nx)) and np.all((1 <= sz) & (sz <= nz))):
        raise ArgumentError(f"source index out of range: need 1 <= sx <= {nx}, 1 <= sz <= {nz}")
    if not (np.all((1 <= rx) & (rx <= nx)) and np.all((1 <= rz) & (rz <= nz))):
        raise ArgumentError(f"receiver index out of range: need 1 <= rx <= {nx}, 1 <= rz <= {nz}")


def check_geometry_3d(nx, ny, nz, sx, sy, sz, rx, ry, rz) -> None:
    for name, pts in (("source", (sx, sy, sz)), ("receiver", (rx, ry, rz))):
        for a, n, ax in zip(pts, (nx, ny, nz), "xyz"):
            a = np.asarray(a)
            if not np.all((1 <= a) & (a <= n)):
                raise ArgumentError(f"{name} index out of range: need 1 <= {ax} <= {n}")


def warn_dispersion(vmin_disp: float, f0: float, dh: float, ppw_min: float = 4.0) -> None:
    """_warn_dispersion, src/utils/checks.jl:22-28 (fmax = 2.5 f0)."""
    fmax = np.float32(2.5) * np.float32(f0)
    ppw = vmin_disp / fmax / dh
    if ppw < ppw_min:
        warnings.warn(f"dispersion risk: shortest wavelength spans only {ppw:.1f} grid points "
                      f"(vmin={vmin_disp}, fmax~{fmax:.1f} Hz); reduce dh or f0", stacklevel=3)


def check_numerics(vmax: float, vmin_disp: float, dh: float, dt: float, f0: float, fd_order: int,
                   ppw_min: float = 4.0, ndim: int = 2) -> None:
    """_check_numerics, src/utils/checks.jl:39-48: dt <= dh / (vmax * sqrt(2) * sum|a_l|), in Float32.

    ndim=3 uses sqrt(3) (the 3D staggered-grid bound; the reference does not check 3D)."""
    a = get_fd_coefficients(fd_order)
    s = np.float32(0.0)
    for c in a:
        s = np.float32(s + np.float32(abs(c)))
    root = np.float32(np.sqrt(np.float32(ndim)))
    dt_max = np.float32(np.float32(dh) / np.float32(np.float32(np.float32(vmax) * root) * s))
    if not np.float32(dt) <= dt_max:
        raise ArgumentError(f"CFL unstable: dt={dt} > dt_max={float(dt_max):.4g} "
                            f"(vmax={vmax}, dh={dh}, fd_order={fd_order})")
    warn_dispersion(vmin_disp, f0, dh, ppw_min)


def check_numerics_centered(vmax: float, vmin_disp: float, dh: float, dt: float, f0: float, fd_order: int,
                            ppw_min: float = 4.0) -> None:
    """_check_numerics_centered, src/utils/checks.jl:61-72, in Float32: lam_max = |c0| + 2 sum|d2_l|,
    dt <= 2 dh / (vmax sqrt(2 lam_max))."""
    c0, d2 = get_centered_d2(fd_order)
    s = np.float32(0.0)
    for c in d2:
        s = np.float32(s + np.float32(abs(c)))
    lam_max = np.float32(np.float32(abs(c0)) + np.float32(np.float32(2.0) * s))
    root = np.float32(np.sqrt(np.float32(np.float32(2.0) * lam_max)))
    dt_max = np.float32(np.float32(np.float32(2.0) * np.float32(dh)) / np.float32(np.float32(vmax) * root))
    if not np.float32(dt) <= dt_max:
        raise ArgumentError(f"CFL unstable (centered): dt={dt} > dt_max={float(dt_max):.4g} "
                            f"(vmax={vmax}, dh={dh}, fd_order={fd_order})")
    warn_dispersion(vmin_disp, f0, dh, ppw_min)
