"""ctypes front end of the CPU oracle (oracle/fomo_oracle*.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by bench.py's
cpu_baseline leg, as the checker.  The product package never imports this module.

Array conventions follow the reference (Fomo.jl): models are (nx, nz[, ny...]) arrays indexed
[i, j] with i along x, passed to C in column-major (x fastest) order; source/receiver indices are
1-based interior indices; wavelets are (n_src, nt); seismograms are (n_rec, nt).

It also restates `ricker_wavelet` (src/utils/ricker_wavelet.jl:6-12) in numpy, including Julia's
lazy range arithmetic for `(0:nt-1) .* dt` (see `ricker`).
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBS: dict[str, C.CDLL] = {}

_f = C.c_float
_i = C.c_int
_p = C.c_void_p


def build() -> None:
    """Compile the oracle libraries (gcc, -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib(precision: str = "f32") -> C.CDLL:
    if precision not in _LIBS:
        path = os.path.join(_HERE, f"liboracle_{precision}.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        real = _f if precision == "f32" else C.c_double
        L.orc_acoustic2d.restype = _i
        L.orc_acoustic2d.argtypes = [_i, _i, _p, _p, _f, _f, _i, _i, _i, _f, _i, _p, _p, _p, _i, _p, _p,
                                     _p, _p, _p, _i, _p, _p, _i, _f, _i]
        L.orc_elastic2d.restype = _i
        L.orc_elastic2d.argtypes = [_i, _i, _p, _p, _p, _f, _f, _i, _i, _i, _f, _i, _i, _p, _p, _p, _i, _p, _p,
                                    _p, _p, _i, _p, _p, _i, _f, _i]
        L.orc_acoustic3d.restype = _i
        L.orc_acoustic3d.argtypes = [_i, _i, _i, _p, _p, _f, _f, _i, _i, _i, _f, _i, _p, _p, _p, _p, _i, _p,
                                     _p, _p, _p, _p, _p, _p, _p, _i]
        L.orc_elastic3d.restype = _i
        L.orc_elastic3d.argtypes = [_i, _i, _i, _p, _p, _p, _f, _f, _i, _i, _i, _f, _i, _p, _p, _p, _p, _i, _p,
                                    _p, _p, _p, _p, _p, _p, _i]
        L.orc_elastic_params3d.restype = None
        L.orc_elastic_params3d.argtypes = [_p, _p, _p, _i, _i, _i, _p, _p, _p, _p, _p]
        L.orc_scalar2d.restype = _i
        L.orc_scalar2d.argtypes = [_i, _i, _p, _f, _f, _i, _i, _i, _f, _i, _p, _p, _p, _i, _p, _p, _p, _p, _i]
        L.orc_centered_d2.restype = _i
        L.orc_centered_d2.argtypes = [_i, _p, _p]
        L.orc_sponge_damp2d.restype = None
        L.orc_sponge_damp2d.argtypes = [_i, _i, _i, _f, _p]
        L.orc_fd_coeffs.restype = _i
        L.orc_fd_coeffs.argtypes = [_i, _p]
        L.orc_habc_coeffs.restype = None
        L.orc_habc_coeffs.argtypes = [_f, _f, _f, _p]
        L.orc_frame_total.restype = C.c_long
        L.orc_frame_total.argtypes = [_i, _i, _i]
        L.orc_frame_ij.restype = None
        L.orc_frame_ij.argtypes = [C.c_long, _i, _i, _i, C.POINTER(_i), C.POINTER(_i)]
        L.orc_pad2d.restype = None
        L.orc_pad2d.argtypes = [_p, _i, _i, _i, _p]
        L.orc_buoyancy2d.restype = None
        L.orc_buoyancy2d.argtypes = [_p, _i, _i, _p, _p]
        L.orc_elastic_params2d.restype = None
        L.orc_elastic_params2d.argtypes = [_p, _p, _p, _i, _i, _p, _p, _p, _p]
        L.orc_habc_weights2d.restype = None
        L.orc_habc_weights2d.argtypes = [_i, _i, _i, C.c_double, _p]
        L.orc_probe_acoustic_velocity.restype = None
        L.orc_probe_acoustic_velocity.argtypes = [_i, _i, _i, real, _p, _p, _p, _p, _p]
        L.orc_probe_acoustic_pressure.restype = None
        L.orc_probe_acoustic_pressure.argtypes = [_i, _i, _i, real, _p, _p, _p, _p]
        L.orc_probe_elastic_velocity.restype = None
        L.orc_probe_elastic_velocity.argtypes = [_i, _i, _i, real] + [_p] * 7
        L.orc_probe_elastic_stress.restype = None
        L.orc_probe_elastic_stress.argtypes = [_i, _i, _i, real] + [_p] * 8
        L.orc_probe_acoustic3d_velocity.restype = None
        L.orc_probe_acoustic3d_velocity.argtypes = [_i, _i, _i, _i, real] + [_p] * 7
        L.orc_probe_acoustic3d_pressure.restype = None
        L.orc_probe_acoustic3d_pressure.argtypes = [_i, _i, _i, _i, real] + [_p] * 5
        L.orc_probe_elastic3d_velocity.restype = None
        L.orc_probe_elastic3d_velocity.argtypes = [_i, _i, _i, _i, real] + [_p] * 4
        L.orc_probe_elastic3d_stress.restype = None
        L.orc_probe_elastic3d_stress.argtypes = [_i, _i, _i, _i, real] + [_p] * 6
        L.orc_probe_habc.restype = None
        L.orc_probe_habc.argtypes = [_i, _i, _i, real, real, real, _p, _p, _p]
        L.orc_max_threads.restype = _i
        L.orc_centered_d1.restype = _i
        L.orc_centered_d1.argtypes = [_i, _p]
        L.orc_coupled_medium.restype = _i
        L.orc_coupled_medium.argtypes = [_i, _i, _p, _p, _f, _i, _i, C.c_double, _p]
        L.orc_probe_coupled_velocity.restype = _i
        L.orc_probe_coupled_velocity.argtypes = [_i, _i, _i, _f, _f] + [_p] * 5
        L.orc_probe_coupled_position.restype = None
        L.orc_probe_coupled_position.argtypes = [_i, _i, _i, _f] + [_p] * 4
        L.orc_coupled2d.restype = _i
        L.orc_coupled2d.argtypes = [_i, _i, _p, _p, _f, _f, _i, _i, _i, _f, _f, C.c_double, _i, _p, _p, _p, _i,
                                    _p, _p, _i, _p, _p, _p, _p, _p, _i]
        _LIBS[precision] = L
    return _LIBS[precision]


def _dtype(precision: str):
    return np.float32 if precision == "f32" else np.float64


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data


def _f32F(a) -> np.ndarray:
    return np.asfortranarray(np.asarray(a, dtype=np.float32))


def _i32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32).ravel())


# ------------------------------------------------------------------------------------------------
# small restatements used by the pinning tests

def fd_coeffs(order: int) -> np.ndarray:
    out = np.zeros(5, np.float32)
    n = lib().orc_fd_coeffs(order, out.ctypes.data)
    if n < 0:
        raise ValueError(f"Unsupported FD order: {order}")
    return out[:n].copy()


def habc_coeffs(v_ref: float, dt: float, dh: float, precision: str = "f32") -> np.ndarray:
    out = np.zeros(3, _dtype(precision))
    lib(precision).orc_habc_coeffs(v_ref, dt, dh, out.ctypes.data)
    return out


def frame_total(nx: int, nz: int, nbc: int) -> int:
    return int(lib().orc_frame_total(nx, nz, nbc))


def frame_ij(idx: int, nx: int, nz: int, nbc: int) -> tuple[int, int]:
    i, j = _i(), _i()
    lib().orc_frame_ij(idx, nx, nz, nbc, C.byref(i), C.byref(j))
    return i.value, j.value


def pad2d(a, pad: int, precision: str = "f32") -> np.ndarray:
    a = _f32F(a)
    nx, nz = a.shape
    out = np.zeros((nx + 2 * pad, nz + 2 * pad), _dtype(precision), order="F")
    lib(precision).orc_pad2d(a.ctypes.data, nx, nz, pad, out.ctypes.data)
    return out


def buoyancy2d(rho_pad: np.ndarray, precision: str = "f32"):
    r = np.asfortranarray(rho_pad, dtype=_dtype(precision))
    bx = np.zeros_like(r)
    bz = np.zeros_like(r)
    lib(precision).orc_buoyancy2d(r.ctypes.data, r.shape[0], r.shape[1], bx.ctypes.data, bz.ctypes.data)
    return bx, bz


def elastic_params2d(vp_pad, vs_pad, rho_pad, precision: str = "f32"):
    dt = _dtype(precision)
    vp, vs, rho = (np.asfortranarray(x, dtype=dt) for x in (vp_pad, vs_pad, rho_pad))
    lam, mu, muxz, l2m = (np.zeros_like(vp) for _ in range(4))
    lib(precision).orc_elastic_params2d(vp.ctypes.data, vs.ctypes.data, rho.ctypes.data, vp.shape[0], vp.shape[1],
                                        lam.ctypes.data, mu.ctypes.data, muxz.ctypes.data, l2m.ctypes.data)
    return lam, mu, muxz, l2m


def habc_weights2d(nx: int, nz: int, pad: int, off: float, precision: str = "f32") -> np.ndarray:
    w = np.zeros((nx, nz), _dtype(precision), order="F")
    lib(precision).orc_habc_weights2d(nx, nz, pad, off, w.ctypes.data)
    return w


def probe_acoustic_velocity(vx, vz, p, bx, bz, fd_order: int, dtx: float, precision: str = "f32"):
    dt = _dtype(precision)
    vx, vz, p, bx, bz = (np.array(x, dtype=dt, order="F", copy=True) for x in (vx, vz, p, bx, bz))
    lib(precision).orc_probe_acoustic_velocity(vx.shape[0], vx.shape[1], fd_order, dtx, vx.ctypes.data,
                                               vz.ctypes.data, p.ctypes.data, bx.ctypes.data, bz.ctypes.data)
    return vx, vz


def probe_acoustic_pressure(p, vx, vz, kappa, fd_order: int, dtx: float, precision: str = "f32"):
    dt = _dtype(precision)
    p, vx, vz, kappa = (np.array(x, dtype=dt, order="F", copy=True) for x in (p, vx, vz, kappa))
    lib(precision).orc_probe_acoustic_pressure(p.shape[0], p.shape[1], fd_order, dtx, p.ctypes.data,
                                               vx.ctypes.data, vz.ctypes.data, kappa.ctypes.data)
    return p


def _fcopy(arrays, precision):
    dt = _dtype(precision)
    return [np.array(x, dtype=dt, order="F", copy=True) for x in arrays]


def probe_elastic_velocity(vx, vz, txx, tzz, txz, bx, bz, fd_order: int, dtx: float, precision: str = "f64"):
    """One elastic2d velocity sweep on a given padded (nx, nz) state; returns new (vx, vz)."""
    vx, vz, txx, tzz, txz, bx, bz = _fcopy((vx, vz, txx, tzz, txz, bx, bz), precision)
    lib(precision).orc_probe_elastic_velocity(vx.shape[0], vx.shape[1], fd_order, dtx,
                                              *(a.ctypes.data for a in (vx, vz, txx, tzz, txz, bx, bz)))
    return vx, vz


def probe_elastic_stress(txx, tzz, txz, vx, vz, lam, lam2mu, muxz, fd_order: int, dtx: float,
                         precision: str = "f64"):
    """One elastic2d stress sweep on a given padded state; returns new (txx, tzz, txz)."""
    arrs = _fcopy((txx, tzz, txz, vx, vz, lam, lam2mu, muxz), precision)
    lib(precision).orc_probe_elastic_stress(arrs[0].shape[0], arrs[0].shape[1], fd_order, dtx,
                                            *(a.ctypes.data for a in arrs))
    return arrs[0], arrs[1], arrs[2]


def probe_acoustic3d_velocity(vx, vy, vz, p, bx, by, bz, fd_order: int, dtdh: float, precision: str = "f64"):
    """One acoustic3d velocity sweep on a given padded (nx, ny, nz) state; returns (vx, vy, vz)."""
    arrs = _fcopy((vx, vy, vz, p, bx, by, bz), precision)
    lib(precision).orc_probe_acoustic3d_velocity(*arrs[0].shape, fd_order, dtdh, *(a.ctypes.data for a in arrs))
    return arrs[0], arrs[1], arrs[2]


def probe_acoustic3d_pressure(p, vx, vy, vz, kappa, fd_order: int, dtdh: float, precision: str = "f64"):
    """One acoustic3d pressure sweep on a given padded state; returns p."""
    arrs = _fcopy((p, vx, vy, vz, kappa), precision)
    lib(precision).orc_probe_acoustic3d_pressure(*arrs[0].shape, fd_order, dtdh, *(a.ctypes.data for a in arrs))
    return arrs[0]


def _fields9(F, precision):
    arrs = _fcopy(F, precision)
    ptrs = (C.c_void_p * 9)(*(a.ctypes.data for a in arrs))
    return arrs, ptrs


def probe_elastic3d_velocity(F, bx, by, bz, fd_order: int, dtdh: float, precision: str = "f64"):
    """One elastic3d velocity sweep.  F: (vx, vy, vz, txx, tyy, tzz, txy, txz, tyz); returns new F."""
    arrs, ptrs = _fields9(F, precision)
    m = _fcopy((bx, by, bz), precision)
    lib(precision).orc_probe_elastic3d_velocity(*arrs[0].shape, fd_order, dtdh, ptrs, *(a.ctypes.data for a in m))
    return arrs


def probe_elastic3d_stress(F, lam, lam2mu, mxy, mxz, myz, fd_order: int, dtdh: float, precision: str = "f64"):
    """One elastic3d stress sweep.  F as in probe_elastic3d_velocity; returns new F."""
    arrs, ptrs = _fields9(F, precision)
    m = _fcopy((lam, lam2mu, mxy, mxz, myz), precision)
    lib(precision).orc_probe_elastic3d_stress(*arrs[0].shape, fd_order, dtdh, ptrs, *(a.ctypes.data for a in m))
    return arrs


def probe_habc(f, fo, w, nbc: int, q: float, qt: float, qxt: float, precision: str = "f32"):
    dt = _dtype(precision)
    f, fo, w = (np.array(x, dtype=dt, order="F", copy=True) for x in (f, fo, w))
    lib(precision).orc_probe_habc(f.shape[0], f.shape[1], nbc, q, qt, qxt, f.ctypes.data, fo.ctypes.data,
                                  w.ctypes.data)
    return f


# ------------------------------------------------------------------------------------------------
# Ricker wavelet with Julia's range semantics

def _julia_rat(x: np.float32) -> tuple[int, int]:
    """Base.rat for a Float32 (denominator bound maxintfloat(Float16) = 2048)."""
    x = np.float32(x)
    y = np.float32(x)
    a, d, b, c = 1, 1, 0, 0
    m = 2048
    while abs(float(y)) <= m:
        f = int(np.trunc(y))
        y = np.float32(y - np.float32(f))
        a, c = f * a + c, a
        b, d = f * b + d, b
        if max(abs(a), abs(b)) > m:
            return c, d
        if np.float32(a) / np.float32(b) == x:
            break
        with np.errstate(divide="ignore"):
            y = np.float32(np.float32(1.0) / y)
    return a, b


def ricker(f0: float, dt: float, nt: int) -> np.ndarray:
    """ricker_wavelet(f0, dt, nt) (src/utils/ricker_wavelet.jl:6-12).

    `(0:nt-1) .* dt` is `range(0f0, step=dt, length=nt)`.  When Base.rat finds n/d with
    Float32(n/d) == dt the range stores ref/step as Float64 with step = n/d; otherwise they stay
    Float32.  `.- c` and `pf .*` keep the storage type (broadcast specialisations for ranges);
    elements are T(ref + (k-1)*step).  The square and product are Float32, exp is rounded once."""
    f0, dt = np.float32(f0), np.float32(dt)
    t0 = np.float32(np.float32(1.0) / f0)
    shift = np.float32(np.float32(1.5) * t0)
    pf = np.float32(np.float32(math.pi) * f0)
    sn, sd = _julia_rat(dt)
    wide = False
    if sd != 0 and np.float32(sn / sd) == dt:
        den = sd
        dst = np.float32(np.float32(den) * dt)
        if abs(float(dst)) <= 16777216.0:
            step_n = int(np.rint(dst))
            step64 = step_n / den
            wide = True
    k = np.arange(nt, dtype=np.float64)
    if wide:
        ref = 0.0 - float(shift)
        u = (float(pf) * ref + k * (float(pf) * step64)).astype(np.float32)
    else:
        ref = np.float32(np.float32(0.0) - shift)
        ur = np.float32(pf * ref)
        us = np.float32(pf * dt)
        u = (ur + (k.astype(np.float32) * us).astype(np.float32)).astype(np.float32)
    u2 = (u * u).astype(np.float32)
    e = np.exp(-u2.astype(np.float64)).astype(np.float32)
    return ((np.float32(1.0) - np.float32(2.0) * u2).astype(np.float32) * e).astype(np.float32)


# ------------------------------------------------------------------------------------------------
# full runs

def acoustic2d(vp, rho, dh, dt, nt, *, sx, sz, rx, rz, wavelet, nbc=50, fd_order=8, v_ref=None,
               snap_interval=0, precision="f32", nthreads=0, final_state=False, habc_width=0, sponge_factor=0.0):
    """acoustic2d (HABC, det path).  wavelet: (n_src, nt).  Returns dict of outputs."""
    vp, rho = _f32F(vp), _f32F(rho)
    nx, nz = vp.shape
    if v_ref is None:
        v_ref = float(vp[vp > 0].min())
    sx, sz, rx, rz = _i32(sx), _i32(sz), _i32(rx), _i32(rz)
    w = np.asfortranarray(np.asarray(wavelet, np.float32).reshape(len(sx), nt))
    nrec = len(rx)
    dt_ = _dtype(precision)
    sp, svx, svz = (np.zeros((nrec, nt), dt_, order="F") for _ in range(3))
    nsnap = nt // snap_interval if snap_interval > 0 else 0
    snaps = np.zeros((nsnap, nz, nx), dt_) if nsnap else None  # snap k is snaps[k].T  (x fastest)
    pad = nbc + fd_order // 2
    fs = np.zeros(3 * (nx + 2 * pad) * (nz + 2 * pad), dt_) if final_state else None
    rc = lib(precision).orc_acoustic2d(nx, nz, vp.ctypes.data, rho.ctypes.data, dh, dt, nt, nbc, fd_order, v_ref,
                                       len(sx), sx.ctypes.data, sz.ctypes.data, w.ctypes.data, nrec, rx.ctypes.data,
                                       rz.ctypes.data, sp.ctypes.data, svx.ctypes.data, svz.ctypes.data,
                                       snap_interval, _ptr(snaps), _ptr(fs), habc_width, sponge_factor, nthreads)
    if rc != 0:
        raise ValueError(f"oracle acoustic2d failed ({rc})")
    out = dict(seis_p=sp, seis_vx=svx, seis_vz=svz,
               snaps=[s.T for s in snaps] if snaps is not None else [])
    if final_state:
        shp = (nx + 2 * pad, nz + 2 * pad)
        n = shp[0] * shp[1]
        out["state"] = {k: fs[m * n:(m + 1) * n].reshape(shp, order="F") for m, k in enumerate(("p", "vx", "vz"))}
    return out


def centered_d2(order: int):
    """(c0, d2) of src/utils/FD_centered.jl:59-62 as the oracle stores them (Float32)."""
    c0 = np.zeros(1, np.float32)
    d2 = np.zeros(5, np.float32)
    n = lib("f32").orc_centered_d2(order, c0.ctypes.data, d2.ctypes.data)
    if n < 0:
        raise ValueError(f"Unsupported FD order: {order} (centered d2)")
    return float(c0[0]), d2[:n].copy()


def scalar2d(vp, dh, dt, nt, *, sx, sz, rx, rz, wavelet, nbc=50, fd_order=8, v_ref=None, precision="f32",
             nthreads=0, final_state=False):
    """scalar2d (src/equations/scalar2d/scalar2d.jl:33-81).  wavelet: (n_src, nt), already scaled as
    source_scale asks.  Returns dict with seis_u (and state["u"] = u^nt when final_state)."""
    vp = _f32F(vp)
    nx, nz = vp.shape
    if v_ref is None:
        v_ref = float(vp[vp > 0].min())
    sx, sz, rx, rz = _i32(sx), _i32(sz), _i32(rx), _i32(rz)
    w = np.asfortranarray(np.asarray(wavelet, np.float32).reshape(len(sx), nt))
    nrec = len(rx)
    dt_ = _dtype(precision)
    su = np.zeros((nrec, nt), dt_, order="F")
    pad = nbc + fd_order // 2
    shp = (nx + 2 * pad, nz + 2 * pad)
    fs = np.zeros(shp[0] * shp[1], dt_) if final_state else None
    rc = lib(precision).orc_scalar2d(nx, nz, vp.ctypes.data, dh, dt, nt, nbc, fd_order, v_ref, len(sx),
                                     sx.ctypes.data, sz.ctypes.data, w.ctypes.data, nrec, rx.ctypes.data,
                                     rz.ctypes.data, su.ctypes.data, _ptr(fs), nthreads)
    if rc != 0:
        raise ValueError(f"oracle scalar2d failed ({rc})")
    out = dict(seis_u=su)
    if final_state:
        out["state"] = {"u": fs.reshape(shp, order="F")}
    return out


def centered_d1(order: int) -> np.ndarray:
    """d1 of src/utils/FD_centered.jl:48-51 as the oracle stores them (Float32)."""
    d1 = np.zeros(5, np.float32)
    n = lib("f32").orc_centered_d1(order, d1.ctypes.data)
    if n < 0:
        raise ValueError(f"Unsupported FD order: {order} (centered d1)")
    return d1[:n].copy()


COUPLED_MEDIUM = ("alpha", "beta", "dalpha_dx", "dalpha_dz", "dbeta_dx", "dbeta_dz", "lap_alpha", "lap_beta")


def coupled_medium(vp, vs, dh, nbc, fd_order, smooth_sigma=3.0, precision="f32"):
    """init_coupled_medium (src/equations/coupled2d/medium.jl:50-117) on the padded grid: a dict of
    the eight (nx+2pad, nz+2pad) arrays named as in COUPLED_MEDIUM."""
    vp, vs = _f32F(vp), _f32F(vs)
    nx, nz = vp.shape
    pad = nbc + fd_order // 2
    shp = (nx + 2 * pad, nz + 2 * pad)
    n = shp[0] * shp[1]
    out = np.zeros(8 * n, _dtype(precision))
    rc = lib(precision).orc_coupled_medium(nx, nz, vp.ctypes.data, vs.ctypes.data, dh, nbc, fd_order,
                                           float(smooth_sigma), out.ctypes.data)
    if rc != 0:
        raise ValueError(f"oracle coupled_medium failed ({rc})")
    return {k: out[m * n:(m + 1) * n].reshape(shp, order="F") for m, k in enumerate(COUPLED_MEDIUM)}


def probe_coupled_velocity(dPdt, dSdt, P, S, med, fd_order, dt, dh, precision="f64"):
    """One velocity half-step (update_fields.jl:27-112) on a given padded state; med: a dict as
    coupled_medium returns.  Returns the new (dPdt, dSdt)."""
    dt_ = _dtype(precision)
    a, b, p, s = (np.asfortranarray(np.asarray(x, dt_)).copy(order="F") for x in (dPdt, dSdt, P, S))
    m = np.concatenate([np.asarray(med[k], dt_).ravel(order="F") for k in COUPLED_MEDIUM])
    nx, nz = a.shape
    rc = lib(precision).orc_probe_coupled_velocity(nx, nz, fd_order, dt, dh, a.ctypes.data, b.ctypes.data,
                                                   p.ctypes.data, s.ctypes.data, m.ctypes.data)
    if rc != 0:
        raise ValueError("oracle probe_coupled_velocity failed")
    return a, b


def probe_coupled_position(P, S, dPdt, dSdt, fd_order, dt, precision="f64"):
    """One position half-step (update_fields.jl:124-144).  Returns the new (P, S)."""
    dt_ = _dtype(precision)
    p, s, a, b = (np.asfortranarray(np.asarray(x, dt_)).copy(order="F") for x in (P, S, dPdt, dSdt))
    nx, nz = p.shape
    lib(precision).orc_probe_coupled_position(nx, nz, fd_order, dt, p.ctypes.data, s.ctypes.data, a.ctypes.data,
                                              b.ctypes.data)
    return p, s


def coupled2d(vp, vs, dh, dt, nt, *, sx, sz, rx, rz, wavelet, nbc=50, fd_order=8, snap_interval=0, v_ref_p=None,
              v_ref_s=None, smooth_sigma=3.0, precision="f32", nthreads=0, final_state=False):
    """coupled2d (src/equations/coupled2d/coupled2d.jl:53-260) with the deterministic two-pass
    Higdon boundary.  wavelet: (n_src, nt).  Returns dict seis_P, seis_S, snaps_P, snaps_S (and
    state P, S, dPdt, dSdt when final_state)."""
    vp, vs = _f32F(vp), _f32F(vs)
    nx, nz = vp.shape
    if v_ref_p is None:
        v_ref_p = float(vp[vp > 0].min())
    if v_ref_s is None:
        v_ref_s = float(vs[vs > 0].min())
    sx, sz, rx, rz = _i32(sx), _i32(sz), _i32(rx), _i32(rz)
    w = np.asfortranarray(np.asarray(wavelet, np.float32).reshape(len(sx), nt))
    nrec = len(rx)
    dt_ = _dtype(precision)
    sP, sS = (np.zeros((nrec, nt), dt_, order="F") for _ in range(2))
    nsnap = nt // snap_interval if snap_interval > 0 else 0
    snP = np.zeros((nsnap, nz, nx), dt_) if nsnap else None
    snS = np.zeros((nsnap, nz, nx), dt_) if nsnap else None
    pad = nbc + fd_order // 2
    shp = (nx + 2 * pad, nz + 2 * pad)
    n = shp[0] * shp[1]
    fs = np.zeros(4 * n, dt_) if final_state else None
    rc = lib(precision).orc_coupled2d(nx, nz, vp.ctypes.data, vs.ctypes.data, dh, dt, nt, nbc, fd_order, v_ref_p,
                                      v_ref_s, float(smooth_sigma), len(sx), sx.ctypes.data, sz.ctypes.data,
                                      w.ctypes.data, nrec, rx.ctypes.data, rz.ctypes.data, snap_interval,
                                      sP.ctypes.data, sS.ctypes.data, _ptr(snP), _ptr(snS), _ptr(fs), nthreads)
    if rc != 0:
        raise ValueError(f"oracle coupled2d failed ({rc})")
    out = dict(seis_P=sP, seis_S=sS, snaps_P=[s.T for s in snP] if snP is not None else [],
               snaps_S=[s.T for s in snS] if snS is not None else [])
    if final_state:
        out["state"] = {k: fs[m * n:(m + 1) * n].reshape(shp, order="F")
                        for m, k in enumerate(("P", "S", "dPdt", "dSdt"))}
    return out


def elastic2d(vp, vs, rho, dh, dt, nt, *, sx, sz, rx, rz, wavelet, nbc=50, fd_order=8, v_ref=None,
              source_type=0, snap_interval=0, precision="f32", nthreads=0, final_state=False, habc_width=0,
              sponge_factor=0.0):
    vp, vs, rho = _f32F(vp), _f32F(vs), _f32F(rho)
    nx, nz = vp.shape
    if v_ref is None:
        v_ref = float(vp[vp > 0].min())
    sx, sz, rx, rz = _i32(sx), _i32(sz), _i32(rx), _i32(rz)
    w = np.asfortranarray(np.asarray(wavelet, np.float32).reshape(len(sx), nt))
    nrec = len(rx)
    dt_ = _dtype(precision)
    svx, svz = (np.zeros((nrec, nt), dt_, order="F") for _ in range(2))
    nsnap = nt // snap_interval if snap_interval > 0 else 0
    snaps = np.zeros((nsnap, nz, nx), dt_) if nsnap else None
    pad = nbc + fd_order // 2
    fs = np.zeros(5 * (nx + 2 * pad) * (nz + 2 * pad), dt_) if final_state else None
    rc = lib(precision).orc_elastic2d(nx, nz, vp.ctypes.data, vs.ctypes.data, rho.ctypes.data, dh, dt, nt, nbc,
                                      fd_order, v_ref, source_type, len(sx), sx.ctypes.data, sz.ctypes.data,
                                      w.ctypes.data, nrec, rx.ctypes.data, rz.ctypes.data, svx.ctypes.data,
                                      svz.ctypes.data, snap_interval, _ptr(snaps), _ptr(fs), habc_width, sponge_factor,
                                      nthreads)
    if rc != 0:
        raise ValueError(f"oracle elastic2d failed ({rc})")
    out = dict(seis_vx=svx, seis_vz=svz, snaps=[s.T for s in snaps] if snaps is not None else [])
    if final_state:
        shp = (nx + 2 * pad, nz + 2 * pad)
        n = shp[0] * shp[1]
        out["state"] = {k: fs[m * n:(m + 1) * n].reshape(shp, order="F")
                        for m, k in enumerate(("vx", "vz", "txx", "tzz", "txz"))}
    return out


def acoustic3d(vp, rho, dh, dt, nt, *, sx, sy, sz, rx, ry, rz, wavelet, nbc=50, fd_order=8, v_ref=None,
               precision="f32", nthreads=0, final_state=False):
    vp, rho = _f32F(vp), _f32F(rho)
    nx, ny, nz = vp.shape
    if v_ref is None:
        v_ref = float(vp[vp > 0].min())
    sx, sy, sz, rx, ry, rz = (_i32(a) for a in (sx, sy, sz, rx, ry, rz))
    w = np.asfortranarray(np.asarray(wavelet, np.float32).reshape(len(sx), nt))
    nrec = len(rx)
    dt_ = _dtype(precision)
    svx, svy, svz, sp = (np.zeros((nrec, nt), dt_, order="F") for _ in range(4))
    pad = nbc + fd_order // 2
    shp = (nx + 2 * pad, ny + 2 * pad, nz + 2 * pad)
    n = shp[0] * shp[1] * shp[2]
    fs = np.zeros(4 * n, dt_) if final_state else None
    rc = lib(precision).orc_acoustic3d(nx, ny, nz, vp.ctypes.data, rho.ctypes.data, dh, dt, nt, nbc, fd_order,
                                       v_ref, len(sx), sx.ctypes.data, sy.ctypes.data, sz.ctypes.data, w.ctypes.data,
                                       nrec, rx.ctypes.data, ry.ctypes.data, rz.ctypes.data, svx.ctypes.data,
                                       svy.ctypes.data, svz.ctypes.data, sp.ctypes.data, _ptr(fs), nthreads)
    if rc != 0:
        raise ValueError(f"oracle acoustic3d failed ({rc})")
    out = dict(seis_vx=svx, seis_vy=svy, seis_vz=svz, seis_p=sp)
    if final_state:
        out["state"] = {k: fs[m * n:(m + 1) * n].reshape(shp, order="F") for m, k in enumerate(("p", "vx", "vy", "vz"))}
    return out


def elastic3d(vp, vs, rho, dh, dt, nt, *, sx, sy, sz, rx, ry, rz, wavelet, nbc=50, fd_order=8, v_ref=None,
              precision="f32", nthreads=0, final_state=False):
    """elastic3d with the two-pass HABC (fomo_oracle3d.c:orc_elastic3d)."""
    vp, vs, rho = _f32F(vp), _f32F(vs), _f32F(rho)
    nx, ny, nz = vp.shape
    if v_ref is None:
        v_ref = float(vp[vp > 0].min())
    sx, sy, sz, rx, ry, rz = (_i32(a) for a in (sx, sy, sz, rx, ry, rz))
    w = np.asfortranarray(np.asarray(wavelet, np.float32).reshape(len(sx), nt))
    nrec = len(rx)
    dt_ = _dtype(precision)
    svx, svy, svz = (np.zeros((nrec, nt), dt_, order="F") for _ in range(3))
    pad = nbc + fd_order // 2
    shp = (nx + 2 * pad, ny + 2 * pad, nz + 2 * pad)
    n = shp[0] * shp[1] * shp[2]
    fs = np.zeros(9 * n, dt_) if final_state else None
    rc = lib(precision).orc_elastic3d(nx, ny, nz, vp.ctypes.data, vs.ctypes.data, rho.ctypes.data, dh, dt, nt, nbc,
                                      fd_order, v_ref, len(sx), sx.ctypes.data, sy.ctypes.data, sz.ctypes.data,
                                      w.ctypes.data, nrec, rx.ctypes.data, ry.ctypes.data, rz.ctypes.data,
                                      svx.ctypes.data, svy.ctypes.data, svz.ctypes.data, _ptr(fs), nthreads)
    if rc != 0:
        raise ValueError(f"oracle elastic3d failed ({rc})")
    out = dict(seis_vx=svx, seis_vy=svy, seis_vz=svz)
    if final_state:
        names = ("vx", "vy", "vz", "txx", "tyy", "tzz", "txy", "txz", "tyz")
        out["state"] = {k: fs[m * n:(m + 1) * n].reshape(shp, order="F") for m, k in enumerate(names)}
    return out


def elastic_params3d(vp, vs, rho, precision="f32"):
    """lam, lam2mu, mu_txy, mu_txz, mu_tyz of padded (nx, ny, nz) arrays."""
    dt_ = _dtype(precision)
    vp, vs, rho = (np.asfortranarray(np.asarray(a, dt_)) for a in (vp, vs, rho))
    outs = [np.zeros(vp.shape, dt_, order="F") for _ in range(5)]
    lib(precision).orc_elastic_params3d(vp.ctypes.data, vs.ctypes.data, rho.ctypes.data, *vp.shape,
                                        *(o.ctypes.data for o in outs))
    return outs


def sponge_damp2d(nx: int, nz: int, nbc: int, factor: float = 0.015, precision: str = "f32") -> np.ndarray:
    """Cerjan damping profile on the padded (nx, nz) grid (src/boundary/sponge.jl:30-50)."""
    out = np.zeros((nx, nz), _dtype(precision), order="F")
    lib(precision).orc_sponge_damp2d(nx, nz, nbc, factor, out.ctypes.data)
    return out
