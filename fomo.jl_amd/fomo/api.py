"""Drop-in mirror of Fomo.jl's forward-modeling API over the native engine.

Same names, keywords, defaults, validation and result fields as the reference:
  acoustic2d  src/equations/acoustic2d/acoustic2d.jl:47-192
  elastic2d   src/equations/elastic2d/elastic2d.jl:45-163
Models are numpy arrays indexed [i, j] (i along x) like the reference's Julia matrices; source
and receiver indices are 1-based.  Seismograms come back as (n_rec, nt) arrays, snapshots as a list
of (nx, nz) arrays.
"""
from __future__ import annotations

import ctypes as C
from collections import namedtuple

import numpy as np

from . import _lib
from ._lib import ArgumentError, Desc
from .checks import check_geometry, check_numerics
from .setup import ricker_wavelet

Stats = namedtuple("Stats", ["kernel_time_s"])
Acoustic2DResult = namedtuple("Acoustic2DResult", ["seis_p", "seis_vx", "seis_vz", "snaps", "stats"])
Elastic2DResult = namedtuple("Elastic2DResult", ["seis_vx", "seis_vz", "snaps", "stats"])


def _sym(x) -> str:
    return str(x).lstrip(":")


def _f32F(a) -> np.ndarray:
    return np.asfortranarray(np.asarray(a, dtype=np.float32))


def _i32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.int64).ravel().astype(np.int32))


def _default_vref(vp: np.ndarray) -> float:
    pos = vp[vp > np.float32(0)]
    if pos.size == 0:
        raise ArgumentError("vp has no positive entry")
    return float(np.float32(pos.min()))


def wavelet_matrix(wavelet, f0, dt, nt, n_src, scale_source_by_dt) -> np.ndarray:
    """acoustic2d.jl:91-97: Ricker(f0) or the user's wavelet, repeated per source, optionally * dt."""
    w = ricker_wavelet(f0, dt, nt) if wavelet is None else np.asarray(wavelet, dtype=np.float64).astype(np.float32)
    w = np.asarray(w, np.float32).ravel()
    if len(w) != nt:
        raise ArgumentError(f"wavelet length {len(w)} != nt={nt}")
    m = np.repeat(w[None, :], n_src, axis=0)
    if scale_source_by_dt:
        m = (m * np.float32(dt)).astype(np.float32)
    return np.asfortranarray(m)


def _desc(eq, nx, nz, nbc, fd_order, nt, dh, dt, v_ref, n_src, n_rec, snap_interval, use_graph,
          source_type=0, ny=1, storage=0) -> Desc:
    d = Desc()
    d.eq, d.nx, d.ny, d.nz = eq, nx, ny, nz
    d.nbc, d.fd_order, d.nt = nbc, fd_order, nt
    d.dh, d.dt, d.v_ref = dh, dt, v_ref
    d.n_src, d.n_rec = n_src, n_rec
    d.source_type, d.storage = source_type, storage
    d.snap_interval = snap_interval
    d.use_graph = 1 if use_graph else 0
    return d


def _common_checks(boundary, nbc, fd_order, nt, snap_interval):
    if _sym(boundary) not in ("habc", "sponge"):
        raise ArgumentError(f"boundary must be :habc or :sponge, got {boundary}")
    if _sym(boundary) == "sponge":
        raise _lib.FomoError("boundary=:sponge is not provided by this engine (HABC only)")
    if nbc < 0 or nt < 0 or snap_interval < 0:
        raise ArgumentError("nbc, nt and snap_interval must be >= 0")
    if fd_order not in (2, 4, 6, 8, 10):
        raise ValueError(f"Unsupported FD order: {fd_order}")


def acoustic2d(vp, rho, dh, dt, nt, f0, *, sx, sz, rx, rz, nbc=50, fd_order=8, snap_interval=0,
               boundary="habc", v_ref=None, sponge_factor=0.015, wavelet=None, scale_source_by_dt=False,
               use_cuda_graph=True, verbose=True):
    """acoustic2d(vp, rho, dh, dt, nt, f0; sx, sz, rx, rz, ...) -> (seis_p, seis_vx, seis_vz, snaps, stats)."""
    del sponge_factor, verbose
    _common_checks(boundary, nbc, fd_order, nt, snap_interval)
    vp, rho = _f32F(vp), _f32F(rho)
    if vp.ndim != 2 or vp.shape != rho.shape:
        raise ArgumentError("vp and rho must be matrices of the same size")
    nx, nz = vp.shape
    check_geometry(nx, nz, sx, sz, rx, rz)
    vmin = _default_vref(vp)
    check_numerics(float(vp.max()), vmin, dh, dt, f0, fd_order)
    v_ref = vmin if v_ref is None else float(np.float32(v_ref))
    sx, sz, rx, rz = _i32(sx), _i32(sz), _i32(rx), _i32(rz)
    if len(sx) != len(sz) or len(rx) != len(rz):
        raise ArgumentError("source / receiver coordinate vectors differ in length")
    wm = wavelet_matrix(wavelet, f0, dt, nt, len(sx), scale_source_by_dt)
    n_rec = len(rx)
    d = _desc(_lib.ACOUSTIC2D, nx, nz, nbc, fd_order, nt, dh, dt, v_ref, len(sx), n_rec, snap_interval,
              use_cuda_graph)
    sp, svx, svz = (np.zeros((n_rec, nt), np.float32, order="F") for _ in range(3))
    nsnap = nt // snap_interval if snap_interval > 0 else 0
    snaps = np.zeros((nsnap, nz, nx), np.float32) if nsnap else None
    t = C.c_double(0.0)
    _lib.check(_lib.lib().fomo_acoustic2d(C.byref(d), vp.ctypes.data, rho.ctypes.data, sx.ctypes.data,
                                          sz.ctypes.data, wm.ctypes.data, rx.ctypes.data, rz.ctypes.data,
                                          sp.ctypes.data, svx.ctypes.data, svz.ctypes.data,
                                          None if snaps is None else snaps.ctypes.data, C.byref(t)))
    snap_list = [s.T for s in snaps] if snaps is not None else []
    return Acoustic2DResult(sp, svx, svz, snap_list, Stats(float(t.value)))


def elastic2d(vp, vs, rho, dh, dt, nt, f0, *, sx, sz, rx, rz, nbc=50, fd_order=8, snap_interval=0,
              boundary="habc", v_ref=None, sponge_factor=0.015, wavelet=None, scale_source_by_dt=False,
              use_cuda_graph=True, verbose=True, source_type="explosive"):
    """elastic2d(vp, vs, rho, dh, dt, nt, f0; ...) -> (seis_vx, seis_vz, snaps, stats).

    `source_type` (extension, default :explosive = the reference's τxx+τzz source); :fz adds the
    wavelet to vz at the end of the velocity phase (a vertical force)."""
    del sponge_factor, verbose
    _common_checks(boundary, nbc, fd_order, nt, snap_interval)
    st = _sym(source_type)
    if st not in ("explosive", "fz"):
        raise ArgumentError(f"source_type must be :explosive or :fz, got {source_type}")
    vp, vs, rho = _f32F(vp), _f32F(vs), _f32F(rho)
    if vp.ndim != 2 or vp.shape != rho.shape or vp.shape != vs.shape:
        raise ArgumentError("vp, vs and rho must be matrices of the same size")
    nx, nz = vp.shape
    check_geometry(nx, nz, sx, sz, rx, rz)
    vs_pos = vs[vs > np.float32(0)]
    vmin_disp = float(vs_pos.min()) if vs_pos.size else _default_vref(vp)
    check_numerics(float(vp.max()), vmin_disp, dh, dt, f0, fd_order)
    v_ref = _default_vref(vp) if v_ref is None else float(np.float32(v_ref))
    sx, sz, rx, rz = _i32(sx), _i32(sz), _i32(rx), _i32(rz)
    if len(sx) != len(sz) or len(rx) != len(rz):
        raise ArgumentError("source / receiver coordinate vectors differ in length")
    wm = wavelet_matrix(wavelet, f0, dt, nt, len(sx), scale_source_by_dt)
    n_rec = len(rx)
    d = _desc(_lib.ELASTIC2D, nx, nz, nbc, fd_order, nt, dh, dt, v_ref, len(sx), n_rec, snap_interval,
              use_cuda_graph, source_type=_lib.SRC_FZ if st == "fz" else _lib.SRC_EXPLOSIVE)
    svx, svz = (np.zeros((n_rec, nt), np.float32, order="F") for _ in range(2))
    nsnap = nt // snap_interval if snap_interval > 0 else 0
    snaps = np.zeros((nsnap, nz, nx), np.float32) if nsnap else None
    t = C.c_double(0.0)
    _lib.check(_lib.lib().fomo_elastic2d(C.byref(d), vp.ctypes.data, vs.ctypes.data, rho.ctypes.data,
                                         sx.ctypes.data, sz.ctypes.data, wm.ctypes.data, rx.ctypes.data,
                                         rz.ctypes.data, svx.ctypes.data, svz.ctypes.data,
                                         None if snaps is None else snaps.ctypes.data, C.byref(t)))
    snap_list = [s.T for s in snaps] if snaps is not None else []
    return Elastic2DResult(svx, svz, snap_list, Stats(float(t.value)))


class Plan:
    """Device-resident run (fomo_plan): setup once, then advance / read / reset.

    Used by the benchmark and the tests; `acoustic2d`/`elastic2d` are one-shot wrappers of the
    same machinery inside the library."""

    def __init__(self, eq, vp, rho, dh, dt, nt, *, sx, sz, rx, rz, wavelet, vs=None, nbc=50, fd_order=8,
                 v_ref=None, source_type=0, use_graph=True, sy=None, ry=None, storage=0, record_p=False):
        vp, rho = _f32F(vp), _f32F(rho)
        self._keep = [vp, rho]
        if vs is not None:
            vs = _f32F(vs)
            self._keep.append(vs)
        shape = vp.shape
        nx, nz = shape[0], shape[-1]
        ny = shape[1] if len(shape) == 3 else 1
        v_ref = _default_vref(vp) if v_ref is None else float(np.float32(v_ref))
        sx, sz, rx, rz = _i32(sx), _i32(sz), _i32(rx), _i32(rz)
        sy = _i32(sy) if sy is not None else None
        ry = _i32(ry) if ry is not None else None
        wm = np.asfortranarray(np.asarray(wavelet, np.float32).reshape(len(sx), nt))
        self.n_rec, self.nt, self.eq = len(rx), nt, eq
        d = _desc(eq, nx, nz, nbc, fd_order, nt, dh, dt, v_ref, len(sx), len(rx), 0, use_graph,
                  source_type=source_type, ny=ny, storage=storage)
        d.record_p = 1 if record_p else 0
        h = C.c_void_p()
        ptr = lambda a: None if a is None else a.ctypes.data  # noqa: E731
        _lib.check(_lib.lib().fomo_plan_create(C.byref(d), vp.ctypes.data, ptr(vs), rho.ctypes.data,
                                               sx.ctypes.data, ptr(sy), sz.ctypes.data, wm.ctypes.data,
                                               rx.ctypes.data, ptr(ry), rz.ctypes.data, C.byref(h)))
        self._h = h
        a, b, c, p = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
        _lib.check(_lib.lib().fomo_plan_dims(h, C.byref(a), C.byref(b), C.byref(c), C.byref(p)))
        self.padded = (a.value, c.value) if ny == 1 else (a.value, b.value, c.value)
        self.pad = p.value

    def reset(self):
        _lib.check(_lib.lib().fomo_plan_reset(self._h))

    def advance(self, n: int):
        _lib.check(_lib.lib().fomo_plan_advance(self._h, int(n)))

    def timed_advance(self, n: int) -> float:
        ms = C.c_double(0.0)
        _lib.check(_lib.lib().fomo_plan_timed_advance(self._h, int(n), C.byref(ms)))
        return ms.value

    def sync(self):
        _lib.check(_lib.lib().fomo_plan_sync(self._h))

    def kernel_times(self, n: int) -> tuple[float, float, float]:
        """Average ms per launch of (velocity sweep, pressure/stress sweep, gather) over n steps."""
        out = (C.c_double * 3)()
        _lib.check(_lib.lib().fomo_plan_kernel_times(self._h, int(n), out))
        return out[0], out[1], out[2]

    def seis(self, nslots: int = 3):
        outs = [np.zeros((self.n_rec, self.nt), np.float32, order="F") for _ in range(nslots)]
        ptrs = [o.ctypes.data for o in outs] + [None] * (4 - nslots)
        _lib.check(_lib.lib().fomo_plan_read_seis(self._h, *ptrs))
        return outs

    def field(self, which: int) -> np.ndarray:
        out = np.zeros(self.padded, np.float32, order="F")
        _lib.check(_lib.lib().fomo_plan_read_field(self._h, int(which), out.ctypes.data))
        return out

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().fomo_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
