"""Drop-in mirror of Fomo.jl's forward-modeling API over the native engine.

Same names, keywords, defaults, validation and result fields as the reference:
  acoustic2d  src/equations/acoustic2d/acoustic2d.jl:47-192
  elastic2d   src/equations/elastic2d/elastic2d.jl:45-163
  coupled2d   src/equations/coupled2d/coupled2d.jl:53-183
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
from .checks import check_geometry, check_numerics, check_numerics_centered
from .setup import ricker_wavelet

Stats = namedtuple("Stats", ["kernel_time_s"])
Acoustic2DResult = namedtuple("Acoustic2DResult", ["seis_p", "seis_vx", "seis_vz", "snaps", "stats"])
Elastic2DResult = namedtuple("Elastic2DResult", ["seis_vx", "seis_vz", "snaps", "stats"])
Scalar2DResult = namedtuple("Scalar2DResult", ["seis_u", "stats"])
Coupled2DResult = namedtuple("Coupled2DResult", ["seis_P", "seis_S", "snaps_P", "snaps_S", "stats"])


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
          source_type=0, ny=1, storage=0, boundary="habc", sponge_factor=0.015) -> Desc:
    d = Desc()
    d.boundary = _lib.BOUNDARY_SPONGE if _sym(boundary) == "sponge" else _lib.BOUNDARY_HABC
    d.sponge_factor = float(np.float32(sponge_factor))
    d.eq, d.nx, d.ny, d.nz = eq, nx, ny, nz
    d.nbc, d.fd_order, d.nt = nbc, fd_order, nt
    d.dh, d.dt, d.v_ref = dh, dt, v_ref
    d.n_src, d.n_rec = n_src, n_rec
    d.source_type, d.storage = source_type, storage
    d.snap_interval = snap_interval
    d.use_graph = 1 if use_graph else 0
    return d


def _common_checks(boundary, nbc, fd_order, nt, snap_interval, learned_ok=False):
    allowed = ("habc", "sponge", "learned") if learned_ok else ("habc", "sponge")
    if _sym(boundary) not in allowed:
        raise ArgumentError(f"boundary must be one of {', '.join(':' + a for a in allowed)}, got {boundary}")
    if nbc < 0 or nt < 0 or snap_interval < 0:
        raise ArgumentError("nbc, nt and snap_interval must be >= 0")
    if fd_order not in (2, 4, 6, 8, 10):
        raise ValueError(f"Unsupported FD order: {fd_order}")


def acoustic2d(vp, rho, dh, dt, nt, f0, *, sx, sz, rx, rz, nbc=50, fd_order=8, snap_interval=0,
               boundary="habc", v_ref=None, sponge_factor=0.015, wavelet=None, scale_source_by_dt=False,
               use_cuda_graph=True, verbose=True, learned_weights=None, storage="f32"):
    """acoustic2d(vp, rho, dh, dt, nt, f0; sx, sz, rx, rz, ...) -> (seis_p, seis_vx, seis_vz, snaps, stats).

    boundary = :habc (deterministic two-pass Higdon), :sponge (Cerjan, `sponge_factor`) or :learned
    (an extension: the Higdon frame with MLP-corrected blend weights, fomo.learned, trained for thin
    layers such as nbc = 16; `learned_weights` overrides the shipped network).
    `storage` (extension): :f32 (default, bit-identical to the reference), :f16 or :bf16 storage of
    wavefields and materials with per-field scales, or :mat16 (Float32 wavefields, Float16
    materials); arithmetic stays Float32."""
    del verbose
    _common_checks(boundary, nbc, fd_order, nt, snap_interval, learned_ok=True)
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
    if _sym(boundary) == "learned":
        return _acoustic2d_learned(vp, rho, dh, dt, nt, sx, sz, rx, rz, wm, nbc, fd_order, v_ref, snap_interval,
                                   use_cuda_graph, learned_weights, storage)
    d = _desc(_lib.ACOUSTIC2D, nx, nz, nbc, fd_order, nt, dh, dt, v_ref, len(sx), n_rec, snap_interval,
              use_cuda_graph, boundary=boundary, sponge_factor=sponge_factor, storage=storage_code(storage))
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


def _acoustic2d_learned(vp, rho, dh, dt, nt, sx, sz, rx, rz, wm, nbc, fd_order, v_ref, snap_interval, use_graph,
                        weights, storage="f32"):
    """boundary = :learned through a device-resident plan: enable the network, run the loop (timed
    with hipEvents), copy snapshots of the interior of p between segments, read the seismograms."""
    from . import learned
    w = learned.load_weights() if weights is None else np.asarray(weights, np.float32)
    p = Plan(_lib.ACOUSTIC2D, vp, rho, dh, dt, nt, sx=sx, sz=sz, rx=rx, rz=rz, wavelet=wm, nbc=nbc,
             fd_order=fd_order, v_ref=v_ref, use_graph=use_graph, storage=storage)
    try:
        learned.enable(p, w)
        snaps = []
        ms = 0.0
        if snap_interval > 0:
            pad = p.pad
            for _ in range(nt // snap_interval):
                ms += p.timed_advance(snap_interval)
                snaps.append(np.array(p.field(0)[pad:-pad, pad:-pad]))
            rest = nt - (nt // snap_interval) * snap_interval
            if rest:
                ms += p.timed_advance(rest)
        else:
            ms = p.timed_advance(nt)
        sp, svx, svz = p.seis(3)
    finally:
        p.close()
    return Acoustic2DResult(sp, svx, svz, snaps, Stats(ms * 1e-3))


def elastic2d(vp, vs, rho, dh, dt, nt, f0, *, sx, sz, rx, rz, nbc=50, fd_order=8, snap_interval=0,
              boundary="habc", v_ref=None, sponge_factor=0.015, wavelet=None, scale_source_by_dt=False,
              use_cuda_graph=True, verbose=True, source_type="explosive", storage="f32"):
    """elastic2d(vp, vs, rho, dh, dt, nt, f0; ...) -> (seis_vx, seis_vz, snaps, stats).

    `source_type` (extension, default :explosive = the reference's τxx+τzz source); :fz adds the
    wavelet to vz at the end of the velocity phase (a vertical force).  `storage` as acoustic2d."""
    del verbose
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
              use_cuda_graph, source_type=_lib.SRC_FZ if st == "fz" else _lib.SRC_EXPLOSIVE, boundary=boundary,
              sponge_factor=sponge_factor, storage=storage_code(storage))
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


def scalar2d(vp, dh, dt, nt, f0, *, sx, sz, rx, rz, nbc=50, fd_order=8, v_ref=None, wavelet=None,
             source_scale="v2dt2", verbose=True, use_cuda_graph=True, storage="f32"):
    """scalar2d(vp, dh, dt, nt, f0; sx, sz, rx, rz, ...) -> (seis_u, stats).

    src/equations/scalar2d/scalar2d.jl:113-189: second-order constant-density equation with the
    deterministic HABC.  source_scale :v2dt2 (default) multiplies each source's wavelet by
    vp(source)^2 dt^2, :dt by dt, :none leaves it (scalar2d.jl:139-146).  Extensions as acoustic2d:
    `use_cuda_graph`, `storage`."""
    del verbose
    scale = _sym(source_scale)
    if scale not in ("v2dt2", "dt", "none"):
        raise ArgumentError(f"source_scale must be :v2dt2, :dt or :none, got {source_scale}")
    if nbc < 0 or nt < 0:
        raise ArgumentError("nbc and nt must be >= 0")
    if fd_order not in (2, 4, 6, 8, 10):
        raise ValueError(f"Unsupported FD order: {fd_order} (centered d2)")
    vp = _f32F(vp)
    if vp.ndim != 2:
        raise ArgumentError("vp must be a matrix")
    nx, nz = vp.shape
    check_geometry(nx, nz, sx, sz, rx, rz)
    vmin = _default_vref(vp)
    check_numerics_centered(float(vp.max()), vmin, dh, dt, f0, fd_order)
    v_ref = vmin if v_ref is None else float(np.float32(v_ref))
    sx, sz, rx, rz = _i32(sx), _i32(sz), _i32(rx), _i32(rz)
    if len(sx) != len(sz) or len(rx) != len(rz):
        raise ArgumentError("source / receiver coordinate vectors differ in length")
    if wavelet is None:
        wav = ricker_wavelet(f0, dt, nt)
    else:
        wav = np.asarray(wavelet, np.float32).ravel()
        if len(wav) != nt:
            raise ArgumentError(f"wavelet length {len(wav)} != nt={nt}")
    wm = np.repeat(wav[None, :nt], len(sx), axis=0).astype(np.float32)
    f32 = np.float32
    if scale == "v2dt2":
        for q in range(len(sx)):
            v = f32(vp[sx[q] - 1, sz[q] - 1])
            wm[q] = wm[q] * f32(f32(v * v) * f32(f32(dt) * f32(dt)))
    elif scale == "dt":
        wm = (wm * f32(dt)).astype(np.float32)
    wm = np.asfortranarray(wm)
    n_rec = len(rx)
    d = _desc(_lib.SCALAR2D, nx, nz, nbc, fd_order, nt, dh, dt, v_ref, len(sx), n_rec, 0, use_cuda_graph,
              storage=storage_code(storage))
    su = np.zeros((n_rec, nt), np.float32, order="F")
    t = C.c_double(0.0)
    _lib.check(_lib.lib().fomo_scalar2d(C.byref(d), vp.ctypes.data, sx.ctypes.data, sz.ctypes.data, wm.ctypes.data,
                                        rx.ctypes.data, rz.ctypes.data, su.ctypes.data, C.byref(t)))
    return Scalar2DResult(su, Stats(float(t.value)))


def coupled2d(vp, vs, dh, dt, nt, f0, *, sx, sz, rx, rz, nbc=50, fd_order=8, snap_interval=0, v_ref_p=None,
              v_ref_s=None, smooth_sigma=3.0, wavelet=None, verbose=True, use_cuda_graph=True):
    """coupled2d(vp, vs, dh, dt, nt, f0; sx, sz, rx, rz, ...) -> (seis_P, seis_S, snaps_P, snaps_S, stats).

    src/equations/coupled2d/coupled2d.jl:53-183: the coupled P-S potential equations (constant
    density) as a velocity-position split with a Higdon blend after each half-step.  v_ref_p /
    v_ref_s (defaults min(vp > 0) / min(vs > 0)) set the P and S fields' Higdon coefficients;
    smooth_sigma (grid points, 0 disables) smooths alpha = vp^2 and beta = vs^2 for the coupling
    terms.  The wavelet (Ricker(f0) by default) is added raw to P.  Extension: `use_cuda_graph`."""
    del verbose
    if nbc < 0 or nt < 0 or snap_interval < 0:
        raise ArgumentError("nbc, nt and snap_interval must be >= 0")
    if fd_order not in (2, 4, 6, 8, 10):
        raise ValueError(f"Unsupported FD order: {fd_order} (centered d1)")
    vp, vs = _f32F(vp), _f32F(vs)
    if vp.ndim != 2 or vp.shape != vs.shape:
        raise ArgumentError("vp and vs must be matrices of the same size")
    nx, nz = vp.shape
    check_geometry(nx, nz, sx, sz, rx, rz)
    vpmin = _default_vref(vp)
    vs_pos = vs[vs > np.float32(0)]
    vmin_disp = float(np.float32(vs_pos.min())) if vs_pos.size else vpmin
    check_numerics_centered(float(vp.max()), vmin_disp, dh, dt, f0, fd_order)
    v_ref_p = vpmin if v_ref_p is None else float(np.float32(v_ref_p))
    if v_ref_s is None:
        if vs_pos.size == 0:
            raise ArgumentError("vs has no positive entry (v_ref_s defaults to min(vs > 0))")
        v_ref_s = float(np.float32(vs_pos.min()))
    else:
        v_ref_s = float(np.float32(v_ref_s))
    sx, sz, rx, rz = _i32(sx), _i32(sz), _i32(rx), _i32(rz)
    if len(sx) != len(sz) or len(rx) != len(rz):
        raise ArgumentError("source / receiver coordinate vectors differ in length")
    wm = wavelet_matrix(wavelet, f0, dt, nt, len(sx), False)
    n_rec = len(rx)
    d = _desc(_lib.COUPLED2D, nx, nz, nbc, fd_order, nt, dh, dt, v_ref_p, len(sx), n_rec, snap_interval,
              use_cuda_graph)
    d.v_ref_s = v_ref_s
    d.smooth_sigma = float(smooth_sigma)
    sP, sS = (np.zeros((n_rec, nt), np.float32, order="F") for _ in range(2))
    nsnap = nt // snap_interval if snap_interval > 0 else 0
    snP = np.zeros((nsnap, nz, nx), np.float32) if nsnap else None
    snS = np.zeros((nsnap, nz, nx), np.float32) if nsnap else None
    t = C.c_double(0.0)
    ptr = lambda a: None if a is None else a.ctypes.data  # noqa: E731
    _lib.check(_lib.lib().fomo_coupled2d(C.byref(d), vp.ctypes.data, vs.ctypes.data, sx.ctypes.data,
                                         sz.ctypes.data, wm.ctypes.data, rx.ctypes.data, rz.ctypes.data,
                                         sP.ctypes.data, sS.ctypes.data, ptr(snP), ptr(snS), C.byref(t)))
    return Coupled2DResult(sP, sS, [s.T for s in snP] if snP is not None else [],
                           [s.T for s in snS] if snS is not None else [], Stats(float(t.value)))


Acoustic2DBatchResult = namedtuple("Acoustic2DBatchResult", ["seis_p", "seis_vx", "seis_vz", "stats"])
Elastic2DBatchResult = namedtuple("Elastic2DBatchResult", ["seis_vx", "seis_vz", "stats"])
BatchStats = namedtuple("BatchStats", ["kernel_time_s", "n_shots"])


def _shot_matrix(sx, sz):
    """batch.jl:194-198: a vector means one source per shot; a matrix is (n_shots, n_src)."""
    sxm = np.asarray(sx, dtype=np.int64)
    szm = np.asarray(sz, dtype=np.int64)
    if sxm.ndim == 1:
        sxm = sxm[:, None]
    if szm.ndim == 1:
        szm = szm[:, None]
    if sxm.shape != szm.shape:
        raise ArgumentError(f"sx and sz differ in size: {sxm.shape} vs {szm.shape}")
    if sxm.shape[0] < 1:
        raise ArgumentError("at least one shot is required")
    return sxm, szm


def _batch(eq, vp, vs, rho, dh, dt, nt, f0, sx, sz, rx, rz, nbc, fd_order, v_ref, wavelet,
           scale_source_by_dt, use_cuda_graph, storage="f32"):
    vp, rho = _f32F(vp), _f32F(rho)
    if vs is not None:
        vs = _f32F(vs)
    nx, nz = vp.shape
    if nbc < 0 or nt < 0:
        raise ArgumentError("nbc and nt must be >= 0")
    if fd_order not in (2, 4, 6, 8, 10):
        raise ValueError(f"Unsupported FD order: {fd_order}")
    sxm, szm = _shot_matrix(sx, sz)
    ns, n_src = sxm.shape
    for s in range(ns):
        check_geometry(nx, nz, sxm[s], szm[s], rx, rz)
    vmin = _default_vref(vp)
    check_numerics(float(vp.max()), vmin, dh, dt, f0, fd_order)
    v_ref = vmin if v_ref is None else float(np.float32(v_ref))
    wm = wavelet_matrix(wavelet, f0, dt, nt, n_src, scale_source_by_dt)
    rx, rz = _i32(rx), _i32(rz)
    n_rec = len(rx)
    # device layout (n_src, n_shots) column-major (batch_acq.jl:27-37) = (n_shots, n_src) row-major
    sxd = np.ascontiguousarray(sxm.astype(np.int32))
    szd = np.ascontiguousarray(szm.astype(np.int32))
    d = _desc(eq, nx, nz, nbc, fd_order, nt, dh, dt, v_ref, n_src, n_rec, 0, use_cuda_graph,
              storage=storage_code(storage))
    nslots = 3 if eq == _lib.ACOUSTIC2D else 2
    outs = [np.zeros((n_rec, nt, ns), np.float32, order="F") for _ in range(nslots)]
    t = C.c_double(0.0)
    if eq == _lib.ACOUSTIC2D:
        _lib.check(_lib.lib().fomo_acoustic2d_batch(C.byref(d), vp.ctypes.data, rho.ctypes.data, ns, sxd.ctypes.data,
                                                    szd.ctypes.data, wm.ctypes.data, rx.ctypes.data, rz.ctypes.data,
                                                    outs[0].ctypes.data, outs[1].ctypes.data, outs[2].ctypes.data,
                                                    C.byref(t)))
    else:
        _lib.check(_lib.lib().fomo_elastic2d_batch(C.byref(d), vp.ctypes.data, vs.ctypes.data, rho.ctypes.data, ns,
                                                   sxd.ctypes.data, szd.ctypes.data, wm.ctypes.data, rx.ctypes.data,
                                                   rz.ctypes.data, outs[0].ctypes.data, outs[1].ctypes.data,
                                                   C.byref(t)))
    return outs, BatchStats(float(t.value), ns)


def acoustic2d_batch(vp, rho, dh, dt, nt, f0, *, sx, sz, rx, rz, nbc=50, fd_order=8, v_ref=None, wavelet=None,
                     scale_source_by_dt=False, verbose=True, use_cuda_graph=True, storage="f32"):
    """acoustic2d_batch (src/equations/acoustic2d/batch.jl:180-261) -> (seis_p, seis_vx, seis_vz, stats).

    Seismograms are (n_rec, nt, n_shots); each shot is bit-identical to its solo acoustic2d run."""
    del verbose
    outs, st = _batch(_lib.ACOUSTIC2D, vp, None, rho, dh, dt, nt, f0, sx, sz, rx, rz, nbc, fd_order, v_ref, wavelet,
                      scale_source_by_dt, use_cuda_graph, storage)
    return Acoustic2DBatchResult(outs[0], outs[1], outs[2], st)


def elastic2d_batch(vp, vs, rho, dh, dt, nt, f0, *, sx, sz, rx, rz, nbc=50, fd_order=8, v_ref=None, wavelet=None,
                    scale_source_by_dt=False, verbose=True, use_cuda_graph=True, storage="f32"):
    """elastic2d_batch (src/equations/elastic2d/batch.jl:197-261) -> (seis_vx, seis_vz, stats)."""
    del verbose
    outs, st = _batch(_lib.ELASTIC2D, vp, vs, rho, dh, dt, nt, f0, sx, sz, rx, rz, nbc, fd_order, v_ref, wavelet,
                      scale_source_by_dt, use_cuda_graph, storage)
    return Elastic2DBatchResult(outs[0], outs[1], st)


Acoustic3DResult = namedtuple("Acoustic3DResult", ["seis_vx", "seis_vy", "seis_vz", "snaps"])
Acoustic3DExtended = namedtuple("Acoustic3DExtended", ["seis_vx", "seis_vy", "seis_vz", "snaps", "seis_p", "stats"])
_PLANES = {"xz": 0, "xy": 1, "yz": 2}
_STORAGE = {"f32": _lib.STORE_F32, "float32": _lib.STORE_F32, "f16": _lib.STORE_F16, "float16": _lib.STORE_F16,
            "bf16": _lib.STORE_BF16, "bfloat16": _lib.STORE_BF16, "mat16": _lib.STORE_F32_MAT16,
            "f32_mat16": _lib.STORE_F32_MAT16}


def storage_code(storage) -> int:
    if isinstance(storage, (int, np.integer)):
        if int(storage) not in _STORAGE.values():
            raise ArgumentError(f"unknown storage code {storage}")
        return int(storage)
    key = _sym(storage).lower()
    if key not in _STORAGE:
        raise ArgumentError(f"storage must be :f32, :f16, :bf16 or :mat16, got {storage}")
    return _STORAGE[key]


def _run3d(eq, vp, vs, rho, dh, dt, nt, f0, sx, sy, sz, rx, ry, rz, nbc, fd_order, snap_interval, snap_plane,
           snap_index, v_ref, storage, wavelet, want_p, use_cuda_graph):
    vp, rho = _f32F(vp), _f32F(rho)
    if vp.ndim != 3 or vp.shape != rho.shape:
        raise ArgumentError("vp and rho must be 3D arrays of the same size")
    if vs is not None:
        vs = _f32F(vs)
        if vs.shape != vp.shape:
            raise ArgumentError("vp, vs and rho must be 3D arrays of the same size")
    if fd_order not in (2, 4, 6, 8, 10):
        raise ValueError(f"Unsupported FD order: {fd_order}")
    if nbc < 0 or nt < 0 or snap_interval < 0:
        raise ArgumentError("nbc, nt and snap_interval must be >= 0")
    nx, ny, nz = vp.shape
    plane = _PLANES.get(_sym(snap_plane))
    if plane is None:
        raise ArgumentError(f"snap_plane must be :xz, :xy or :yz, got {snap_plane}")
    if snap_index == 0:
        snap_index = ny // 2  # acoustic3d.jl:53-55, elastic3d.jl:49-51 (ny/2 for every plane)
    from .checks import check_geometry_3d
    check_geometry_3d(nx, ny, nz, sx, sy, sz, rx, ry, rz)
    v_ref = _default_vref(vp) if v_ref is None else float(np.float32(v_ref))
    sx, sy, sz, rx, ry, rz = (_i32(x) for x in (sx, sy, sz, rx, ry, rz))
    w = ricker_wavelet(f0, dt, nt) if wavelet is None else np.asarray(wavelet, np.float64).astype(np.float32)
    if len(w) != nt:
        raise ArgumentError(f"wavelet length {len(w)} != nt={nt}")
    wm = np.asfortranarray(np.repeat(w[None, :], len(sx), axis=0))
    d = _desc(eq, nx, nz, nbc, fd_order, nt, dh, dt, v_ref, len(sx), len(rx), snap_interval,
              use_cuda_graph, ny=ny, storage=storage_code(storage))
    d.snap_plane, d.snap_index = plane, snap_index
    n_rec = len(rx)
    outs = [np.zeros((n_rec, nt), np.float32, order="F") for _ in range(4)]
    nsnap = nt // snap_interval if snap_interval > 0 else 0
    a_n, b_n = {0: (nx, nz), 1: (nx, ny), 2: (ny, nz)}[plane]
    snaps = np.zeros((nsnap, b_n, a_n), np.float32) if nsnap else None
    t = C.c_double(0.0)
    common = (sx.ctypes.data, sy.ctypes.data, sz.ctypes.data, wm.ctypes.data, rx.ctypes.data, ry.ctypes.data,
              rz.ctypes.data, outs[0].ctypes.data, outs[1].ctypes.data, outs[2].ctypes.data)
    tail = (None if snaps is None else snaps.ctypes.data, C.byref(t))
    if eq == _lib.ACOUSTIC3D:
        _lib.check(_lib.lib().fomo_acoustic3d(C.byref(d), vp.ctypes.data, rho.ctypes.data, *common,
                                              outs[3].ctypes.data if want_p else None, *tail))
    else:
        _lib.check(_lib.lib().fomo_elastic3d(C.byref(d), vp.ctypes.data, vs.ctypes.data, rho.ctypes.data, *common,
                                             *tail))
    snap_list = [s.T for s in snaps] if snaps is not None else []
    return outs, snap_list, float(t.value)


def acoustic3d(vp, rho, dh, dt, nt, f0, *, sx, sy, sz, rx, ry, rz, nbc=50, fd_order=8, snap_interval=0,
               snap_plane="xz", snap_index=0, v_ref=None, storage="f32", wavelet=None, record_p=False,
               use_cuda_graph=True, extended=False):
    """acoustic3d (src/equations/acoustic3d/acoustic3d.jl:31-129) -> (seis_vx, seis_vy, seis_vz, snaps).

    Extensions: `storage` (:f32, :f16, :bf16 storage of wavefields and materials, FP32 arithmetic),
    `wavelet` (custom source time function), `record_p` and `extended=True` (also return seis_p and
    stats).  Every source gets the wavelet (the reference reshapes one wavelet to 1 x nt)."""
    outs, snap_list, t = _run3d(_lib.ACOUSTIC3D, vp, None, rho, dh, dt, nt, f0, sx, sy, sz, rx, ry, rz, nbc,
                                fd_order, snap_interval, snap_plane, snap_index, v_ref, storage, wavelet,
                                record_p or extended, use_cuda_graph)
    if extended:
        return Acoustic3DExtended(outs[0], outs[1], outs[2], snap_list, outs[3], Stats(t))
    return Acoustic3DResult(outs[0], outs[1], outs[2], snap_list)


Elastic3DResult = namedtuple("Elastic3DResult", ["seis_vx", "seis_vy", "seis_vz", "snaps"])
Elastic3DExtended = namedtuple("Elastic3DExtended", ["seis_vx", "seis_vy", "seis_vz", "snaps", "stats"])


def elastic3d(vp, vs, rho, dh, dt, nt, f0, *, sx, sy, sz, rx, ry, rz, nbc=50, fd_order=8, snap_interval=0,
              snap_plane="xz", snap_index=0, v_ref=None, storage="f32", wavelet=None, use_cuda_graph=True,
              extended=False):
    """elastic3d (src/equations/elastic3d/elastic3d.jl:31-117) -> (seis_vx, seis_vy, seis_vz, snaps).

    Explosive source (the wavelet is added to txx, tyy and tzz after the stress HABC), snapshots of
    vz.  Extensions as acoustic3d: `storage`, `wavelet`, `extended=True` (adds stats)."""
    if vs is None:
        raise ArgumentError("elastic3d needs vs")
    outs, snap_list, t = _run3d(_lib.ELASTIC3D, vp, vs, rho, dh, dt, nt, f0, sx, sy, sz, rx, ry, rz, nbc,
                                fd_order, snap_interval, snap_plane, snap_index, v_ref, storage, wavelet, False,
                                use_cuda_graph)
    if extended:
        return Elastic3DExtended(outs[0], outs[1], outs[2], snap_list, Stats(t))
    return Elastic3DResult(outs[0], outs[1], outs[2], snap_list)


class Plan:
    """Device-resident run (fomo_plan): setup once, then advance / read / reset.

    A solo run is a plan of one shot; `Plan.windows` builds a multi-shot plan whose shots are
    windows of one model (each bit-identical to its solo run on the cropped model)."""

    def __init__(self, eq, vp, rho, dh, dt, nt, *, sx=None, sz=None, rx=None, rz=None, wavelet=None, vs=None,
                 nbc=50, fd_order=8, v_ref=None, source_type=0, use_graph=True, sy=None, ry=None, storage=0,
                 record_p=False, tile_cfg=0, shots=None, snap_interval=0, v_ref_s=None, smooth_sigma=3.0):
        vp = _f32F(vp)
        rho = None if rho is None else _f32F(rho)  # scalar2d and coupled2d take no density
        self._keep = [vp, rho]
        if vs is not None:
            vs = _f32F(vs)
            self._keep.append(vs)
        shape = vp.shape
        nx, nz = shape[0], shape[-1]
        ny = shape[1] if len(shape) == 3 else 1
        self.nt, self.eq = nt, eq
        ptr = lambda a: None if a is None else a.ctypes.data  # noqa: E731
        h = C.c_void_p()
        if shots is None:
            v_ref = _default_vref(vp) if v_ref is None else float(np.float32(v_ref))
            sx, sz, rx, rz = _i32(sx), _i32(sz), _i32(rx), _i32(rz)
            sy = _i32(sy) if sy is not None else None
            ry = _i32(ry) if ry is not None else None
            wm = np.asfortranarray(np.asarray(wavelet, np.float32).reshape(len(sx), nt))
            self._keep += [sx, sz, rx, rz, sy, ry, wm]
            self.n_rec = [len(rx)]
            d = _desc(eq, nx, nz, nbc, fd_order, nt, dh, dt, v_ref, len(sx), len(rx), snap_interval, use_graph,
                      source_type=source_type, ny=ny, storage=storage_code(storage))
            d.record_p = 1 if record_p else 0
            d.tile_cfg = tile_cfg
            d.v_ref_s = 0.0 if v_ref_s is None else float(np.float32(v_ref_s))
            d.smooth_sigma = float(smooth_sigma)
            _lib.check(_lib.lib().fomo_plan_create(C.byref(d), vp.ctypes.data, ptr(vs), ptr(rho),
                                                   sx.ctypes.data, ptr(sy), sz.ctypes.data, wm.ctypes.data,
                                                   rx.ctypes.data, ptr(ry), rz.ctypes.data, C.byref(h)))
        else:
            table = (_lib.Shot * len(shots))()
            self.n_rec = []
            for k, s in enumerate(shots):
                ssx, ssz = _i32(s["sx"]), _i32(s["sz"])
                srx, srz = _i32(s["rx"]), _i32(s["rz"])
                w = np.asfortranarray(np.asarray(s["wavelet"], np.float32).reshape(len(ssx), nt))
                self._keep += [ssx, ssz, srx, srz, w]
                t = table[k]
                t.x0, t.nx, t.n_src, t.n_rec = int(s["x0"]), int(s["nx"]), len(ssx), len(srx)
                t.sx, t.sz, t.wavelet = ssx.ctypes.data, ssz.ctypes.data, w.ctypes.data
                t.rx, t.rz = srx.ctypes.data, srz.ctypes.data
                t.v_ref = float(s.get("v_ref") or 0.0)
                self.n_rec.append(len(srx))
            d = _desc(eq, nx, nz, nbc, fd_order, nt, dh, dt, 0.0, 0, 0, 0, use_graph, source_type=source_type,
                      storage=storage_code(storage))
            d.tile_cfg = tile_cfg
            d.v_ref_s = 0.0 if v_ref_s is None else float(np.float32(v_ref_s))
            d.smooth_sigma = float(smooth_sigma)
            _lib.check(_lib.lib().fomo_batch_create(C.byref(d), vp.ctypes.data, ptr(vs), ptr(rho),
                                                    len(shots), table, C.byref(h)))
        self._h = h
        n = C.c_int32()
        _lib.check(_lib.lib().fomo_plan_n_shots(h, C.byref(n)))
        self.n_shots = n.value
        self.padded, self.pads = [], []
        for s in range(self.n_shots):
            a, b, c, p = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
            _lib.check(_lib.lib().fomo_plan_shot_dims(h, s, C.byref(a), C.byref(b), C.byref(c), C.byref(p)))
            self.padded.append((a.value, c.value) if ny == 1 else (a.value, b.value, c.value))
            self.pads.append(p.value)
        self.pad = self.pads[0]

    @property
    def padded_cells(self) -> int:
        return int(sum(np.prod(p) for p in self.padded))

    def reset(self):
        _lib.check(_lib.lib().fomo_plan_reset(self._h))

    def advance(self, n: int):
        _lib.check(_lib.lib().fomo_plan_advance(self._h, int(n)))

    def timed_advance(self, n: int) -> float:
        ms = C.c_double(0.0)
        _lib.check(_lib.lib().fomo_plan_timed_advance(self._h, int(n), C.byref(ms)))
        return ms.value

    def timed_advance_status(self, n: int) -> tuple[float, bool]:
        """timed_advance that reports a 16-bit storage overflow (FOMO_EOVERFLOW: some stored value
        saturated, so the wavefields are invalid) instead of raising: (ms, overflowed)."""
        ms = C.c_double(0.0)
        rc = _lib.lib().fomo_plan_timed_advance(self._h, int(n), C.byref(ms))
        if rc == _lib.FOMO_EOVERFLOW:
            return ms.value, True
        _lib.check(rc)
        return ms.value, False

    def sync(self):
        _lib.check(_lib.lib().fomo_plan_sync(self._h))

    def kernel_times(self, n: int) -> tuple[float, float, float]:
        """Average ms per launch of (velocity sweep, pressure/stress sweep) over n back-to-back
        launches each, and the rest of an ordinary step's time (gather, counter updates, gaps)
        over n steps (the state advances n steps)."""
        out = (C.c_double * 3)()
        _lib.check(_lib.lib().fomo_plan_kernel_times(self._h, int(n), out))
        return out[0], out[1], out[2]

    def snaps(self, shot: int = 0) -> np.ndarray:
        """All snapshot slots of a 2D shot, (count, nx * nz) Float32 interior cells, x fastest;
        slots of steps not yet taken since the last reset are zero."""
        shp = self.padded[shot]
        if len(shp) != 2:
            raise ArgumentError("Plan.snaps reads 2D plans; 3D slices come from acoustic3d / elastic3d")
        n = C.c_int32(0)
        _lib.check(_lib.lib().fomo_plan_read_snaps(self._h, shot, C.byref(n), None))
        cells = (shp[0] - 2 * self.pads[shot]) * (shp[1] - 2 * self.pads[shot])
        out = np.zeros((n.value, cells), np.float32)
        if n.value:
            _lib.check(_lib.lib().fomo_plan_read_snaps(self._h, shot, C.byref(n), out.ctypes.data))
        return out

    def tuning(self) -> dict:
        """The work layout in use (fomo_plan_tuning): family ("sweeps" or "fused"), shape index,
        extent (2D strip rows; 3D z-planes per segment) and source ("timed", "process", "file",
        "fixed")."""
        out = (C.c_int32 * 4)()
        _lib.check(_lib.lib().fomo_plan_tuning(self._h, out))
        return {"family": ("sweeps", "fused")[out[0]], "shape": out[1], "extent": out[2],
                "source": ("timed", "process", "file", "fixed")[out[3]]}

    def launches(self) -> list[dict]:
        """The kernel launches of one step: name, cells updated, algorithmic bytes per cell."""
        L = _lib.lib()
        n = C.c_int32(0)
        _lib.check(L.fomo_plan_launch_count(self._h, C.byref(n)))
        out = []
        for k in range(n.value):
            name, cells, bpc = C.c_char_p(), C.c_int64(0), C.c_int32(0)
            _lib.check(L.fomo_plan_launch_info(self._h, k, C.byref(name), C.byref(cells), C.byref(bpc)))
            out.append({"name": name.value.decode(), "cells": cells.value, "bytes_per_cell": bpc.value})
        return out

    def launch_times(self, n: int) -> list[float]:
        """Average ms per launch of each launch of a step, each repeated n times back to back."""
        k = len(self.launches())
        ms = (C.c_double * k)()
        _lib.check(_lib.lib().fomo_plan_launch_times(self._h, int(n), ms))
        return list(ms)

    def seis(self, nslots: int = 3, shot: int = 0):
        outs = [np.zeros((self.n_rec[shot], self.nt), np.float32, order="F") for _ in range(nslots)]
        ptrs = [o.ctypes.data for o in outs] + [None] * (4 - nslots)
        _lib.check(_lib.lib().fomo_plan_read_shot_seis(self._h, shot, *ptrs))
        return outs

    def trace_misfit(self, other=None, *, slot: int = 0, shot: int = 0, other_shot: int = 0, t0: int = 0,
                     t1=None):
        """FP64 sums over one trace slot, reduced on the GPU (fomo_trace_misfit): per receiver
        num = sum_t (a - b)^2 and den = sum_t b^2 over samples t0 <= t < t1, where a is this plan's
        shot and b the other plan's shot (b = 0 when other is None: num is then each trace's
        energy).  Returns (num, den) as float64 arrays of the shot's receiver count."""
        if not 0 <= shot < self.n_shots:
            raise _lib.ArgumentError("shot index out of range")
        n = self.n_rec[shot]
        num = np.zeros(n, np.float64)
        den = np.zeros(n, np.float64)
        t1 = self.nt if t1 is None else int(t1)
        if other is not None and not getattr(other, "_h", None):
            raise _lib.ArgumentError("the other plan is closed")
        ob = None if other is None else other._h
        _lib.check(_lib.lib().fomo_trace_misfit(self._h, int(shot), ob, int(other_shot), int(slot), int(t0), t1,
                                                num.ctypes.data, den.ctypes.data))
        return num, den

    def relative_residual(self, other, *, slot: int = 0, shot: int = 0, other_shot: int = 0):
        """||a - b|| / ||b|| over the whole record of one slot (the reference's reflection measure,
        test/verify_patch.jl:86-87), from the GPU's FP64 sums."""
        num, den = self.trace_misfit(other, slot=slot, shot=shot, other_shot=other_shot)
        return float(np.sqrt(num.sum() / den.sum())) if den.sum() > 0 else float("inf")

    def field(self, which: int, shot: int = 0) -> np.ndarray:
        out = np.zeros(self.padded[shot], np.float32, order="F")
        _lib.check(_lib.lib().fomo_plan_read_shot_field(self._h, shot, int(which), out.ctypes.data))
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
