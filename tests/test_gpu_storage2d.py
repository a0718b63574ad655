"""Float16 / BFloat16 storage of the 2D solvers, and Float16 materials under Float32 wavefields
(extensions; arithmetic stays Float32).

Fields are stored as RN(v / dec) with dec = the shot's peak source amplitude for p and stresses and
that over the smallest impedance for velocities; materials with dec = max |value|.  The bar is an
error budget against the FP64 oracle (the reference has no such mode); Float32 storage stays
bit-identical, which the parity tests cover.  Every figure is printed so the run log records it.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NX, NZ, NT, NBC = 160, 110, 500, 16
DT, DH, F0 = 1e-3, 10.0, 15.0


def model():
    from fomo.models import layered_2d
    return layered_2d(NX, NZ, seed=5, n_interfaces=4, jitter=3.0, vacuum_top=True)


def geometry():
    sx, sz = [NX // 2], [6]
    rx = np.arange(1, NX + 1, 2)
    rz = np.full(len(rx), 4)
    return sx, sz, rx, rz


def strong_trace_errors(got, ref):
    """Per-trace relative L2 against `ref` over traces holding > 1 % of the strongest trace's
    energy, and the global relative L2 over all traces."""
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    den = np.linalg.norm(ref, axis=1)
    ok = den > 1e-2 * den.max()
    e = np.linalg.norm(got - ref, axis=1)[ok] / den[ok]
    glob = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    return float(np.median(e)), float(e.max()), float(glob)


@pytest.mark.parametrize("storage,tol", [("f16", 4e-2), ("bf16", 2e-1)])
def test_acoustic2d_storage_error(gpu, orc, storage, tol):
    import fomo
    vp, _, rho = model()
    sx, sz, rx, rz = geometry()
    w = fomo.ricker_wavelet(F0, DT, NT)
    r = fomo.acoustic2d(vp, rho, DH, DT, NT, F0, sx=sx, sz=sz, rx=rx, rz=rz, nbc=NBC, storage=storage)
    r32 = fomo.acoustic2d(vp, rho, DH, DT, NT, F0, sx=sx, sz=sz, rx=rx, rz=rz, nbc=NBC)
    ref = orc.acoustic2d(vp, rho, DH, DT, NT, sx=sx, sz=sz, rx=rx, rz=rz, wavelet=w[None], nbc=NBC,
                         precision="f64")
    for k in ("seis_p", "seis_vx", "seis_vz"):
        med, mx, glob = strong_trace_errors(getattr(r, k), ref[k])
        _, mx32, _ = strong_trace_errors(getattr(r32, k), ref[k])
        print(f"acoustic2d {storage} {k}: per-trace rel L2 vs FP64 median {med:.2e} max {mx:.2e} "
              f"global {glob:.2e} (fp32 max {mx32:.2e})", flush=True)
        assert np.all(np.isfinite(getattr(r, k)))
        assert mx32 < 1e-4
        assert mx < tol, (k, mx)


@pytest.mark.parametrize("source_type", ["explosive", "fz"])
@pytest.mark.parametrize("storage,tol", [("f16", 6e-2), ("bf16", 3e-1)])
def test_elastic2d_storage_error(gpu, orc, storage, tol, source_type):
    """Both source kinds: a vertical force feeds vz directly, so its scales put velocities in units
    of the source amplitude (stresses would overflow Float16 in the other convention)."""
    import fomo
    vp, vs, rho = model()
    sx, sz, rx, rz = geometry()
    w = fomo.ricker_wavelet(F0, DT, NT)
    kw = dict(sx=sx, sz=sz, rx=rx, rz=rz, nbc=NBC, source_type=source_type)
    r = fomo.elastic2d(vp, vs, rho, DH, DT, NT, F0, storage=storage, **kw)
    r32 = fomo.elastic2d(vp, vs, rho, DH, DT, NT, F0, **kw)
    ref = orc.elastic2d(vp, vs, rho, DH, DT, NT, sx=sx, sz=sz, rx=rx, rz=rz, wavelet=w[None], nbc=NBC,
                        precision="f64", source_type=1 if source_type == "fz" else 0)
    for k in ("seis_vx", "seis_vz"):
        med, mx, glob = strong_trace_errors(getattr(r, k), ref[k])
        _, mx32, _ = strong_trace_errors(getattr(r32, k), ref[k])
        print(f"elastic2d {storage} {source_type} {k}: per-trace rel L2 vs FP64 median {med:.2e} max {mx:.2e} "
              f"global {glob:.2e} (fp32 max {mx32:.2e})", flush=True)
        assert np.all(np.isfinite(getattr(r, k)))
        assert mx32 < 1e-4
        assert mx < tol, (k, mx)


@pytest.mark.parametrize("storage", ["f16", "bf16", "mat16"])
def test_batch_shot_equals_solo_reduced_storage(gpu, storage):
    """Scales are per shot and depend only on the shot's own wavelet and window, so a batched shot
    stays bit-identical to its solo run in 16-bit storage too."""
    import fomo
    vp, vs, rho = model()
    _, _, rx, rz = geometry()
    sxs, szs = [30, NX // 2, NX - 20], [6, 9, 5]
    b = fomo.elastic2d_batch(vp, vs, rho, DH, DT, 300, F0, sx=sxs, sz=szs, rx=rx, rz=rz, nbc=NBC,
                             storage=storage)
    for k in range(3):
        s = fomo.elastic2d(vp, vs, rho, DH, DT, 300, F0, sx=[sxs[k]], sz=[szs[k]], rx=rx, rz=rz, nbc=NBC,
                           storage=storage)
        assert np.array_equal(b.seis_vx[:, :, k], s.seis_vx)
        assert np.array_equal(b.seis_vz[:, :, k], s.seis_vz)


@pytest.mark.parametrize("storage", ["f16", "bf16", "mat16"])
def test_graph_stream_snapshots_reduced_storage(gpu, storage):
    """Graph and stream launches agree bitwise; snapshots are the decoded stored wavefield, equal
    to what a plan reads back after the same number of steps."""
    import fomo
    from fomo import _lib
    vp, _, rho = model()
    sx, sz, rx, rz = geometry()
    nt, si = 240, 60
    a = fomo.acoustic2d(vp, rho, DH, DT, nt, F0, sx=sx, sz=sz, rx=rx, rz=rz, nbc=NBC, storage=storage,
                        snap_interval=si)
    b = fomo.acoustic2d(vp, rho, DH, DT, nt, F0, sx=sx, sz=sz, rx=rx, rz=rz, nbc=NBC, storage=storage,
                        snap_interval=si, use_cuda_graph=False)
    for k in ("seis_p", "seis_vx", "seis_vz"):
        assert np.array_equal(getattr(a, k), getattr(b, k))
    assert len(a.snaps) == nt // si
    for x, y in zip(a.snaps, b.snaps):
        assert np.array_equal(x, y)
    w = fomo.ricker_wavelet(F0, DT, nt)
    p = fomo.Plan(_lib.ACOUSTIC2D, vp, rho, DH, DT, nt, sx=sx, sz=sz, rx=rx, rz=rz, wavelet=w[None], nbc=NBC,
                  storage=storage)
    try:
        p.advance(si)
        pad = p.pad
        f = p.field(0)[pad:-pad, pad:-pad]
        assert f.shape == a.snaps[0].shape
        assert np.array_equal(f, a.snaps[0])
        assert np.abs(f).max() > 0
        p.advance(nt - si)
        sp, _, _ = p.seis(3)
        assert np.array_equal(sp, a.seis_p)
    finally:
        p.close()


def test_learned_boundary_f16_close_to_f32(gpu):
    """The learned boundary on Float16 storage against the same network on Float32 storage."""
    import fomo
    vp, _, rho = model()
    sx, sz, rx, rz = geometry()
    r16 = fomo.acoustic2d(vp, rho, DH, DT, NT, F0, sx=sx, sz=sz, rx=rx, rz=rz, nbc=16, boundary="learned",
                          storage="f16")
    r32 = fomo.acoustic2d(vp, rho, DH, DT, NT, F0, sx=sx, sz=sz, rx=rx, rz=rz, nbc=16, boundary="learned")
    med, mx, glob = strong_trace_errors(r16.seis_p, r32.seis_p)
    print(f"learned f16 vs learned f32 seis_p: per-trace rel L2 median {med:.2e} max {mx:.2e} global {glob:.2e}",
          flush=True)
    assert mx < 4e-2



@pytest.mark.parametrize("eq", ["acoustic2d", "elastic2d/explosive", "elastic2d/fz"])
def test_mat16_storage_error(gpu, orc, eq):
    """Float32 wavefields with Float16 materials (storage :mat16): each material is rounded once to
    11 significant bits, a fixed perturbation of the model, while every step still rounds only in
    Float32.  Measured on this model, over the solvers, sources and trace slots: 0.97-5.1e-3
    per-trace max (medians 0.6-1.9e-3), 0.53-2.0e-3 global, about ten times closer to FP64 than
    Float16 wavefields.  The bars: 1e-2 per trace, 4e-3 global."""
    import fomo
    vp, vs, rho = model()
    sx, sz, rx, rz = geometry()
    w = fomo.ricker_wavelet(F0, DT, NT)
    kw = dict(sx=sx, sz=sz, rx=rx, rz=rz, nbc=NBC)
    if eq == "acoustic2d":
        r = fomo.acoustic2d(vp, rho, DH, DT, NT, F0, storage="mat16", **kw)
        ref = orc.acoustic2d(vp, rho, DH, DT, NT, wavelet=w[None], precision="f64", **kw)
        keys = ("seis_p", "seis_vx", "seis_vz")
    else:
        st = eq.split("/")[1]
        r = fomo.elastic2d(vp, vs, rho, DH, DT, NT, F0, storage="mat16", source_type=st, **kw)
        ref = orc.elastic2d(vp, vs, rho, DH, DT, NT, wavelet=w[None], precision="f64",
                            source_type=1 if st == "fz" else 0, **kw)
        keys = ("seis_vx", "seis_vz")
    for k in keys:
        med, mx, glob = strong_trace_errors(getattr(r, k), ref[k])
        print(f"{eq} mat16 {k}: per-trace rel L2 vs FP64 median {med:.2e} max {mx:.2e} global {glob:.2e}",
              flush=True)
        assert np.all(np.isfinite(getattr(r, k)))
        assert mx < 1e-2 and glob < 4e-3, (k, mx, glob)


def test_mat16_constant_model_is_bitwise_f32(gpu):
    """On a homogeneous model every material array is one value, which Float16 with its
    max |value| scale stores exactly (1.0 times the scale): :mat16 must then give Float32's bits,
    through the separate mixed-type kernels, for both equations."""
    import fomo
    vp = np.full((NX, NZ), 2500.0, np.float32)
    vs = np.full((NX, NZ), 1400.0, np.float32)
    rho = np.full((NX, NZ), 2100.0, np.float32)
    sx, sz, rx, rz = geometry()
    kw = dict(sx=sx, sz=[30], rx=rx, rz=rz, nbc=NBC)
    a = fomo.acoustic2d(vp, rho, DH, DT, 300, F0, storage="mat16", **kw)
    b = fomo.acoustic2d(vp, rho, DH, DT, 300, F0, **kw)
    for k in ("seis_p", "seis_vx", "seis_vz"):
        assert np.array_equal(getattr(a, k), getattr(b, k)), k
    a = fomo.elastic2d(vp, vs, rho, DH, DT, 300, F0, storage="mat16", **kw)
    b = fomo.elastic2d(vp, vs, rho, DH, DT, 300, F0, **kw)
    for k in ("seis_vx", "seis_vz"):
        assert np.array_equal(getattr(a, k), getattr(b, k)), k
    assert np.abs(b.seis_vz).max() > 0


def test_mat16_rejected_where_fields_and_materials_share_buffers(gpu):
    """scalar2d keeps u^{n-1} among its stencil inputs' "materials", and the 3D engine has no
    mixed kernels: both refuse :mat16."""
    import fomo
    sx, sz, rx, rz = geometry()
    vp = np.full((NX, NZ), 2500.0, np.float32)
    with pytest.raises(fomo.ArgumentError):
        fomo.scalar2d(vp, DH, DT, 50, F0, sx=sx, sz=sz, rx=rx, rz=rz, nbc=NBC, storage="mat16")
    n = 24
    v3 = np.full((n, n, n), 2500.0, np.float32)
    with pytest.raises(fomo.ArgumentError):
        fomo.acoustic3d(v3, np.full((n, n, n), 2000.0, np.float32), DH, DT, 20, F0, sx=[12], sy=[12], sz=[4],
                        rx=[5], ry=[5], rz=[3], nbc=8, storage="mat16")


def test_fp16_overflow_is_reported(gpu):
    """16-bit overflow is loud.  On the reference's absorption-experiment model, plain HABC at
    nbc 100 grows without bound in the bottom frame after about 700 steps (the oracle shows the
    same growth in FP32, where it stays finite for 1500 steps); with Float16 storage the growth
    leaves the format's range.  The overflowing stores saturate and set the plan's flag, and the
    next call that waits for the plan raises StorageOverflowError; a reset clears the flag, and
    the one-shot API raises the same error.  Float32 storage runs the same case without error."""
    import fomo
    import fomo._lib as L
    from fomo import learned
    from fomo.api import Plan
    vp, rho, sx, sz, rx, rz = learned.patch_geometry()
    nt = 1500
    w = fomo.ricker_wavelet(10.0, 1e-3, nt)
    p = Plan(L.ACOUSTIC2D, vp, rho, 10.0, 1e-3, nt, sx=[sx], sz=[sz], rx=rx, rz=rz, wavelet=w[None], nbc=100,
             fd_order=8, storage="f16")
    try:
        p.advance(300)
        p.sync()  # early steps stay in range
        p.advance(nt - 300)
        with pytest.raises(L.StorageOverflowError, match="overflow"):
            p.sync()
        with pytest.raises(L.StorageOverflowError):
            p.seis(3)
        p.reset()
        p.advance(300)
        p.sync()
    finally:
        p.close()
    with pytest.raises(L.StorageOverflowError):
        fomo.acoustic2d(vp, rho, 10.0, 1e-3, nt, 10.0, sx=[sx], sz=[sz], rx=rx, rz=rz, nbc=100, storage="f16")
    r = fomo.acoustic2d(vp, rho, 10.0, 1e-3, nt, 10.0, sx=[sx], sz=[sz], rx=rx, rz=rz, nbc=100)
    assert np.all(np.isfinite(r.seis_p))
