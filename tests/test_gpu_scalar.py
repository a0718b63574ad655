"""GPU parity of scalar2d (src/equations/scalar2d/scalar2d.jl:113-189) against the CPU oracle.

FP32 is bit-identical to the oracle (seismograms and the final field), for every order, with and
without graphs; 16-bit storage is held to an error budget against the FP64 oracle.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def model(nx, nz):
    """The reference's own check model (test/verify_scalar.jl:90-92): 3000 m/s over 3600 m/s."""
    vp = np.full((nx, nz), 3000.0, np.float32)
    vp[:, nz // 2 - 1:] = 3600.0
    return vp


def scaled(vp, w, sx, sz, dt):
    """source_scale :v2dt2 (scalar2d.jl:140-143) in Float32, one row per source."""
    f32 = np.float32
    rows = []
    for x, z in zip(sx, sz):
        v = f32(vp[x - 1, z - 1])
        rows.append(w * f32(f32(v * v) * f32(f32(dt) * f32(dt))))
    return np.stack(rows).astype(np.float32)


@pytest.mark.parametrize("fd_order", [2, 4, 6, 8, 10])
def test_scalar2d_bitwise_vs_oracle(gpu, orc, fd_order):
    import fomo
    nx, nz, nt, nbc = 120, 100, 150, 30
    vp = model(nx, nz)
    sx, sz = [nx // 2, 30], [8, 40]
    rx = np.arange(10, nx - 9, 10)
    rz = np.full(len(rx), 6)
    dt, dh, f0 = 1e-3, 10.0, 15.0
    w = fomo.ricker_wavelet(f0, dt, nt)
    r = fomo.scalar2d(vp, dh, dt, nt, f0, sx=sx, sz=sz, rx=rx, rz=rz, nbc=nbc, fd_order=fd_order)
    ref = orc.scalar2d(vp, dh, dt, nt, sx=sx, sz=sz, rx=rx, rz=rz, wavelet=scaled(vp, w, sx, sz, dt), nbc=nbc,
                       fd_order=fd_order)
    assert r.seis_u.shape == (len(rx), nt)
    assert np.abs(r.seis_u).max() > 0
    assert np.array_equal(r.seis_u, ref["seis_u"]), np.abs(r.seis_u - ref["seis_u"]).max()
    assert r.stats.kernel_time_s > 0


@pytest.mark.parametrize("nt", [1, 37, 64])
def test_scalar2d_final_state_graph_and_stream(gpu, orc, nt):
    """The final field through a plan (three-buffer rotation read back at u^nt), with graphs and
    with direct launches, for step counts that end inside and at the end of a graph replay."""
    import fomo
    from fomo import _lib
    nx, nz, nbc = 90, 70, 12
    vp = model(nx, nz)
    sx, sz, rx, rz = [45], [20], np.arange(1, nx + 1, 4), np.full(len(range(1, nx + 1, 4)), 5)
    dt = 1e-3
    w = scaled(vp, fomo.ricker_wavelet(15.0, dt, nt), sx, sz, dt)
    ref = orc.scalar2d(vp, 10.0, dt, nt, sx=sx, sz=sz, rx=rx, rz=rz, wavelet=w, nbc=nbc, final_state=True)
    for graph in (True, False):
        p = fomo.Plan(_lib.SCALAR2D, vp, vp, 10.0, dt, nt, sx=sx, sz=sz, rx=rx, rz=rz, wavelet=w, nbc=nbc,
                      fd_order=8, use_graph=graph)
        try:
            p.advance(nt)
            u = p.field(0)
            (su,) = p.seis(1)
        finally:
            p.close()
        assert np.array_equal(u, ref["state"]["u"])
        assert np.array_equal(su, ref["seis_u"])


@pytest.mark.parametrize("source_scale", ["dt", "none"])
def test_scalar2d_source_scales_and_edges(gpu, orc, source_scale):
    """source_scale :dt / :none, nbc = 0 (pad = M), sources and receivers on the interior's corner
    cells, duplicate source cells."""
    import fomo
    nx, nz, nt = 64, 48, 90
    vp = model(nx, nz)
    dt = 1e-3
    sx, sz = [1, nx, 32, 32], [1, nz, 20, 20]
    rx = np.array([1, nx, 32, 1])
    rz = np.array([1, nz, 24, nz])
    w = fomo.ricker_wavelet(20.0, dt, nt)
    wm = np.stack([w] * 4)
    if source_scale == "dt":
        wm = (wm * np.float32(dt)).astype(np.float32)
    for nbc in (0, 6):
        r = fomo.scalar2d(vp, 10.0, dt, nt, 20.0, sx=sx, sz=sz, rx=rx, rz=rz, nbc=nbc, fd_order=8,
                          source_scale=source_scale)
        ref = orc.scalar2d(vp, 10.0, dt, nt, sx=sx, sz=sz, rx=rx, rz=rz, wavelet=wm, nbc=nbc, fd_order=8)
        assert np.array_equal(r.seis_u, ref["seis_u"])


@pytest.mark.parametrize("storage,tol", [("f16", 3e-2), ("bf16", 2e-1)])
def test_scalar2d_reduced_storage_error(gpu, orc, storage, tol):
    import fomo
    nx, nz, nt = 140, 110, 400
    vp = model(nx, nz)
    sx, sz = [70], [10]
    rx = np.arange(1, nx + 1, 2)
    rz = np.full(len(rx), 6)
    dt = 1e-3
    w = scaled(vp, fomo.ricker_wavelet(15.0, dt, nt), sx, sz, dt)
    r = fomo.scalar2d(vp, 10.0, dt, nt, 15.0, sx=sx, sz=sz, rx=rx, rz=rz, nbc=20, storage=storage)
    ref = orc.scalar2d(vp, 10.0, dt, nt, sx=sx, sz=sz, rx=rx, rz=rz, wavelet=w, nbc=20, precision="f64")
    got = r.seis_u.astype(np.float64)
    den = np.linalg.norm(ref["seis_u"], axis=1)
    ok = den > 1e-2 * den.max()
    e = np.linalg.norm(got - ref["seis_u"], axis=1)[ok] / den[ok]
    print(f"scalar2d {storage}: per-trace rel L2 vs FP64 median {np.median(e):.2e} max {e.max():.2e}", flush=True)
    assert np.all(np.isfinite(got))
    assert e.max() < tol
