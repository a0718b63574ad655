"""GPU acoustic3d and elastic3d: bit-identity with the FP32 oracle (seismograms and final
wavefields), graph == stream, tiling independence, snapshots, and the error of Float16 / BFloat16
storage against the FP64 oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def model3d(nx, ny, nz, seed):
    from fomo.models import layered_3d
    return layered_3d(nx, ny, nz, seed=seed, n_interfaces=3, jitter=3.0, n_blobs=2, sigma=6.0)


def geom3d(nx, ny, nz):
    sx, sy, sz = [nx // 2, nx // 3], [ny // 2, ny // 2 + 3], [6, nz // 2]
    rx = np.concatenate([np.arange(1, nx + 1, 4), [1, nx, nx]])
    ry = np.concatenate([np.full(len(np.arange(1, nx + 1, 4)), ny // 2), [1, ny, 1]])
    rz = np.concatenate([np.full(len(np.arange(1, nx + 1, 4)), 3), [1, nz, nz]])
    return sx, sy, sz, rx, ry, rz


def run_plan(vp, rho, nt, w, geo, nbc, order, storage=0, use_graph=True):
    import fomo._lib as L
    from fomo.api import Plan
    sx, sy, sz, rx, ry, rz = geo
    p = Plan(L.ACOUSTIC3D, vp, rho, 10.0, 1e-3, nt, sx=sx, sy=sy, sz=sz, rx=rx, ry=ry, rz=rz,
             wavelet=np.stack([w] * len(sx)), nbc=nbc, fd_order=order, storage=storage, record_p=True,
             use_graph=use_graph)
    p.advance(nt)
    p.sync()
    seis = p.seis(4)
    state = [p.field(k) for k in range(4)]
    p.close()
    return seis, state


@pytest.mark.parametrize("order", [2, 4, 8])
def test_acoustic3d_bitwise_vs_oracle(gpu, orc, order):
    import fomo
    nx, ny, nz, nt, nbc = 45, 38, 41, 70, 5
    vp, rho = model3d(nx, ny, nz, 3)
    geo = geom3d(nx, ny, nz)
    w = fomo.ricker_wavelet(15.0, 1e-3, nt)
    seis, state = run_plan(vp, rho, nt, w, geo, nbc, order)
    sx, sy, sz, rx, ry, rz = geo
    ref = orc.acoustic3d(vp, rho, 10.0, 1e-3, nt, sx=sx, sy=sy, sz=sz, rx=rx, ry=ry, rz=rz,
                         wavelet=np.stack([w, w]), nbc=nbc, fd_order=order, final_state=True)
    for got, key in zip(seis, ("seis_vx", "seis_vy", "seis_vz", "seis_p")):
        assert np.all(np.isfinite(got))
        assert np.array_equal(got, ref[key]), (key, np.abs(got - ref[key]).max())
    for got, key in zip(state, ("p", "vx", "vy", "vz")):
        assert np.array_equal(got, ref["state"][key]), key


def test_acoustic3d_graph_equals_stream_and_api(gpu, orc):
    import fomo
    nx, ny, nz, nt = 40, 33, 36, 51
    vp, rho = model3d(nx, ny, nz, 4)
    geo = geom3d(nx, ny, nz)
    w = fomo.ricker_wavelet(15.0, 1e-3, nt)
    a, sa = run_plan(vp, rho, nt, w, geo, 6, 8, use_graph=True)
    b, sb = run_plan(vp, rho, nt, w, geo, 6, 8, use_graph=False)
    c, sc = run_plan(vp, rho, nt, w, geo, 6, 8, use_graph=True)
    for x, y, z in zip(a + sa, b + sb, c + sc):
        assert np.array_equal(x, y) and np.array_equal(x, z)
    sx, sy, sz, rx, ry, rz = geo
    r = fomo.acoustic3d(vp, rho, 10.0, 1e-3, nt, 15.0, sx=sx, sy=sy, sz=sz, rx=rx, ry=ry, rz=rz, nbc=6,
                        fd_order=8)
    assert len(r) == 4
    for x, y in zip(r[:3], a[:3]):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("plane", ["xz", "xy", "yz"])
def test_acoustic3d_snapshots(gpu, orc, plane):
    import fomo
    nx, ny, nz, nt = 34, 30, 28, 40
    vp, rho = model3d(nx, ny, nz, 5)
    geo = geom3d(nx, ny, nz)
    sx, sy, sz, rx, ry, rz = geo
    r = fomo.acoustic3d(vp, rho, 10.0, 1e-3, nt, 15.0, sx=sx, sy=sy, sz=sz, rx=rx, ry=ry, rz=rz, nbc=4,
                        fd_order=4, snap_interval=20, snap_plane=plane, snap_index=7)
    assert len(r.snaps) == 2
    w = fomo.ricker_wavelet(15.0, 1e-3, nt)
    ref = orc.acoustic3d(vp, rho, 10.0, 1e-3, nt, sx=sx, sy=sy, sz=sz, rx=rx, ry=ry, rz=rz,
                         wavelet=np.stack([w, w]), nbc=4, fd_order=4, final_state=True)
    pad = 4 + 2
    p = ref["state"]["p"][pad:-pad, pad:-pad, pad:-pad]
    want = {"xz": p[:, 6, :], "xy": p[:, :, 6], "yz": p[6, :, :]}[plane]
    assert r.snaps[-1].shape == want.shape
    assert np.array_equal(r.snaps[-1], want)


@pytest.mark.parametrize("si", [1, 3, 16])
@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_acoustic3d_snapshots_inside_graph_replays(gpu, si, storage):
    """Snapshots are written by the per-step gather launch, also inside 16-step graph replays:
    every snapshot k equals the p a plan holds after (k+1)*si steps, and graph == stream."""
    import fomo
    import fomo._lib as L
    from fomo.api import Plan
    nx, ny, nz, nt = 30, 26, 24, 37
    vp, rho = model3d(nx, ny, nz, 11)
    sx, sy, sz, rx, ry, rz = geom3d(nx, ny, nz)
    kw = dict(sx=sx, sy=sy, sz=sz, rx=rx, ry=ry, rz=rz, nbc=5, fd_order=8, snap_interval=si, snap_plane="xy",
              snap_index=9, storage=storage)
    a = fomo.acoustic3d(vp, rho, 10.0, 1e-3, nt, 15.0, **kw)
    b = fomo.acoustic3d(vp, rho, 10.0, 1e-3, nt, 15.0, use_cuda_graph=False, **kw)
    assert len(a.snaps) == nt // si
    for x, y in zip(a.snaps, b.snaps):
        assert np.array_equal(x, y)
    w = fomo.ricker_wavelet(15.0, 1e-3, nt)
    p = Plan(L.ACOUSTIC3D, vp, rho, 10.0, 1e-3, nt, sx=sx, sy=sy, sz=sz, rx=rx, ry=ry, rz=rz,
             wavelet=np.stack([w] * len(sx)), nbc=5, fd_order=8, storage=storage)
    try:
        done = 0
        pad = p.pad
        for k in sorted({0, len(a.snaps) // 2, len(a.snaps) - 1}):
            p.advance((k + 1) * si - done)
            done = (k + 1) * si
            f = p.field(0)[pad:-pad, pad:-pad, pad:-pad]
            assert np.array_equal(a.snaps[k], f[:, :, 8]), k
        assert np.abs(a.snaps[-1]).max() > 0
    finally:
        p.close()


@pytest.mark.parametrize("storage,tol", [(1, 2e-2), (2, 1e-1)])
def test_reduced_precision_storage_error(gpu, orc, storage, tol):
    """Float16 / BFloat16 storage with per-array scaling, FP32 arithmetic: per-trace relative L2
    against the FP64 oracle (reported; bound set loosely, the figure is what matters)."""
    import fomo
    nx, ny, nz, nt, nbc = 48, 44, 46, 200, 8
    vp, rho = model3d(nx, ny, nz, 6)
    geo = geom3d(nx, ny, nz)
    w = fomo.ricker_wavelet(15.0, 1e-3, nt)
    seis, _ = run_plan(vp, rho, nt, w, geo, nbc, 8, storage=storage)
    sx, sy, sz, rx, ry, rz = geo
    ref = orc.acoustic3d(vp, rho, 10.0, 1e-3, nt, sx=sx, sy=sy, sz=sz, rx=rx, ry=ry, rz=rz,
                         wavelet=np.stack([w, w]), nbc=nbc, fd_order=8, precision="f64")
    f32, _ = run_plan(vp, rho, nt, w, geo, nbc, 8, storage=0)
    errs = {}
    for got, g32, key in zip(seis, f32, ("seis_vx", "seis_vy", "seis_vz", "seis_p")):
        r = ref[key]
        den = np.linalg.norm(r, axis=1)
        ok = den > 1e-6 * den.max()
        e = np.linalg.norm(got.astype(np.float64) - r, axis=1)[ok] / den[ok]
        e32 = np.linalg.norm(g32.astype(np.float64) - r, axis=1)[ok] / den[ok]
        errs[key] = (float(np.median(e)), float(e.max()), float(e32.max()))
    print(f"storage {storage}: per-trace rel L2 vs FP64 (median, max, fp32 max) {errs}")
    for key, (med, mx, m32) in errs.items():
        assert m32 < 1e-4
        assert mx < tol, (key, mx)


# ---------------------------------------------------------------------------------------------
# elastic3d

E3_FIELDS = ("vx", "vy", "vz", "txx", "tyy", "tzz", "txy", "txz", "tyz")


def run_plan_el(vp, vs, rho, nt, w, geo, nbc, order, storage=0, use_graph=True, tile_cfg=0):
    import fomo._lib as L
    from fomo.api import Plan
    sx, sy, sz, rx, ry, rz = geo
    p = Plan(L.ELASTIC3D, vp, rho, 10.0, 1e-3, nt, vs=vs, sx=sx, sy=sy, sz=sz, rx=rx, ry=ry, rz=rz,
             wavelet=np.stack([w] * len(sx)), nbc=nbc, fd_order=order, storage=storage, use_graph=use_graph,
             tile_cfg=tile_cfg)
    p.advance(nt)
    p.sync()
    seis = p.seis(3)
    state = [p.field(k) for k in range(9)]
    p.close()
    return seis, state


def el_model(nx, ny, nz, seed):
    vp, rho = model3d(nx, ny, nz, seed)
    vs = (vp / 1.75).astype(np.float32)
    return vp, vs, rho


@pytest.mark.parametrize("order", [2, 4, 8])
def test_elastic3d_bitwise_vs_oracle(gpu, orc, order):
    import fomo
    nx, ny, nz, nt, nbc = 37, 34, 33, 50, 5
    vp, vs, rho = el_model(nx, ny, nz, 7)
    geo = geom3d(nx, ny, nz)
    w = fomo.ricker_wavelet(15.0, 1e-3, nt)
    seis, state = run_plan_el(vp, vs, rho, nt, w, geo, nbc, order)
    sx, sy, sz, rx, ry, rz = geo
    ref = orc.elastic3d(vp, vs, rho, 10.0, 1e-3, nt, sx=sx, sy=sy, sz=sz, rx=rx, ry=ry, rz=rz,
                        wavelet=np.stack([w, w]), nbc=nbc, fd_order=order, final_state=True)
    for got, key in zip(seis, ("seis_vx", "seis_vy", "seis_vz")):
        assert np.all(np.isfinite(got)) and np.abs(got).max() > 0
        assert np.array_equal(got, ref[key]), (key, np.abs(got - ref[key]).max())
    for got, key in zip(state, E3_FIELDS):
        assert np.array_equal(got, ref["state"][key]), key


def test_elastic3d_graph_stream_tiling_and_api(gpu, orc):
    """Graph replay == direct launches == other z-segment lengths (16 and 128 planes), bit for
    bit; the one-shot API returns the same seismograms."""
    import fomo
    nx, ny, nz, nt = 40, 31, 44, 41
    vp, vs, rho = el_model(nx, ny, nz, 8)
    geo = geom3d(nx, ny, nz)
    w = fomo.ricker_wavelet(15.0, 1e-3, nt)
    runs = [run_plan_el(vp, vs, rho, nt, w, geo, 7, 8, use_graph=g, tile_cfg=c)
            for g, c in ((True, 0), (False, 0), (True, 1), (True, 4))]
    a, sa = runs[0]
    for b, sb in runs[1:]:
        for x, y in zip(a + sa, b + sb):
            assert np.array_equal(x, y)
    sx, sy, sz, rx, ry, rz = geo
    r = fomo.elastic3d(vp, vs, rho, 10.0, 1e-3, nt, 15.0, sx=sx, sy=sy, sz=sz, rx=rx, ry=ry, rz=rz, nbc=7,
                       fd_order=8, snap_interval=20, snap_plane="xy", snap_index=9)
    assert len(r) == 4 and len(r.snaps) == 2
    for x, y in zip(r[:3], a):
        assert np.array_equal(x, y)
    # snapshot 2 (step 40) is vz at interior z = 9; compare with a 40-step plan's final vz
    s40, st40 = run_plan_el(vp, vs, rho, 40, w[:40], geo, 7, 8)
    pad = 7 + 4
    vz = st40[2][pad:-pad, pad:-pad, pad:-pad]
    assert np.array_equal(r.snaps[-1], vz[:, :, 8])


def test_acoustic3d_tiling_independence(gpu):
    """z-segments of 16, 64 and 128 planes give the same bits."""
    import fomo
    nx, ny, nz, nt = 36, 35, 70, 30
    vp, rho = model3d(nx, ny, nz, 9)
    geo = geom3d(nx, ny, nz)
    w = fomo.ricker_wavelet(15.0, 1e-3, nt)
    import fomo._lib as L
    from fomo.api import Plan
    sx, sy, sz, rx, ry, rz = geo
    outs = []
    for cfg in (1, 3, 4):
        p = Plan(L.ACOUSTIC3D, vp, rho, 10.0, 1e-3, nt, sx=sx, sy=sy, sz=sz, rx=rx, ry=ry, rz=rz,
                 wavelet=np.stack([w] * len(sx)), nbc=9, fd_order=8, record_p=True, tile_cfg=cfg)
        p.advance(nt)
        outs.append(p.seis(4) + [p.field(k) for k in range(4)])
        p.close()
    for other in outs[1:]:
        for x, y in zip(outs[0], other):
            assert np.array_equal(x, y)


@pytest.mark.parametrize("storage,tol,tol_vy", [(1, 5e-2, 5e-2), (2, 2e-1, 1.0)])
def test_elastic3d_reduced_precision_storage_error(gpu, orc, storage, tol, tol_vy):
    """vy on the receiver line is a small difference of large stress derivatives, so 16-bit
    storage costs it more than vx and vz: FP16 ~2 % global on vy, ~0.4 % on vx and vz; BFloat16
    ~12 % global on vy (too coarse for it), ~4 % on vx and vz."""
    import fomo
    nx, ny, nz, nt, nbc = 40, 38, 42, 160, 8
    vp, vs, rho = el_model(nx, ny, nz, 10)
    geo = geom3d(nx, ny, nz)
    w = fomo.ricker_wavelet(15.0, 1e-3, nt)
    seis, _ = run_plan_el(vp, vs, rho, nt, w, geo, nbc, 8, storage=storage)
    sx, sy, sz, rx, ry, rz = geo
    ref = orc.elastic3d(vp, vs, rho, 10.0, 1e-3, nt, sx=sx, sy=sy, sz=sz, rx=rx, ry=ry, rz=rz,
                        wavelet=np.stack([w, w]), nbc=nbc, fd_order=8, precision="f64")
    # Traces whose energy is below 1 % of the strongest trace of the component sit near the 16-bit
    # resolution of the field's single scale; they are reported, not gated.
    errs = {}
    for got, key in zip(seis, ("seis_vx", "seis_vy", "seis_vz")):
        r = ref[key]
        den = np.linalg.norm(r, axis=1)
        err = np.linalg.norm(got.astype(np.float64) - r, axis=1)
        strong = den > 1e-2 * den.max()
        weak = (den > 1e-6 * den.max()) & ~strong
        e = err[strong] / den[strong]
        glob = float(np.linalg.norm(got.astype(np.float64) - r) / np.linalg.norm(r))
        ew = float((err[weak] / den[weak]).max()) if weak.any() else 0.0
        errs[key] = (float(np.median(e)), float(e.max()), glob, ew)
    print(f"elastic3d storage {storage}: per-trace rel L2 vs FP64 over strong traces (median, max), "
          f"global rel L2, max over weak traces: {errs}")
    for key, (med, mx, glob, ew) in errs.items():
        t = tol_vy if key == "seis_vy" else tol
        assert mx < t and glob < t, (key, mx, glob)


@pytest.mark.parametrize("order,storage", [(2, "f32"), (4, "f32"), (8, "f32"), (8, "f16"), (10, "bf16")])
def test_acoustic3d_fused_interior_equals_two_sweeps(gpu, order, storage, monkeypatch):
    """The one-launch interior step (velocity then pressure, new velocities kept on chip) gives the
    bits of the two separate sweeps (FOMO_3D_UNFUSED; FOMO_3D_FUSED forces it on 16-bit storage,
    whose default is the separate sweeps), for every FD order's halo widths, for 16-bit
    storage, and for interior sizes that leave partial tiles in x (50 = 64 - 14), y (37 = 2 x 16 +
    5) and z (segments of 32 planes over 45); the launch list names the fused launch."""
    import fomo
    import fomo._lib as L
    from fomo.api import Plan
    nx, ny, nz, nt = 50, 37, 45, 24
    vp, rho = model3d(nx, ny, nz, 12)
    geo = geom3d(nx, ny, nz)
    w = fomo.ricker_wavelet(15.0, 1e-3, nt)
    sx, sy, sz, rx, ry, rz = geo
    outs = []
    for unfused in (False, True):
        if unfused:
            monkeypatch.setenv("FOMO_3D_UNFUSED", "1")
            monkeypatch.delenv("FOMO_3D_FUSED", raising=False)
        else:
            monkeypatch.delenv("FOMO_3D_UNFUSED", raising=False)
            monkeypatch.setenv("FOMO_3D_FUSED", "1")  # 16-bit plans run the separate sweeps by default
        p = Plan(L.ACOUSTIC3D, vp, rho, 10.0, 1e-3, nt, sx=sx, sy=sy, sz=sz, rx=rx, ry=ry, rz=rz,
                 wavelet=np.stack([w] * len(sx)), nbc=6, fd_order=order, storage=storage, record_p=True, tile_cfg=1)
        try:
            names = [x["name"] for x in p.launches()]
            assert any("interior step" in n for n in names) != unfused, names
            p.advance(nt)
            outs.append(p.seis(4) + [p.field(k) for k in range(4)])
        finally:
            p.close()
    for x, y in zip(outs[0], outs[1]):
        assert np.array_equal(x, y)
    assert np.abs(outs[0][4]).max() > 0


# ---------------------------------------------------------------------------------------------
# row layout


@pytest.mark.parametrize("nbc,nx", [(28, 40), (11, 49), (13, 52), (3, 62)])
def test_row_layouts_bitwise_vs_oracle(gpu, orc, nbc, nx):
    """3D arrays sit in rows of a 128-byte pitch whose base is offset so that the interior's first
    column starts a line.  Pads that put that column at different positions of a line (offsets 0,
    17, 15 and 25 floats), a grid whose offset rows fill their pitch exactly (nbc 11, nx 49) and
    small pads all give the oracle's bits in traces and final fields, through the fused interior
    and through the separate sweeps (FD8), and for elastic3d (FD4)."""
    import fomo
    ny, nz, nt = 30, 34, 40
    vp, rho = model3d(nx, ny, nz, 21)
    geo = geom3d(nx, ny, nz)
    w = fomo.ricker_wavelet(15.0, 1e-3, nt)
    sx, sy, sz, rx, ry, rz = geo
    print(f"row layout: nx {nx} nbc {nbc}", flush=True)
    ref = orc.acoustic3d(vp, rho, 10.0, 1e-3, nt, sx=sx, sy=sy, sz=sz, rx=rx, ry=ry, rz=rz,
                         wavelet=np.stack([w, w]), nbc=nbc, fd_order=8, final_state=True)
    seis, state = run_plan(vp, rho, nt, w, geo, nbc, 8)
    for got, key in zip(seis, ("seis_vx", "seis_vy", "seis_vz", "seis_p")):
        assert np.array_equal(got, ref[key]), key
    for got, key in zip(state, ("p", "vx", "vy", "vz")):
        assert np.array_equal(got, ref["state"][key]), key
    vs = (vp / 1.75).astype(np.float32)
    ref = orc.elastic3d(vp, vs, rho, 10.0, 1e-3, nt, sx=sx, sy=sy, sz=sz, rx=rx, ry=ry, rz=rz,
                        wavelet=np.stack([w, w]), nbc=nbc, fd_order=4, final_state=True)
    seis, state = run_plan_el(vp, vs, rho, nt, w, geo, nbc, 4)
    for got, key in zip(seis, ("seis_vx", "seis_vy", "seis_vz")):
        assert np.array_equal(got, ref[key]), key
    for got, key in zip(state, E3_FIELDS):
        assert np.array_equal(got, ref["state"][key]), key


@pytest.mark.parametrize("storage", [1, 2])
def test_row_layout_16bit_traces_equal_field_reads(gpu, storage):
    """With 16-bit storage the rows are padded to 64 elements (128 bytes) and the base offset
    differs from Float32's: the receiver gather (cells addressed through the pitch) and the
    decoded field read (which drops the pitch) return the same values at the receivers."""
    import fomo
    nx, ny, nz, nt, nbc = 45, 31, 33, 30, 13
    vp, rho = model3d(nx, ny, nz, 22)
    geo = geom3d(nx, ny, nz)
    w = fomo.ricker_wavelet(15.0, 1e-3, nt)
    seis, state = run_plan(vp, rho, nt, w, geo, nbc, 8, storage=storage)
    sx, sy, sz, rx, ry, rz = geo
    pad = nbc + 4
    for k, f in ((0, 1), (1, 2), (2, 3), (3, 0)):  # vx, vy, vz, p traces against fields vx, vy, vz, p
        at = state[f][rx - 1 + pad, ry - 1 + pad, rz - 1 + pad]
        assert np.abs(seis[k][:, -1]).max() > 0
        assert np.array_equal(seis[k][:, -1], at), k
