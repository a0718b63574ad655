"""elastic2d: the one-launch interior step (new vx, vz kept in LDS and registers on their way to the
new stresses; the frame's velocities and stresses in launches of their own) gives the bits of the
two separate sweeps (FOMO_2D_UNFUSED=1) and of the CPU oracle -- receivers, the five final fields
and snapshots -- for every FD order, both source types (explosive into txx/tzz, vertical force into
vz, which lanes recomputing a neighbouring strip's halo velocities inject too), 16-bit storage,
strips shorter than the halo and single-strip columns, a frame of nbc = 0, a batch of uneven
windows and the sponge boundary."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("vx", "vz", "txx", "tzz", "txz")


def _model(nx, nz, seed):
    from fomo.models import layered_2d
    vp, _, rho = layered_2d(nx, nz, seed=seed, n_interfaces=4, vacuum_top=False)
    vs = (vp / np.float32(1.73)).astype(np.float32)
    return vp, vs, rho


def _set(monkeypatch, unfused, rows):
    if unfused:
        monkeypatch.setenv("FOMO_2D_UNFUSED", "1")
        monkeypatch.delenv("FOMO_2D_FUSED", raising=False)
    else:
        monkeypatch.delenv("FOMO_2D_UNFUSED", raising=False)
        monkeypatch.setenv("FOMO_2D_FUSED", "1")
    if rows:
        monkeypatch.setenv("FOMO_FUSED2D_ROWS", str(rows))
    else:
        monkeypatch.delenv("FOMO_FUSED2D_ROWS", raising=False)


def _strip_width(nx, order):
    """Columns of the fused launch's strips over an interior of nx columns: at most 256 - 2N, the
    strips of about equal width."""
    n = -(-nx // (256 - order))
    return -(-nx // n)


def _geometry(nx, nz, order):
    """Sources at an interior corner cell, inside, in the thin band next to the frame, two in one
    column (they add in source order) and one on each side of the seam of the first two strips
    (a velocity there is recomputed by the neighbouring strip's halo lanes)."""
    edge = _strip_width(nx, order)  # columns of the first strip (1-based: the next strip starts at edge + 1)
    sx = [1, 2 * nx // 5, nx - 2, 2 * nx // 5, min(edge, nx - 1), min(edge + 1, nx)]
    sz = [1, nz // 3, nz - 3, nz // 2, nz // 4, nz // 5]
    return sx, sz


def _solo(monkeypatch, unfused, rows, nx, nz, order, storage, nbc, src, nt=60):
    import fomo
    import fomo._lib as L
    from fomo.api import Plan
    _set(monkeypatch, unfused, rows)
    vp, vs, rho = _model(nx, nz, 23)
    w = fomo.ricker_wavelet(18.0, 8e-4, nt)
    rx = np.arange(1, nx + 1, 3)
    sx, sz = _geometry(nx, nz, order)
    amps = [1.0, 0.5, -1.0, 0.25, 0.75, -0.6][: len(sx)]
    p = Plan(L.ELASTIC2D, vp, rho, 10.0, 8e-4, nt, vs=vs, sx=sx, sz=sz, rx=rx, rz=np.full(len(rx), 2),
             wavelet=np.stack([a * w for a in amps]), nbc=nbc, fd_order=order, storage=storage, tile_cfg=1,
             snap_interval=7, use_graph=True, source_type=src)
    try:
        names = [x["name"] for x in p.launches()]
        assert any("interior step" in n for n in names) != unfused, names
        p.advance(nt)
        p.sync()
        return p.seis(2) + [p.field(k) for k in range(5)] + [p.snaps()]
    finally:
        p.close()


def _same(a, b):
    assert len(a) == len(b)
    for k, (x, y) in enumerate(zip(a, b)):
        assert x.shape == y.shape, k
        assert np.array_equal(x, y), (k, float(np.abs(x.astype(np.float64) - y).max()))
    assert np.abs(a[0]).max() > 0 and np.abs(a[-1]).max() > 0


@pytest.mark.parametrize("src", [0, 1])
@pytest.mark.parametrize("order,storage", [(2, "f32"), (4, "f32"), (6, "f32"), (8, "f32"), (10, "f32"), (8, "f16"),
                                           (8, "bf16"), (8, "mat16"), (2, "f16")])
def test_fused_elastic_equals_two_sweeps(gpu, monkeypatch, order, storage, src):
    """560 x 171: strips of at most 256 - 2N columns, of about equal width (187 + 187 + 186 at every
    order), by 24 rows (a partial last one)."""
    import fomo._lib as L
    st = L.SRC_EXPLOSIVE if src == 0 else L.SRC_FZ
    print(f"fused elastic2d order {order} storage {storage} source {src}", flush=True)
    a = _solo(monkeypatch, False, 24, 560, 171, order, storage, nbc=12, src=st)
    b = _solo(monkeypatch, True, None, 560, 171, order, storage, nbc=12, src=st)
    _same(a, b)


@pytest.mark.parametrize("src", [0, 1])
def test_fused_elastic_bitwise_vs_oracle(gpu, orc, monkeypatch, src):
    """The fused family against the CPU oracle directly (FD8, both sources)."""
    import fomo
    import fomo._lib as L
    from fomo.api import Plan
    _set(monkeypatch, False, 32)
    nx, nz, nt, order = 300, 140, 70, 8
    vp, vs, rho = _model(nx, nz, 31)
    w = fomo.ricker_wavelet(18.0, 8e-4, nt)
    rx = np.arange(1, nx + 1, 3)
    rz = np.full(len(rx), 2)
    sx, sz = _geometry(nx, nz, order)
    wav = np.stack([w] * len(sx))
    st = 0 if src == 0 else L.SRC_FZ
    p = Plan(L.ELASTIC2D, vp, rho, 10.0, 8e-4, nt, vs=vs, sx=sx, sz=sz, rx=rx, rz=rz, wavelet=wav, nbc=10,
             fd_order=order, source_type=st, tile_cfg=1)
    try:
        assert any("interior step" in x["name"] for x in p.launches())
        p.advance(nt)
        p.sync()
        seis = p.seis(2)
        state = [p.field(k) for k in range(5)]
    finally:
        p.close()
    ref = orc.elastic2d(vp, vs, rho, 10.0, 8e-4, nt, sx=sx, sz=sz, rx=rx, rz=rz, wavelet=wav, nbc=10, fd_order=order,
                        source_type=st, final_state=True)
    for got, key in zip(seis, ("seis_vx", "seis_vz")):
        assert np.array_equal(got, ref[key]), key
    for got, key in zip(state, FIELDS):
        assert np.array_equal(got, ref["state"][key]), key


@pytest.mark.parametrize("rows", [8, 4096])
def test_fused_elastic_strip_lengths_and_thin_frame(gpu, monkeypatch, rows):
    """Strips of 8 rows (shorter than the rows of velocities recomputed past each end at FD10) and
    one strip per column (4096 rows), with nbc = 0, over 492 columns: two strips of the full 246
    columns; vertical-force sources."""
    import fomo._lib as L
    a = _solo(monkeypatch, False, rows, 492, 203, 10, "f32", nbc=0, src=L.SRC_FZ)
    b = _solo(monkeypatch, True, None, 492, 203, 10, "f32", nbc=0, src=L.SRC_FZ)
    _same(a, b)


@pytest.mark.parametrize("nbc", [0, 3])
@pytest.mark.parametrize("src", [0, 1])
def test_fused_elastic_tiny_grid(gpu, monkeypatch, nbc, src):
    """7 x 5 cells at FD10: one strip of 7 columns in a 256-lane block, 5 rows (fewer than the
    rows of velocities recomputed past each end), two sources in one column."""
    import fomo._lib as L
    st = L.SRC_EXPLOSIVE if src == 0 else L.SRC_FZ
    a = _solo(monkeypatch, False, 8, 7, 5, 10, "f32", nbc=nbc, src=st, nt=30)
    b = _solo(monkeypatch, True, None, 7, 5, 10, "f32", nbc=nbc, src=st, nt=30)
    _same(a, b)


def test_fused_elastic_batch_of_uneven_windows(gpu, monkeypatch):
    """Three windows of different widths (one, two and one strips in x) in one batch."""
    import fomo
    import fomo._lib as L
    from fomo.api import Plan
    nx, nz, nt = 420, 130, 50
    vp, vs, rho = _model(nx, nz, 5)
    w = fomo.ricker_wavelet(15.0, 8e-4, nt)
    shots = []
    for x0, wd in ((1, 180), (97, 301), (260, 161)):
        rxl = np.arange(1, wd + 1, 2)
        shots.append(dict(x0=x0, nx=wd, sx=[wd // 2], sz=[5], rx=rxl, rz=np.full(len(rxl), 3), wavelet=w[None]))
    outs = []
    for unfused in (False, True):
        _set(monkeypatch, unfused, 40)
        p = Plan(L.ELASTIC2D, vp, rho, 10.0, 8e-4, nt, vs=vs, shots=shots, nbc=10, fd_order=8, tile_cfg=1)
        try:
            p.advance(nt)
            o = []
            for s in range(3):
                o += p.seis(2, shot=s) + [p.field(k, shot=s) for k in range(5)]
            outs.append(o)
        finally:
            p.close()
    for x, y in zip(*outs):
        assert np.array_equal(x, y)
    assert np.abs(outs[0][0]).max() > 0


def test_fused_elastic_sponge_boundary(gpu, monkeypatch):
    """The Cerjan sponge (frame-only damping) through the one-call solver, receivers and snapshots."""
    import fomo
    nx, nz, nt = 200, 150, 80
    vp, vs, rho = _model(nx, nz, 8)
    rx = np.arange(1, nx + 1, 2)
    res = []
    for unfused in (False, True):
        _set(monkeypatch, unfused, 32)
        r = fomo.elastic2d(vp, vs, rho, 10.0, 8e-4, nt, 15.0, sx=[nx // 2], sz=[4], rx=rx, rz=np.full(len(rx), 2),
                           nbc=20, fd_order=8, boundary="sponge", snap_interval=10)
        res.append([r.seis_vx, r.seis_vz] + list(r.snaps))
    for x, y in zip(*res):
        assert np.array_equal(x, y)
