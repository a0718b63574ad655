"""scalar2d: the row-march interior launch (u^n of a column in a register queue, its row staged in
LDS, one barrier per row, strips of odd row bands marching up; the frame and the receivers in a
launch of their own) gives the bits of the separate sweep (FOMO_2D_UNFUSED=1) and of the CPU
oracle -- receivers and the final field (scalar2d records no snapshots) -- for every FD order,
16-bit storage, sources
on both sides of a strip seam and two in one column, strips shorter than the stencil, a 7 x 5 grid
and a frame of nbc = 0."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _model(nx, nz):
    vp = np.full((nx, nz), 3000.0, np.float32)
    vp[:, nz // 2:] = 3600.0
    vp[nx // 3:, : nz // 3] = 2700.0
    return vp


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


def _geometry(nx, nz, order):
    """Sources at an interior corner cell, two in one column, one in the thin band next to the frame
    and one on each side of the seam of the first two strips (at most 256 - 2N columns each, of
    about equal width)."""
    n = -(-nx // (256 - order))
    edge = -(-nx // n)
    sx = [1, 2 * nx // 5, 2 * nx // 5, nx - 1, min(edge, nx - 1), min(edge + 1, nx)]
    sz = [1, nz // 3, 2 * nz // 3, nz - 2, max(1, nz // 4), max(1, nz // 5)]
    return sx, sz


def _run(monkeypatch, unfused, rows, nx, nz, order, storage, nbc, nt=60):
    import fomo
    import fomo._lib as L
    from fomo.api import Plan
    _set(monkeypatch, unfused, rows)
    vp = _model(nx, nz)
    w = fomo.ricker_wavelet(18.0, 1e-3, nt)
    sx, sz = _geometry(nx, nz, order)
    amps = [1.0, 0.5, -1.0, 0.25, 0.75, -0.6]
    wm = np.stack([np.float32(a) * w for a in amps]).astype(np.float32)
    rx = np.arange(1, nx + 1, 3)
    p = Plan(L.SCALAR2D, vp, vp, 10.0, 1e-3, nt, sx=sx, sz=sz, rx=rx, rz=np.full(len(rx), 2), wavelet=wm, nbc=nbc,
             fd_order=order, storage=storage, tile_cfg=1, use_graph=True)
    try:
        names = [x["name"] for x in p.launches()]
        assert any("interior step" in n for n in names) != unfused, names
        p.advance(nt)
        p.sync()
        return p.seis(1) + [p.field(0)]
    finally:
        p.close()


def _same(a, b):
    assert len(a) == len(b)
    for k, (x, y) in enumerate(zip(a, b)):
        x, y = np.asarray(x), np.asarray(y)
        assert x.shape == y.shape, k
        assert np.array_equal(x, y), (k, float(np.abs(x.astype(np.float64) - y).max()))
    assert np.abs(np.asarray(a[0])).max() > 0 and np.abs(np.asarray(a[-1])).max() > 0


@pytest.mark.parametrize("order,storage", [(2, "f32"), (4, "f32"), (6, "f32"), (8, "f32"), (10, "f32"), (8, "f16"),
                                           (8, "bf16"), (2, "f16")])
def test_fused_scalar_equals_the_sweep(gpu, monkeypatch, order, storage):
    """560 x 171: three strips of about 187 columns by 24 rows (a partial last one), odd row bands
    marching up."""
    a = _run(monkeypatch, False, 24, 560, 171, order, storage, nbc=12)
    b = _run(monkeypatch, True, None, 560, 171, order, storage, nbc=12)
    _same(a, b)


def test_fused_scalar_bitwise_vs_oracle(gpu, orc, monkeypatch):
    """The fused family against the CPU oracle directly (FD8): receivers and the final field."""
    import fomo
    import fomo._lib as L
    from fomo.api import Plan
    _set(monkeypatch, False, 32)
    nx, nz, nt, order, nbc = 300, 140, 70, 8, 10
    vp = _model(nx, nz)
    w = fomo.ricker_wavelet(18.0, 1e-3, nt)
    sx, sz = _geometry(nx, nz, order)
    wm = np.stack([w] * len(sx)).astype(np.float32)
    rx = np.arange(1, nx + 1, 3)
    rz = np.full(len(rx), 2)
    p = Plan(L.SCALAR2D, vp, vp, 10.0, 1e-3, nt, sx=sx, sz=sz, rx=rx, rz=rz, wavelet=wm, nbc=nbc, fd_order=order,
             tile_cfg=1)
    try:
        assert any("interior step" in x["name"] for x in p.launches())
        p.advance(nt)
        p.sync()
        (su,) = p.seis(1)
        u = p.field(0)
    finally:
        p.close()
    ref = orc.scalar2d(vp, 10.0, 1e-3, nt, sx=sx, sz=sz, rx=rx, rz=rz, wavelet=wm, nbc=nbc, fd_order=order,
                       final_state=True)
    assert np.abs(su).max() > 0
    assert np.array_equal(su, ref["seis_u"])
    assert np.array_equal(u, ref["state"]["u"])


@pytest.mark.parametrize("rows", [8, 4096])
def test_fused_scalar_strip_lengths_and_thin_frame(gpu, monkeypatch, rows):
    """Strips of 8 rows (shorter than the z-stencil at FD10) and one strip per column, with nbc = 0,
    over 492 columns: two strips of the full 246."""
    a = _run(monkeypatch, False, rows, 492, 203, 10, "f32", nbc=0)
    b = _run(monkeypatch, True, None, 492, 203, 10, "f32", nbc=0)
    _same(a, b)


@pytest.mark.parametrize("nbc", [0, 3])
def test_fused_scalar_tiny_grid(gpu, monkeypatch, nbc):
    """7 x 5 cells at FD10: one strip of 7 columns in a 256-lane block, 5 rows."""
    a = _run(monkeypatch, False, 8, 7, 5, 10, "f32", nbc=nbc, nt=30)
    b = _run(monkeypatch, True, None, 7, 5, 10, "f32", nbc=nbc, nt=30)
    _same(a, b)
