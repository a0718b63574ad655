"""acoustic2d: the one-launch interior step (new velocities kept in LDS on their way to the new
pressure; the frame's velocities and pressure in launches of their own) gives the bits of the two
separate sweeps (FOMO_2D_UNFUSED=1) -- receivers, final fields and snapshots -- for every FD
order's halo widths, for 16-bit storage, for strips shorter than the halo and single-strip
columns, for a frame of nbc = 0, for a batch of uneven windows, for the sponge boundary and for
the learned boundary.  The autotuner picks the fused family where it is faster (large grids);
FOMO_2D_FUSED=1 / FOMO_2D_UNFUSED=1 keep one family, as here."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _model(nx, nz, seed):
    from fomo.models import layered_2d
    vp, _, rho = layered_2d(nx, nz, seed=seed, n_interfaces=4, vacuum_top=False)
    return vp, rho


def _set(monkeypatch, unfused, rows):
    # one family each: the autotuner would otherwise pick whichever is faster at these small sizes
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


def _solo(monkeypatch, unfused, rows, nx, nz, order, storage, nbc, nt=60, learned=None):
    import fomo
    import fomo._lib as L
    from fomo.api import Plan
    _set(monkeypatch, unfused, rows)
    vp, rho = _model(nx, nz, 21)
    w = fomo.ricker_wavelet(18.0, 8e-4, nt)
    rx = np.arange(1, nx + 1, 3)
    # sources at an interior corner cell, inside, in the thin band next to the frame and on each
    # side of the seam of the first two strips (strips of at most 257 - 2N columns, of about equal
    # width)
    n = -(-nx // (257 - order))
    edge = -(-nx // n)
    sx, sz = [1, nx // 2, nx - 2, min(edge, nx - 1), min(edge + 1, nx)], [1, nz // 3, nz - 3, nz // 4, nz // 5]
    p = Plan(L.ACOUSTIC2D, vp, rho, 10.0, 8e-4, nt, sx=sx, sz=sz, rx=rx, rz=np.full(len(rx), 2),
             wavelet=np.stack([w, 0.5 * w, -w, 0.75 * w, -0.6 * w]), nbc=nbc, fd_order=order, storage=storage, tile_cfg=1,
             snap_interval=7, use_graph=True)
    try:
        if learned is not None:
            from fomo.learned import enable
            enable(p, learned)
        names = [x["name"] for x in p.launches()]
        assert any("interior step" in n for n in names) != unfused, names
        p.advance(nt)
        p.sync()
        return p.seis(3) + [p.field(k) for k in range(3)] + [p.snaps()]
    finally:
        p.close()


def _same(a, b):
    assert len(a) == len(b)
    for k, (x, y) in enumerate(zip(a, b)):
        assert x.shape == y.shape, k
        assert np.array_equal(x, y), (k, float(np.abs(x.astype(np.float64) - y).max()))
    assert np.abs(a[0]).max() > 0 and np.abs(a[-1]).max() > 0


@pytest.mark.parametrize("order,storage", [(2, "f32"), (4, "f32"), (6, "f32"), (8, "f32"), (10, "f32"), (8, "f16"),
                                           (8, "bf16"), (8, "mat16"), (10, "bf16"), (2, "f16")])
def test_fused_equals_two_sweeps(gpu, monkeypatch, order, storage):
    """560 x 171: strips of at most 257 - 2N columns, of about equal width (187 + 187 + 186 at every
    order), by 24 rows (a partial last one)."""
    a = _solo(monkeypatch, False, 24, 560, 171, order, storage, nbc=12)
    b = _solo(monkeypatch, True, None, 560, 171, order, storage, nbc=12)
    _same(a, b)


@pytest.mark.parametrize("rows", [8, 4096])
def test_fused_strip_lengths_and_thin_frame(gpu, monkeypatch, rows):
    """Strips of 8 rows (shorter than the 2N - 1 = 9 rows of velocities recomputed past each end at
    FD10) and one strip per column (4096 rows), with nbc = 0 (the frame is the N-cell stencil pad),
    over 494 columns: two strips of the full 247 columns."""
    a = _solo(monkeypatch, False, rows, 494, 203, 10, "f32", nbc=0)
    b = _solo(monkeypatch, True, None, 494, 203, 10, "f32", nbc=0)
    _same(a, b)


@pytest.mark.parametrize("nbc", [0, 3])
def test_fused_tiny_grid(gpu, monkeypatch, nbc):
    """7 x 5 cells at FD10: one strip of 7 columns in a 256-lane block, 5 rows (fewer than the
    rows of velocities recomputed past each end), every source cell within the stencil's reach of
    the frame."""
    a = _solo(monkeypatch, False, 8, 7, 5, 10, "f32", nbc=nbc, nt=30)
    b = _solo(monkeypatch, True, None, 7, 5, 10, "f32", nbc=nbc, nt=30)
    _same(a, b)


def test_fused_batch_of_uneven_windows(gpu, monkeypatch):
    """Three windows of different widths (one, two and one strips in x) in one batch."""
    import fomo
    import fomo._lib as L
    from fomo.api import Plan
    nx, nz, nt = 420, 130, 50
    vp, rho = _model(nx, nz, 5)
    w = fomo.ricker_wavelet(15.0, 8e-4, nt)
    shots = []
    for x0, wd in ((1, 180), (97, 301), (260, 161)):
        rxl = np.arange(1, wd + 1, 2)
        shots.append(dict(x0=x0, nx=wd, sx=[wd // 2], sz=[5], rx=rxl, rz=np.full(len(rxl), 3), wavelet=w[None]))
    outs = []
    for unfused in (False, True):
        _set(monkeypatch, unfused, 40)
        p = Plan(L.ACOUSTIC2D, vp, rho, 10.0, 8e-4, nt, shots=shots, nbc=10, fd_order=8, tile_cfg=1)
        try:
            p.advance(nt)
            o = []
            for s in range(3):
                o += p.seis(3, shot=s) + [p.field(k, shot=s) for k in range(3)]
            outs.append(o)
        finally:
            p.close()
    for x, y in zip(*outs):
        assert np.array_equal(x, y)
    assert np.abs(outs[0][0]).max() > 0


def test_fused_sponge_boundary(gpu, monkeypatch):
    """The Cerjan sponge (frame-only damping) through the one-call solver, receivers and snapshots."""
    import fomo
    nx, nz, nt = 200, 150, 80
    vp, rho = _model(nx, nz, 8)
    rx = np.arange(1, nx + 1, 2)
    res = []
    for unfused in (False, True):
        _set(monkeypatch, unfused, 32)
        r = fomo.acoustic2d(vp, rho, 10.0, 8e-4, nt, 15.0, sx=[nx // 2], sz=[4], rx=rx, rz=np.full(len(rx), 2),
                            nbc=20, fd_order=8, boundary="sponge", snap_interval=10)
        res.append([r.seis_p, r.seis_vx, r.seis_vz] + list(r.snaps))
    for x, y in zip(*res):
        assert np.array_equal(x, y)


def test_fused_learned_boundary(gpu, monkeypatch):
    """The learned thin boundary (blend-weight corrections at frame cells from the network) beside
    the fused interior: same bits as beside the separate sweeps."""
    from fomo.learned import load_weights
    wts = load_weights()
    a = _solo(monkeypatch, False, 48, 260, 180, 8, "f32", nbc=16, learned=wts)
    b = _solo(monkeypatch, True, None, 260, 180, 8, "f32", nbc=16, learned=wts)
    _same(a, b)
