"""GPU batch engine: batched shots are bit-identical to their solo runs (verify_batch.jl:50-151),
for the reference's shared-grid batch and for uneven per-shot windows; results do not depend on
the tile shape the autotuner picks."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def model(nx, nz, seed):
    from fomo.models import layered_2d
    return layered_2d(nx, nz, seed=seed, n_interfaces=5, jitter=4.0)


def test_acoustic2d_batch_equals_solo(gpu):
    """verify_batch.jl:116-151: *_batch API vs per-shot API, exactly."""
    import fomo
    nx, nz, nt = 150, 90, 220
    vp, vs, rho = model(nx, nz, 31)
    sx = np.array([[20, 40], [75, 76], [130, 10]])
    sz = np.array([[5, 30], [8, 8], [60, 12]])
    rx = np.arange(1, nx + 1, 4)
    rz = np.full(len(rx), 6)
    b = fomo.acoustic2d_batch(vp, rho, 10.0, 1e-3, nt, 15.0, sx=sx, sz=sz, rx=rx, rz=rz, nbc=15, fd_order=8)
    assert b.seis_p.shape == (len(rx), nt, 3) and b.stats.n_shots == 3
    for s in range(3):
        r = fomo.acoustic2d(vp, rho, 10.0, 1e-3, nt, 15.0, sx=list(sx[s]), sz=list(sz[s]), rx=rx, rz=rz, nbc=15,
                            fd_order=8)
        assert np.array_equal(b.seis_p[:, :, s], r.seis_p)
        assert np.array_equal(b.seis_vx[:, :, s], r.seis_vx)
        assert np.array_equal(b.seis_vz[:, :, s], r.seis_vz)


def test_elastic2d_batch_equals_solo(gpu):
    import fomo
    nx, nz, nt = 140, 100, 200
    vp, vs, rho = model(nx, nz, 32)
    sx = [30, 70, 110, 71]
    sz = [10, 50, 20, 9]
    rx = np.arange(1, nx + 1, 3)
    rz = np.full(len(rx), 5)
    b = fomo.elastic2d_batch(vp, vs, rho, 5.0, 4e-4, nt, 20.0, sx=sx, sz=sz, rx=rx, rz=rz, nbc=12, fd_order=6)
    for s in range(4):
        r = fomo.elastic2d(vp, vs, rho, 5.0, 4e-4, nt, 20.0, sx=[sx[s]], sz=[sz[s]], rx=rx, rz=rz, nbc=12,
                           fd_order=6)
        assert np.array_equal(b.seis_vx[:, :, s], r.seis_vx)
        assert np.array_equal(b.seis_vz[:, :, s], r.seis_vz)


@pytest.mark.parametrize("eq", [1, 2])
def test_uneven_windows_equal_solo_on_cropped_model(gpu, orc, eq):
    """Windowed batch (uneven widths, per-window v_ref): each shot equals the solo run on its
    cropped model, and the first shot equals the oracle on that crop."""
    import fomo
    import fomo._lib as L
    from fomo.api import Plan
    nx, nz, nt = 420, 96, 180
    vp, vs, rho = model(nx, nz, 33)
    w = fomo.ricker_wavelet(15.0, 1e-3, nt)
    wins = [(1, 130), (100, 257), (250, 171), (37, 64), (300, 121)]
    shots = []
    for k, (x0, wd) in enumerate(wins):
        sxl = wd // 2 + (k % 3)
        rxl = np.arange(1, wd + 1, 2)
        shots.append(dict(x0=x0, nx=wd, sx=[sxl], sz=[4 + k], rx=rxl, rz=np.full(len(rxl), 3), wavelet=w[None]))
    kw = dict(vs=vs if eq == 2 else None, nbc=10, fd_order=8)
    plan = Plan(eq, vp, rho, 10.0, 1e-3, nt, shots=shots, **kw)
    plan.advance(nt)
    plan.sync()
    nsl = 3 if eq == 1 else 2
    for k, (x0, wd) in enumerate(wins):
        got = plan.seis(nsl, shot=k)
        crop = lambda a: np.asfortranarray(a[x0 - 1:x0 - 1 + wd])  # noqa: E731
        s = shots[k]
        solo = Plan(eq, crop(vp), crop(rho), 10.0, 1e-3, nt, sx=s["sx"], sz=s["sz"], rx=s["rx"], rz=s["rz"],
                    wavelet=w[None], vs=crop(vs) if eq == 2 else None, nbc=10, fd_order=8)
        solo.advance(nt)
        ref = solo.seis(nsl)
        solo.close()
        for a, b in zip(got, ref):
            assert np.array_equal(a, b), (k, x0, wd)
        if k == 1:
            fn = orc.acoustic2d if eq == L.ACOUSTIC2D else orc.elastic2d
            args = (crop(vp), crop(rho)) if eq == 1 else (crop(vp), crop(vs), crop(rho))
            o = fn(*args, 10.0, 1e-3, nt, sx=s["sx"], sz=s["sz"], rx=s["rx"], rz=s["rz"], wavelet=w[None], nbc=10,
                   fd_order=8)
            keys = ("seis_p", "seis_vx", "seis_vz") if eq == 1 else ("seis_vx", "seis_vz")
            for a, key in zip(got, keys):
                assert np.array_equal(a, o[key])
    plan.close()


def test_tile_shapes_agree_bitwise(gpu):
    import fomo
    import fomo._lib as L
    from fomo.api import Plan
    nx, nz, nt = 200, 150, 120
    vp, vs, rho = model(nx, nz, 34)
    w = fomo.ricker_wavelet(15.0, 1e-3, nt)
    outs = []
    for cfg in (1, 2, 3):
        p = Plan(L.ELASTIC2D, vp, rho, 10.0, 1e-3, nt, vs=vs, sx=[100], sz=[20], rx=np.arange(1, nx + 1),
                 rz=np.full(nx, 4), wavelet=w[None], nbc=20, fd_order=8, tile_cfg=cfg)
        p.advance(nt)
        outs.append(p.seis(2) + [p.field(k) for k in range(5)])
        p.close()
    for a, b in zip(outs[0], outs[1]):
        assert np.array_equal(a, b)
    for a, b in zip(outs[0], outs[2]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("weight", ["0.5", "1.5", "4"])
def test_frame_weight_of_the_layout_changes_no_bits(gpu, monkeypatch, weight):
    """FOMO_FRAME_COST only reorders the work list across XCDs: an uneven elastic batch gives the
    same bits under any weight as under the default."""
    import fomo
    import fomo._lib as L
    from fomo.api import Plan
    nx, nz, nt = 400, 150, 90
    vp, vs, rho = model(nx, nz, 35)
    w = fomo.ricker_wavelet(15.0, 1e-3, nt)
    shots = []
    for x0, wd in ((1, 400), (37, 150), (200, 201)):
        rx = np.arange(1, wd + 1, 3)
        shots.append(dict(x0=x0, nx=wd, sx=[wd // 2], sz=[10], rx=rx, rz=np.full(len(rx), 5), wavelet=w[None]))
    outs = []
    for env in (None, weight):
        if env is None:
            monkeypatch.delenv("FOMO_FRAME_COST", raising=False)
        else:
            monkeypatch.setenv("FOMO_FRAME_COST", env)
        p = Plan(L.ELASTIC2D, vp, rho, 10.0, 1e-3, nt, vs=vs, shots=shots, nbc=20, fd_order=8, tile_cfg=2)
        try:
            p.advance(nt)
            outs.append([a for s in range(len(shots)) for a in p.seis(2, shot=s)])
        finally:
            p.close()
    monkeypatch.delenv("FOMO_FRAME_COST", raising=False)
    assert np.abs(outs[0][0]).max() > 0
    for a, b in zip(outs[0], outs[1]):
        assert np.array_equal(a, b)
