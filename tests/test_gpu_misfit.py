"""FP64 trace misfit / energy reduced on the GPU (fomo_trace_misfit) against numpy's FP64 sums of
the downloaded traces.  The GPU adds in another order than numpy, so the bar is 1e-12 relative;
repeated calls must give the same bits (fixed reduction order)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RTOL = 1e-12


def sums(a, b=None):
    a = a.astype(np.float64)
    b = np.zeros_like(a) if b is None else b.astype(np.float64)
    return ((a - b) ** 2).sum(axis=1), (b ** 2).sum(axis=1)


def plan2d(nbc, nt, nx=150, nz=90, **kw):
    import fomo
    from fomo import _lib
    vp = np.full((nx, nz), 2500.0, np.float32)
    vp[:, nz // 2:] = 3300.0
    rho = np.full((nx, nz), 2100.0, np.float32)
    rx = np.arange(1, nx + 1)  # 150 receivers: three blocks of 64, the last one ragged
    w = fomo.ricker_wavelet(15.0, 1e-3, nt)
    p = fomo.Plan(_lib.ACOUSTIC2D, vp, rho, 10.0, 1e-3, nt, sx=[nx // 2], sz=[10], rx=rx, rz=np.full(nx, 6),
                  wavelet=w[None], nbc=nbc, fd_order=8, **kw)
    p.advance(nt)
    return p


@pytest.mark.parametrize("nt", [1, 255, 700])
def test_misfit_and_energy_match_numpy(gpu, nt):
    a, b = plan2d(20, nt), plan2d(6, nt)
    try:
        sa, sb = a.seis(3), b.seis(3)
        for slot in range(3):
            num, den = a.trace_misfit(b, slot=slot)
            rn, rd = sums(sa[slot], sb[slot])
            assert np.allclose(num, rn, rtol=RTOL, atol=0), slot
            assert np.allclose(den, rd, rtol=RTOL, atol=0), slot
            n2, d2 = a.trace_misfit(b, slot=slot)
            assert np.array_equal(num, n2) and np.array_equal(den, d2)
            e, z = a.trace_misfit(None, slot=slot)
            assert np.allclose(e, sums(sa[slot])[0], rtol=RTOL, atol=0)
            assert not z.any()
        if nt > 1:
            assert num.max() > 0
            r = a.relative_residual(b)
            ref = np.linalg.norm(sa[0].astype(np.float64) - sb[0]) / np.linalg.norm(sb[0].astype(np.float64))
            assert abs(r - ref) <= 1e-12 * ref
    finally:
        a.close()
        b.close()


def test_misfit_time_window_and_same_plan(gpu):
    nt = 600
    a, b = plan2d(20, nt), plan2d(10, nt)
    try:
        sa, sb = a.seis(1)[0], b.seis(1)[0]
        for t0, t1 in ((0, nt), (37, 611 - 11), (299, 300), (400, 400)):
            num, den = a.trace_misfit(b, t0=t0, t1=t1)
            rn, rd = sums(sa[:, t0:t1], sb[:, t0:t1])
            assert np.allclose(num, rn, rtol=RTOL, atol=0)
            assert np.allclose(den, rd, rtol=RTOL, atol=0)
        num, den = a.trace_misfit(a)
        assert not num.any()
        assert np.allclose(den, sums(sa)[0], rtol=RTOL, atol=0)
    finally:
        a.close()
        b.close()


def test_misfit_between_batch_shots_and_3d(gpu):
    """Windows of one batch plan with equal receiver counts, and acoustic3d's p slot."""
    import fomo
    from fomo import _lib
    nx, nz, nt = 200, 80, 300
    vp = np.full((nx, nz), 2800.0, np.float32)
    rho = np.full((nx, nz), 2000.0, np.float32)
    w = fomo.ricker_wavelet(15.0, 1e-3, nt)
    shots = [dict(x0=x0, nx=100, sx=[50], sz=[8], rx=np.arange(1, 101), rz=np.full(100, 5), wavelet=w[None])
             for x0 in (1, 60, 101)]
    p = fomo.Plan(_lib.ACOUSTIC2D, vp, rho, 10.0, 1e-3, nt, shots=shots, nbc=10, fd_order=8)
    try:
        p.advance(nt)
        s0, s2 = p.seis(1, shot=0)[0], p.seis(1, shot=2)[0]
        num, den = p.trace_misfit(p, shot=2, other_shot=0)
        rn, rd = sums(s2, s0)
        assert np.allclose(num, rn, rtol=RTOL, atol=0) and np.allclose(den, rd, rtol=RTOL, atol=0)
    finally:
        p.close()
    n3 = 40
    vp3 = np.full((n3, n3, n3), 2500.0, np.float32)
    rho3 = np.full((n3, n3, n3), 2000.0, np.float32)
    rx, ry = np.meshgrid(np.arange(1, n3 + 1, 3), np.arange(1, n3 + 1, 3), indexing="ij")
    rx, ry = rx.ravel(), ry.ravel()
    w3 = fomo.ricker_wavelet(15.0, 1e-3, 120)
    q = fomo.Plan(_lib.ACOUSTIC3D, vp3, rho3, 10.0, 1e-3, 120, sx=[20], sy=[20], sz=[6], rx=rx, ry=ry,
                  rz=np.full(len(rx), 4), wavelet=w3[None], nbc=8, fd_order=8, record_p=True)
    try:
        q.advance(120)
        s = q.seis(4)
        for slot in (0, 3):
            num, den = q.trace_misfit(None, slot=slot)
            assert np.allclose(num, sums(s[slot])[0], rtol=RTOL, atol=0)
        assert num.max() > 0
    finally:
        q.close()


def test_misfit_argument_errors(gpu):
    import fomo
    a, b, c = plan2d(10, 100), plan2d(10, 120), plan2d(10, 100, nx=120)
    try:
        with pytest.raises(fomo.ArgumentError):
            a.trace_misfit(b)  # different nt
        with pytest.raises(fomo.ArgumentError):
            a.trace_misfit(c)  # different receiver counts
        with pytest.raises(fomo.ArgumentError):
            a.trace_misfit(None, slot=3)
        with pytest.raises(fomo.ArgumentError):
            a.trace_misfit(None, t0=50, t1=40)
        with pytest.raises(fomo.ArgumentError):
            a.trace_misfit(None, t1=101)
        with pytest.raises(fomo.ArgumentError):
            a.trace_misfit(None, shot=1)
        c.close()
        with pytest.raises(fomo.ArgumentError):
            a.trace_misfit(c)  # a closed plan is an error, not a missing one
    finally:
        for p in (a, b, c):
            p.close()
