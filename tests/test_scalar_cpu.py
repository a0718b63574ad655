"""scalar2d on the CPU: centred second-derivative weights and the oracle's scalar solver.

The oracle's scalar2d restates src/equations/scalar2d/scalar2d.jl:33-81 (the reference's own CPU
step is test/verify_scalar.jl:33-86).  It is pinned here by:
  * closed forms: the weights differentiate x^0 .. x^(2M) exactly (moment conditions);
  * the reference's own acceptance check, grid self-convergence (test/verify_scalar.jl:152-178:
    dh 10 m vs 5 m at the same physical source and receivers, relative L2 <= 0.1), with a
    planted coefficient mistake that fails it;
  * invariants: x-mirror symmetry, linearity in the source, FP32 against FP64;
  * absorption: the HABC leaves a small residual against a boundary-free edge-extended domain, and
    removing the boundary's effect (nbc tiny) makes it large.
"""
import math

import numpy as np
import pytest


def test_centered_d2_moment_conditions(orc):
    """f'' ~ c0 f(0) + sum_l d2_l (f(l) + f(-l)) is exact for x^0 .. x^(2M): c0 + 2 sum d2 = 0,
    sum d2_l l^2 = 1, sum d2_l l^(2k) = 0 for k = 2..M (to Float32 rounding)."""
    for order in (2, 4, 6, 8, 10):
        c0, d2 = orc.centered_d2(order)
        m = order // 2
        assert len(d2) == m
        l = np.arange(1, m + 1, dtype=np.float64)
        d = d2.astype(np.float64)
        assert abs(c0 + 2 * d.sum()) < 2e-6
        assert abs((d * l ** 2).sum() - 1.0) < 2e-6
        for k in range(2, m + 1):
            assert abs((d * l ** (2 * k)).sum()) < 2e-5 * m ** (2 * k), (order, k)
    with pytest.raises(ValueError):
        orc.centered_d2(12)


def test_host_centered_d2_matches_oracle(orc):
    import fomo
    for order in (2, 4, 6, 8, 10):
        c0, d2 = fomo.get_centered_d2(order)
        oc0, od2 = orc.centered_d2(order)
        assert np.float32(c0) == np.float32(oc0)
        assert np.array_equal(d2, od2)


def _ricker(orc, f0, dt, nt):
    return orc.ricker(f0, dt, nt)


@pytest.mark.slow
def test_grid_self_convergence(orc):
    """test/verify_scalar.jl:152-178: constant 3000 m/s, 200x160 at dh 10 m / dt 1 ms / 500 steps
    against 399x319 at dh 5 m / dt 0.5 ms / 1000 steps (x = (i-1) dh maps to i2 = 2i-1), 12 Hz,
    source_scale :none, default nbc 50, FD8; the reference accepts relative L2 <= 0.1."""
    nx, nz, nt = 200, 160, 500
    sx1, sz1 = nx // 2, 10
    rx1 = np.arange(30, nx - 29, 20)
    rz1 = np.full(len(rx1), 12)
    w1 = _ricker(orc, 12.0, 1e-3, nt)
    w2 = _ricker(orc, 12.0, 5e-4, 2 * nt)
    vp1 = np.full((nx, nz), 3000.0, np.float32)
    vp2 = np.full((2 * nx - 1, 2 * nz - 1), 3000.0, np.float32)
    coarse = orc.scalar2d(vp1, 10.0, 1e-3, nt, sx=[sx1], sz=[sz1], rx=rx1, rz=rz1, wavelet=w1[None], nbc=50)
    fine = orc.scalar2d(vp2, 5.0, 5e-4, 2 * nt, sx=[2 * sx1 - 1], sz=[2 * sz1 - 1], rx=2 * rx1 - 1,
                        rz=2 * rz1 - 1, wavelet=w2[None], nbc=50)
    sc = coarse["seis_u"].astype(np.float64)
    sf = fine["seis_u"][:, 1::2].astype(np.float64)
    rel = math.sqrt(((sc - sf) ** 2).sum() / max((sf ** 2).sum(), 1e-30))
    print(f"scalar2d grid self-convergence rel L2 {rel:.3e}")
    assert rel <= 0.1
    assert np.abs(sc).max() > 0
    # planted mistake: c2 off by 6 % (a 3 % speed error, what a wrong dt/dh factor in c2 gives) on
    # the fine grid must fail the same check
    bad = orc.scalar2d(vp2 * np.float32(1.03), 5.0, 5e-4, 2 * nt, sx=[2 * sx1 - 1], sz=[2 * sz1 - 1],
                       rx=2 * rx1 - 1, rz=2 * rz1 - 1, wavelet=w2[None], nbc=50)
    sb = bad["seis_u"][:, 1::2].astype(np.float64)
    rel_bad = math.sqrt(((sc - sb) ** 2).sum() / (sb ** 2).sum())
    print(f"planted c2 error: rel L2 {rel_bad:.3e}")
    assert rel_bad > 0.1


def test_mirror_symmetry_linearity_and_precision(orc):
    nx, nz, nt = 121, 90, 300
    vp = np.full((nx, nz), 2500.0, np.float32)
    vp[:, 45:] = 3300.0
    w = _ricker(orc, 15.0, 1e-3, nt)
    rx = np.arange(1, nx + 1)
    rz = np.full(nx, 8)
    kw = dict(rx=rx, rz=rz, nbc=20, fd_order=8)
    a = orc.scalar2d(vp, 10.0, 1e-3, nt, sx=[61], sz=[10], wavelet=w[None], **kw)
    d = a["seis_u"]
    assert np.linalg.norm(d - d[::-1]) / np.linalg.norm(d) < 1e-5
    off = orc.scalar2d(vp, 10.0, 1e-3, nt, sx=[61], sz=[10], wavelet=w[None], **{**kw, "rx": np.clip(rx + 1, 1, nx)})
    assert np.linalg.norm(off["seis_u"][:-1] - off["seis_u"][:-1][::-1]) / np.linalg.norm(off["seis_u"]) > 1e-3
    b = orc.scalar2d(vp, 10.0, 1e-3, nt, sx=[30], sz=[20], wavelet=w[None], **kw)
    ab = orc.scalar2d(vp, 10.0, 1e-3, nt, sx=[61, 30], sz=[10, 20], wavelet=np.stack([w, w]), **kw)
    lin = np.linalg.norm(ab["seis_u"] - (a["seis_u"] + b["seis_u"])) / np.linalg.norm(ab["seis_u"])
    assert lin < 1e-5
    a64 = orc.scalar2d(vp, 10.0, 1e-3, nt, sx=[61], sz=[10], wavelet=w[None], precision="f64", **kw)
    e = np.linalg.norm(a["seis_u"] - a64["seis_u"], axis=1) / np.linalg.norm(a64["seis_u"], axis=1)
    assert e.max() < 1e-4


def test_habc_absorbs(orc):
    """Against a boundary-free, edge-extended domain (the method of test/verify_patch.jl:233-249),
    source and receivers 50 cells deep: nbc 20 leaves 1.4e-2 (nbc 1: 9.2e-2).  The scalar HABC is
    weaker than the first-order one (acoustic2d in the same geometry: 8.2e-4), a property of the
    reference's method -- one Higdon blend of u with f_old = u^n (scalar2d.jl:74-75) -- which this
    oracle restates term for term from the reference's own CPU step."""
    nx, nz, nt, L = 161, 121, 600, 260
    vp = np.full((nx, nz), 2500.0, np.float32)
    vp[:, 70:] = 3100.0
    w = _ricker(orc, 12.0, 1e-3, nt)
    rx = np.arange(1, nx + 1)
    rz = np.full(nx, 50)
    ext = np.pad(vp, L, mode="edge")
    ref = orc.scalar2d(ext, 10.0, 1e-3, nt, sx=[81 + L], sz=[50 + L], rx=rx + L, rz=rz + L, wavelet=w[None], nbc=20,
                       v_ref=2500.0)
    den = np.linalg.norm(ref["seis_u"])
    h = orc.scalar2d(vp, 10.0, 1e-3, nt, sx=[81], sz=[50], rx=rx, rz=rz, wavelet=w[None], nbc=20)
    resid = np.linalg.norm(h["seis_u"] - ref["seis_u"]) / den
    thin = orc.scalar2d(vp, 10.0, 1e-3, nt, sx=[81], sz=[50], rx=rx, rz=rz, wavelet=w[None], nbc=1)
    resid_thin = np.linalg.norm(thin["seis_u"] - ref["seis_u"]) / den
    print(f"scalar2d HABC residual nbc 20 {resid:.3e}, nbc 1 {resid_thin:.3e}")
    assert resid < 0.02
    assert resid_thin > 4 * resid


def test_scalar2d_argument_errors():
    import fomo
    vp = np.full((60, 40), 2000.0, np.float32)
    kw = dict(sx=[30], sz=[10], rx=[5, 6], rz=[5, 5], nbc=10, fd_order=8)
    with pytest.raises(fomo.ArgumentError):
        fomo.scalar2d(vp, 10.0, 9e-3, 100, 10.0, **kw)  # CFL (centred bound)
    with pytest.raises(fomo.ArgumentError):
        fomo.scalar2d(vp, 10.0, 1e-3, 100, 10.0, **{**kw, "sx": [61]})
    with pytest.raises(fomo.ArgumentError):
        fomo.scalar2d(vp, 10.0, 1e-3, 100, 10.0, **kw, source_scale="rho")
    with pytest.raises(fomo.ArgumentError):
        fomo.scalar2d(vp, 10.0, 1e-3, 100, 10.0, **kw, wavelet=np.zeros(99))
    with pytest.raises(ValueError):
        fomo.scalar2d(vp, 10.0, 1e-3, 100, 10.0, **{**kw, "fd_order": 12})
