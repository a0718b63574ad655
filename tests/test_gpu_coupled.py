"""GPU parity of coupled2d (src/equations/coupled2d/coupled2d.jl:53-260) against the CPU oracle.

The engine builds the medium on the device (Gaussian smoothing and derivatives of alpha = vp^2,
beta = vs^2), runs the velocity half-step of the interior with its position half-step in the
same pass and the frame's position half-step in a second launch; the oracle restates the
reference's two half-steps with in-place updates and the two-pass Higdon blend.  FP32 results
are bit-identical: traces, snapshots of P and S, and the final four fields.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DH, DT, F0 = 10.0, 1e-3, 15.0


def model(nx, nz, seed=7):
    """Layered vp / vs with a dipping vs interface (mode conversion) and a flat vp-only one."""
    rng = np.random.default_rng(seed)
    vp = np.full((nx, nz), 2500.0, np.float32)
    vs = np.full((nx, nz), 1200.0, np.float32)
    vp[:, nz // 3:] = 3000.0
    for i in range(nx):
        j = int(nz * 0.6 + 0.1 * i)
        vp[i, j:] = 3500.0
        vs[i, j:] = 1800.0
    vp += rng.uniform(-20, 20, vp.shape).astype(np.float32)
    return vp.astype(np.float32), vs.astype(np.float32)


def run_both(orc, vp, vs, nt, *, fd_order=8, nbc=20, sigma=3.0, snap=0, sx=None, sz=None, **kw):
    import fomo
    nx, nz = vp.shape
    sx = [nx // 2] if sx is None else sx
    sz = [10] if sz is None else sz
    rx = np.arange(1, nx + 1, 3)
    rz = np.full(len(rx), 6)
    print(f"coupled2d {nx}x{nz} nt={nt} fd{fd_order} nbc={nbc} sigma={sigma} snap={snap}", flush=True)
    r = fomo.coupled2d(vp, vs, DH, DT, nt, F0, sx=sx, sz=sz, rx=rx, rz=rz, nbc=nbc, fd_order=fd_order,
                       snap_interval=snap, smooth_sigma=sigma, **kw)
    w = fomo.ricker_wavelet(F0, DT, nt)
    o = orc.coupled2d(vp, vs, DH, DT, nt, sx=sx, sz=sz, rx=rx, rz=rz, wavelet=np.repeat(w[None], len(sx), 0),
                      nbc=nbc, fd_order=fd_order, snap_interval=snap, smooth_sigma=sigma,
                      v_ref_p=kw.get("v_ref_p"), v_ref_s=kw.get("v_ref_s"))
    return r, o


@pytest.mark.parametrize("fd_order", [2, 4, 6, 8, 10])
def test_coupled2d_bitwise_vs_oracle(gpu, orc, fd_order):
    vp, vs = model(110, 90)
    r, o = run_both(orc, vp, vs, 260, fd_order=fd_order)
    assert np.abs(r.seis_P).max() > 0 and np.abs(r.seis_S).max() > 0
    assert np.array_equal(r.seis_P, o["seis_P"]), np.abs(r.seis_P - o["seis_P"]).max()
    assert np.array_equal(r.seis_S, o["seis_S"]), np.abs(r.seis_S - o["seis_S"]).max()
    assert r.stats.kernel_time_s > 0


@pytest.mark.parametrize("sigma", [0.0, 1.3, 3.0, 6.0])
def test_coupled2d_smoothing_widths(gpu, orc, sigma):
    """Unsmoothed (sigma 0), a non-integer width, the reference's default and a wide kernel: the
    device medium equals the oracle's host restatement bit for bit."""
    vp, vs = model(96, 80, seed=3)
    r, o = run_both(orc, vp, vs, 150, sigma=sigma)
    assert np.array_equal(r.seis_P, o["seis_P"])
    assert np.array_equal(r.seis_S, o["seis_S"])


def test_coupled2d_snapshots_and_reference_velocities(gpu, orc):
    """Snapshots of P and S at every 40th step, two sources, explicit v_ref_p / v_ref_s."""
    vp, vs = model(100, 84, seed=5)
    r, o = run_both(orc, vp, vs, 200, snap=40, sx=[30, 70], sz=[8, 40], v_ref_p=2600.0, v_ref_s=1300.0)
    assert len(r.snaps_P) == len(r.snaps_S) == 5
    for a, b in zip(r.snaps_P, o["snaps_P"]):
        assert a.shape == (100, 84)
        assert np.array_equal(a, b)
    for a, b in zip(r.snaps_S, o["snaps_S"]):
        assert np.array_equal(a, b)
    assert np.array_equal(r.seis_P, o["seis_P"]) and np.array_equal(r.seis_S, o["seis_S"])


@pytest.mark.parametrize("nt", [1, 37, 64])
def test_coupled2d_final_state_graph_and_stream(gpu, orc, nt):
    """The four fields after nt steps (P, S, dP/dt, dS/dt) through a plan, with graph replays and
    direct launches, and every strip length of the work layout."""
    import fomo
    from fomo import _lib
    vp, vs = model(90, 70, seed=11)
    nx, nz = vp.shape
    sx, sz = [45], [20]
    rx = np.arange(1, nx + 1, 4)
    rz = np.full(len(rx), 5)
    w = fomo.ricker_wavelet(F0, DT, nt)
    o = orc.coupled2d(vp, vs, DH, DT, nt, sx=sx, sz=sz, rx=rx, rz=rz, wavelet=w[None], nbc=12, final_state=True)
    for graph in (True, False):
        for tile in (1, 2, 3):
            print(f"plan nt={nt} graph={graph} tile={tile}", flush=True)
            p = fomo.Plan(_lib.COUPLED2D, vp, None, DH, DT, nt, vs=vs, sx=sx, sz=sz, rx=rx, rz=rz, wavelet=w[None],
                          nbc=12, use_graph=graph, tile_cfg=tile)
            p.advance(nt)
            for k, name in enumerate(("P", "S", "dPdt", "dSdt")):
                assert np.array_equal(p.field(k), o["state"][name]), (name, graph, tile)
            sP, sS = p.seis(2)
            assert np.array_equal(sP, o["seis_P"]) and np.array_equal(sS, o["seis_S"])
            p.close()


def test_coupled2d_windows_equal_solo(gpu, orc):
    """Uneven windows of one model in one plan: each shot equals the oracle on its cropped model
    (medium, Higdon coefficients and v_ref_s per window)."""
    import fomo
    from fomo import _lib
    vp, vs = model(300, 70, seed=13)
    nt = 120
    w = fomo.ricker_wavelet(F0, DT, nt)
    wins = [(1, 96), (60, 150), (201, 100)]
    shots = []
    for k, (x0, wd) in enumerate(wins):
        rxl = np.arange(1, wd + 1, 2)
        shots.append(dict(x0=x0, nx=wd, sx=[wd // 2], sz=[6 + k], rx=rxl, rz=np.full(len(rxl), 4), wavelet=w[None]))
    plan = fomo.Plan(_lib.COUPLED2D, vp, None, DH, DT, nt, vs=vs, shots=shots, nbc=10)
    plan.advance(nt)
    for k, (x0, wd) in enumerate(wins):
        s = shots[k]
        crop = lambda a: np.asfortranarray(a[x0 - 1:x0 - 1 + wd])  # noqa: E731
        o = orc.coupled2d(crop(vp), crop(vs), DH, DT, nt, sx=s["sx"], sz=s["sz"], rx=s["rx"], rz=s["rz"],
                          wavelet=w[None], nbc=10)
        sP, sS = plan.seis(2, shot=k)
        assert np.array_equal(sP, o["seis_P"]), k
        assert np.array_equal(sS, o["seis_S"]), k
    plan.close()


def test_coupled2d_reference_demo_size(gpu, orc):
    """The reference's demo model (README.md:160-190: 2500/1200 m/s over 3500/1800 m/s at row 150,
    nbc 100, FD8, source at the top centre, receivers x = 1:2:nx at z = 10) on a 400 x 300 grid over
    1400 steps (the reflection from the interface reaches the receivers after about 1.2 s),
    bit-identical to the oracle."""
    nx, nz = 400, 300
    vp = np.full((nx, nz), 2500.0, np.float32)
    vs = np.full((nx, nz), 1200.0, np.float32)
    vp[:, 149:] = 3500.0
    vs[:, 149:] = 1800.0
    import fomo
    nt = 1400
    rx = np.arange(1, nx + 1, 2)
    rz = np.full(len(rx), 10)
    print("coupled2d demo 400x300 nbc 100", flush=True)
    r = fomo.coupled2d(vp, vs, DH, DT, nt, F0, sx=[nx // 2], sz=[10], rx=rx, rz=rz, nbc=100, fd_order=8)
    w = fomo.ricker_wavelet(F0, DT, nt)
    o = orc.coupled2d(vp, vs, DH, DT, nt, sx=[nx // 2], sz=[10], rx=rx, rz=rz, wavelet=w[None], nbc=100)
    assert np.array_equal(r.seis_P, o["seis_P"])
    assert np.array_equal(r.seis_S, o["seis_S"])
    # the S potential arises at the vs interface only: it is much weaker than P at the surface
    assert 0 < np.abs(r.seis_S).max() < np.abs(r.seis_P).max()


def test_coupled2d_refuses_what_the_reference_lacks(gpu):
    import fomo
    from fomo import _lib
    vp, vs = model(40, 30)
    w = fomo.ricker_wavelet(F0, DT, 10)
    with pytest.raises(_lib.FomoError):
        fomo.Plan(_lib.COUPLED2D, vp, None, DH, DT, 10, vs=vs, sx=[20], sz=[5], rx=[1], rz=[1], wavelet=w[None],
                  nbc=10, storage=1)
    with pytest.raises(_lib.ArgumentError):
        fomo.coupled2d(vp, np.zeros_like(vs), DH, DT, 10, F0, sx=[20], sz=[5], rx=[1], rz=[1], nbc=10)


def test_coupled2d_mirror_linearity_and_repeatability(gpu):
    """The reference's solver invariants (test/verify_patch.jl:198-229, :254-261;
    test/verify_det_graph.jl:152-296) for coupled2d: a model and source symmetric about the middle
    column give mirrored traces to 1e-5 -- the P potential even, the S potential odd (a mirror
    reverses the sense of rotation, so the shear potential changes sign); two sources give the sum
    of their solo runs to 1e-5; a repeated run gives the same bits."""
    import fomo
    nx, nz, nt = 121, 90, 300
    vp = np.full((nx, nz), 2600.0, np.float32)
    vs = np.full((nx, nz), 1300.0, np.float32)
    vp[:, 45:] = 3300.0
    vs[:, 45:] = 1900.0
    rx = np.arange(1, nx + 1)
    rz = np.full(nx, 40)  # five cells above the interface: the converted S is 3% of P there
    kw = dict(rx=rx, rz=rz, nbc=20, fd_order=8)
    r = fomo.coupled2d(vp, vs, DH, DT, nt, F0, sx=[61], sz=[12], **kw)
    assert np.abs(r.seis_S).max() > 0.01 * np.abs(r.seis_P).max()
    for key, parity in (("seis_P", 1.0), ("seis_S", -1.0)):
        a = getattr(r, key)
        scale = np.abs(a).max()
        assert scale > 0
        assert np.abs(a - parity * a[::-1]).max() <= 1e-5 * scale, key
    r2 = fomo.coupled2d(vp, vs, DH, DT, nt, F0, sx=[61], sz=[12], **kw)
    assert np.array_equal(r.seis_P, r2.seis_P) and np.array_equal(r.seis_S, r2.seis_S)
    a = fomo.coupled2d(vp, vs, DH, DT, nt, F0, sx=[30], sz=[12], **kw)
    b = fomo.coupled2d(vp, vs, DH, DT, nt, F0, sx=[90], sz=[30], **kw)
    ab = fomo.coupled2d(vp, vs, DH, DT, nt, F0, sx=[30, 90], sz=[12, 30], **kw)
    for key in ("seis_P", "seis_S"):
        s = getattr(a, key) + getattr(b, key)
        assert np.abs(getattr(ab, key) - s).max() <= 1e-5 * np.abs(s).max(), key
