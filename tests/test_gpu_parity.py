"""GPU parity: the HIP path (through the C-ABI) against the CPU oracle.

The FP32 oracle restates the reference's Float32 updates in the reference's evaluation order
and the engine is compiled with -ffp-contract=off, so the bar here is bit-identity for
seismograms and final wavefields (stronger than the per-trace relative L2 <= 1e-5 the target
asks for, which is also asserted).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TRACE_TOL = 1e-5  # per-trace relative L2, FP32


def trace_rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.linalg.norm(b, axis=1)
    num = np.linalg.norm(a - b, axis=1)
    return np.where(den > 0, num / np.where(den > 0, den, 1), num)


def assert_parity(got, want):
    assert got.shape == want.shape
    assert np.all(np.isfinite(got))
    assert trace_rel_l2(got, want).max() <= TRACE_TOL
    assert np.array_equal(got, want), f"max |diff| = {np.abs(got.astype(np.float64) - want).max()}"


def small_model(nx, nz, seed=3, vacuum_top=True):
    from fomo.models import layered_2d
    vp, vs, rho = layered_2d(nx, nz, seed=seed, n_interfaces=4, jitter=3.0, vacuum_top=vacuum_top)
    return vp, vs, rho


def geometry(nx, nz):
    sx = [nx // 2, nx // 3]
    sz = [5, 9]
    rx = np.concatenate([np.arange(1, nx + 1, 3), [1, nx, nx // 2]])
    rz = np.concatenate([np.full(len(np.arange(1, nx + 1, 3)), 4), [nz, 1, nz // 2]])
    return sx, sz, rx, rz


@pytest.mark.parametrize("fd_order", [2, 4, 6, 8, 10])
def test_acoustic2d_bitwise_vs_oracle(gpu, orc, fd_order):
    import fomo
    nx, nz, nt, nbc = 97, 71, 260, 12
    vp, _, rho = small_model(nx, nz)
    sx, sz, rx, rz = geometry(nx, nz)
    dt, dh, f0 = 1e-3, 10.0, 15.0
    w = fomo.ricker_wavelet(f0, dt, nt)
    res = fomo.acoustic2d(vp, rho, dh, dt, nt, f0, sx=sx, sz=sz, rx=rx, rz=rz, nbc=nbc, fd_order=fd_order)
    ref = orc.acoustic2d(vp, rho, dh, dt, nt, sx=sx, sz=sz, rx=rx, rz=rz, wavelet=np.stack([w, w]), nbc=nbc,
                         fd_order=fd_order)
    for k in ("seis_p", "seis_vx", "seis_vz"):
        assert_parity(getattr(res, k), ref[k])
    assert res.stats.kernel_time_s > 0


@pytest.mark.parametrize("fd_order,source_type", [(2, "explosive"), (4, "fz"), (8, "explosive"), (8, "fz"),
                                                  (10, "explosive")])
def test_elastic2d_bitwise_vs_oracle(gpu, orc, fd_order, source_type):
    import fomo
    nx, nz, nt, nbc = 89, 77, 240, 10
    vp, vs, rho = small_model(nx, nz, seed=5)
    sx, sz, rx, rz = geometry(nx, nz)
    dt, dh, f0 = 5e-4, 5.0, 20.0
    w = fomo.ricker_wavelet(f0, dt, nt)
    res = fomo.elastic2d(vp, vs, rho, dh, dt, nt, f0, sx=sx, sz=sz, rx=rx, rz=rz, nbc=nbc, fd_order=fd_order,
                         source_type=source_type)
    ref = orc.elastic2d(vp, vs, rho, dh, dt, nt, sx=sx, sz=sz, rx=rx, rz=rz, wavelet=np.stack([w, w]), nbc=nbc,
                        fd_order=fd_order, source_type=1 if source_type == "fz" else 0)
    for k in ("seis_vx", "seis_vz"):
        assert_parity(getattr(res, k), ref[k])


def _plan_states(eq, vp, vs, rho, nt, kw, use_graph):
    from fomo.api import Plan
    import fomo._lib as L
    p = Plan(eq, vp, rho, 10.0, 1e-3, nt, vs=vs, use_graph=use_graph, **kw)
    p.advance(nt)
    p.sync()
    nf = 3 if eq == L.ACOUSTIC2D else 5
    fields = [p.field(k) for k in range(nf)]
    seis = p.seis(3 if eq == L.ACOUSTIC2D else 2)
    p.close()
    return fields, seis


@pytest.mark.parametrize("eq", [1, 2])
def test_final_wavefields_bitwise_and_graph_equals_stream(gpu, orc, eq):
    """Whole padded state after nt steps equals the oracle bitwise; the hipGraph replay equals
    direct launches bitwise (verify_det_graph.jl:261-296); reruns are bitwise reproducible
    (verify_det_graph.jl:152-185).  nt is odd so the tail step outside the graph is covered."""
    import fomo
    nx, nz, nt = 70, 66, 151
    vp, vs, rho = small_model(nx, nz, seed=11)
    sx, sz, rx, rz = geometry(nx, nz)
    w = fomo.ricker_wavelet(15.0, 1e-3, nt)
    kw = dict(sx=sx, sz=sz, rx=rx, rz=rz, wavelet=np.stack([w, w]), nbc=8, fd_order=8)
    g1, s1 = _plan_states(eq, vp, vs if eq == 2 else None, rho, nt, kw, True)
    g2, s2 = _plan_states(eq, vp, vs if eq == 2 else None, rho, nt, kw, False)
    g3, s3 = _plan_states(eq, vp, vs if eq == 2 else None, rho, nt, kw, True)
    for a, b, c in zip(g1 + s1, g2 + s2, g3 + s3):
        assert np.array_equal(a, b) and np.array_equal(a, c)
    if eq == 1:
        ref = orc.acoustic2d(vp, rho, 10.0, 1e-3, nt, final_state=True, **kw)
        names = ("p", "vx", "vz")
    else:
        ref = orc.elastic2d(vp, vs, rho, 10.0, 1e-3, nt, final_state=True, **kw)
        names = ("vx", "vz", "txx", "tzz", "txz")
    for k, name in enumerate(names):
        assert np.array_equal(g1[k], ref["state"][name]), name


def test_snapshots_match_oracle(gpu, orc):
    import fomo
    nx, nz, nt = 80, 60, 120
    vp, vs, rho = small_model(nx, nz, seed=2)
    sx, sz, rx, rz = geometry(nx, nz)
    w = fomo.ricker_wavelet(15.0, 1e-3, nt)
    kw = dict(sx=sx, sz=sz, rx=rx, rz=rz, nbc=10, fd_order=8)
    res = fomo.acoustic2d(vp, rho, 10.0, 1e-3, nt, 15.0, snap_interval=25, **kw)
    ref = orc.acoustic2d(vp, rho, 10.0, 1e-3, nt, snap_interval=25, wavelet=np.stack([w, w]), **kw)
    assert len(res.snaps) == nt // 25 == len(ref["snaps"])
    for a, b in zip(res.snaps, ref["snaps"]):
        assert a.shape == (nx, nz)
        assert np.array_equal(a, b)
    e = fomo.elastic2d(vp, vs, rho, 10.0, 1e-3, nt, 15.0, snap_interval=40, **kw)
    er = orc.elastic2d(vp, vs, rho, 10.0, 1e-3, nt, snap_interval=40, wavelet=np.stack([w, w]), **kw)
    for a, b in zip(e.snaps, er["snaps"]):
        assert np.array_equal(a, b)


def test_snapshots_after_partial_advance(gpu, orc):
    """A plan whose tile shape is autotuned (steps run before the reset that starts the user's
    run): after a partial advance, the slots reached equal the oracle and the others are zero."""
    import fomo
    import fomo._lib as L
    from fomo.api import Plan
    nx, nz, nt, si = 72, 54, 100, 10
    vp, vs, rho = small_model(nx, nz, seed=6)
    sx, sz, rx, rz = geometry(nx, nz)
    w = fomo.ricker_wavelet(15.0, 1e-3, nt)
    ref = orc.acoustic2d(vp, rho, 10.0, 1e-3, nt, snap_interval=si, wavelet=np.stack([w, w]), sx=sx, sz=sz, rx=rx,
                         rz=rz, nbc=8, fd_order=8)
    p = Plan(L.ACOUSTIC2D, vp, rho, 10.0, 1e-3, nt, sx=sx, sz=sz, rx=rx, rz=rz, wavelet=np.stack([w, w]), nbc=8,
             fd_order=8, tile_cfg=0, snap_interval=si)
    try:
        p.advance(35)
        got = p.snaps()
        assert got.shape == (nt // si, nx * nz)
        for k in range(3):
            assert np.array_equal(got[k].reshape(nz, nx).T, ref["snaps"][k])
        assert not np.any(got[3:])
        p.reset()
        p.advance(9)
        assert not np.any(p.snaps())
    finally:
        p.close()


@pytest.mark.parametrize("si,nt", [(1, 37), (7, 100), (16, 64)])
def test_snapshots_inside_graph_replays(gpu, orc, si, nt):
    """Snapshots are written by the sweeps on the device, also inside 16-step graph replays and
    their odd tails: every interval (1, one not dividing 16, 16) matches the oracle bitwise, with
    and without graphs."""
    import fomo
    nx, nz = 70, 50
    vp, vs, rho = small_model(nx, nz, seed=4)
    sx, sz, rx, rz = geometry(nx, nz)
    w = fomo.ricker_wavelet(15.0, 1e-3, nt)
    kw = dict(sx=sx, sz=sz, rx=rx, rz=rz, nbc=8, fd_order=8)
    ref = orc.acoustic2d(vp, rho, 10.0, 1e-3, nt, snap_interval=si, wavelet=np.stack([w, w]), **kw)
    for graph in (True, False):
        res = fomo.acoustic2d(vp, rho, 10.0, 1e-3, nt, 15.0, snap_interval=si, use_cuda_graph=graph, **kw)
        assert len(res.snaps) == nt // si == len(ref["snaps"])
        for a, b in zip(res.snaps, ref["snaps"]):
            assert np.array_equal(a, b)
    er = orc.elastic2d(vp, vs, rho, 10.0, 1e-3, nt, snap_interval=si, wavelet=np.stack([w, w]), **kw)
    e = fomo.elastic2d(vp, vs, rho, 10.0, 1e-3, nt, 15.0, snap_interval=si, **kw)
    for a, b in zip(e.snaps, er["snaps"]):
        assert np.array_equal(a, b)


def test_edge_cases(gpu, orc):
    """Thin boundary (nbc=0), sources and receivers on the interior's extreme cells, no receivers,
    duplicate source cells, nt = 1 and an odd step count."""
    import fomo
    nx, nz, nt = 65, 40, 90
    vp, vs, rho = small_model(nx, nz, seed=9, vacuum_top=False)
    w = fomo.ricker_wavelet(15.0, 1e-3, nt)
    kw = dict(sx=[1, nx], sz=[nz, 1], rx=[1, nx, 1, nx], rz=[1, 1, nz, nz], nbc=0, fd_order=4)
    res = fomo.acoustic2d(vp, rho, 10.0, 1e-3, nt, 15.0, **kw)
    ref = orc.acoustic2d(vp, rho, 10.0, 1e-3, nt, wavelet=np.stack([w, w]), **kw)
    for k in ("seis_p", "seis_vx", "seis_vz"):
        assert np.array_equal(getattr(res, k), ref[k])
    r0 = fomo.acoustic2d(vp, rho, 10.0, 1e-3, nt, 15.0, sx=[30], sz=[20], rx=[], rz=[], nbc=5)
    assert r0.seis_p.shape == (0, nt)
    # duplicate source cells add in order
    kw2 = dict(sx=[30, 30], sz=[20, 20], rx=[31], rz=[20], nbc=5, fd_order=8)
    res = fomo.acoustic2d(vp, rho, 10.0, 1e-3, nt, 15.0, **kw2)
    ref = orc.acoustic2d(vp, rho, 10.0, 1e-3, nt, wavelet=np.stack([w, w]), **kw2)
    assert np.array_equal(res.seis_p, ref["seis_p"])
    # one step, and an odd step count (the two-step graph plus a direct tail launch)
    for n in (1, 7):
        wn = fomo.ricker_wavelet(15.0, 1e-3, n)
        kw3 = dict(sx=[30], sz=[20], rx=[30, 31], rz=[20, 21], nbc=5, fd_order=8)
        res = fomo.acoustic2d(vp, rho, 10.0, 1e-3, n, 15.0, **kw3)
        ref = orc.acoustic2d(vp, rho, 10.0, 1e-3, n, wavelet=wn[None], **kw3)
        assert np.array_equal(res.seis_p, ref["seis_p"]) and res.seis_p.shape == (2, n)


def test_custom_wavelet_equals_default_and_linearity(gpu):
    """verify_patch.jl:254-274 on the GPU path."""
    import fomo
    nx, nz, nt = 301, 121, 500
    vp, vs, rho = small_model(nx, nz, seed=4, vacuum_top=False)
    kw = dict(rx=np.arange(20, 280), rz=np.full(260, 41), nbc=30, fd_order=4)
    w = fomo.ricker_wavelet(10.0, 1e-3, nt)
    r0 = fomo.acoustic2d(vp, rho, 10.0, 1e-3, nt, 10.0, sx=[100], sz=[41], **kw)
    r1 = fomo.acoustic2d(vp, rho, 10.0, 1e-3, nt, 10.0, sx=[100], sz=[41], wavelet=w, **kw)
    assert np.array_equal(r0.seis_p, r1.seis_p)
    b = fomo.acoustic2d(vp, rho, 10.0, 1e-3, nt, 10.0, sx=[200], sz=[41], **kw)
    ab = fomo.acoustic2d(vp, rho, 10.0, 1e-3, nt, 10.0, sx=[100, 200], sz=[41, 41], **kw)
    r = np.linalg.norm(ab.seis_p - (r0.seis_p + b.seis_p)) / np.linalg.norm(ab.seis_p)
    assert r < 1e-5


def test_mirror_symmetry_gpu(gpu):
    """verify_patch.jl:198-229 at the reference's size on the GPU path (< 1e-5)."""
    import fomo
    from test_oracle_pins import make_model, rel_asym_vx, rel_sym
    vp, vs, rho = make_model(601, 201)
    rx = np.arange(150, 453)
    kw = dict(sx=[301], sz=[41], rx=rx, rz=np.full(len(rx), 41), nbc=30, fd_order=4)
    r = fomo.acoustic2d(vp, rho, 10.0, 1e-3, 1200, 10.0, **kw)
    assert rel_sym(r.seis_p) < 1e-5 and rel_sym(r.seis_vz) < 1e-5 and rel_asym_vx(r.seis_vx) < 1e-5
    e = fomo.elastic2d(vp, vs, rho, 10.0, 1e-3, 1200, 10.0, **kw)
    assert rel_sym(e.seis_vz) < 1e-5 and rel_asym_vx(e.seis_vx) < 1e-5


def test_acoustic_400x300_full_bitwise(gpu, orc):
    """Plain acoustic2d 400x300, dh 10, dt 1 ms, nt 2000, fd8, nbc 100, vp 2500 / rho 2000, vacuum
    top row, Ricker 15 Hz at (200,10), receivers x = 1:2:400 at z = 10 (p, vx, vz)."""
    import fomo
    nx, nz, nt = 400, 300, 2000
    vp = np.full((nx, nz), 2500.0, np.float32)
    rho = np.full((nx, nz), 2000.0, np.float32)
    vp[:, 0] = 0.0
    rho[:, 0] = 0.0
    rx = np.arange(1, nx + 1, 2)
    rz = np.full(len(rx), 10)
    res = fomo.acoustic2d(vp, rho, 10.0, 1e-3, nt, 15.0, sx=[200], sz=[10], rx=rx, rz=rz, nbc=100, fd_order=8,
                          v_ref=2500.0)
    w = fomo.ricker_wavelet(15.0, 1e-3, nt)
    ref = orc.acoustic2d(vp, rho, 10.0, 1e-3, nt, sx=[200], sz=[10], rx=rx, rz=rz, wavelet=w[None], nbc=100,
                         fd_order=8, v_ref=2500.0)
    for k in ("seis_p", "seis_vx", "seis_vz"):
        assert_parity(getattr(res, k), ref[k])


def test_elastic_2000x800_full_run_bitwise(gpu, orc):
    """Plain elastic2d, the bench workload in full: 2000x800 grid (dh 5, dt 0.4 ms, fd8, nbc 50,
    seeded layered model, vacuum top, vertical-force Ricker 20 Hz at (1000,5), receivers at every
    cell of z=5), all 4000 steps bitwise against the oracle (about 15 s of oracle time on 16 host
    threads); the traces are finite and decay by the end of the record."""
    import fomo
    from fomo.models import layered_2d
    nx, nz, nt = 2000, 800, 4000
    vp, vs, rho = layered_2d(nx, nz, seed=20260101)
    rx = np.arange(1, nx + 1)
    rz = np.full(nx, 5)
    kw = dict(sx=[1000], sz=[5], rx=rx, rz=rz, nbc=50, fd_order=8)
    res = fomo.elastic2d(vp, vs, rho, 5.0, 4e-4, nt, 20.0, source_type="fz", **kw)
    w = fomo.ricker_wavelet(20.0, 4e-4, nt)
    ref = orc.elastic2d(vp, vs, rho, 5.0, 4e-4, nt, wavelet=w[None], source_type=1, **kw)
    for k in ("seis_vx", "seis_vz"):
        assert_parity(getattr(res, k), ref[k])
    assert np.all(np.isfinite(res.seis_vz)) and np.all(np.isfinite(res.seis_vx))
    peak = np.abs(res.seis_vz).max()
    assert np.abs(res.seis_vz[:, -200:]).max() < peak


# ---------------------------------------------------------------------------------------------
# boundary = :sponge (src/boundary/sponge.jl, acoustic2d.jl:213-228, elastic2d.jl:180-195)

@pytest.mark.parametrize("fd_order", [4, 8])
def test_sponge_acoustic2d_bitwise_vs_oracle(gpu, orc, fd_order):
    import fomo
    nx, nz, nt, nbc = 97, 71, 260, 12
    vp, _, rho = small_model(nx, nz)
    sx, sz, rx, rz = geometry(nx, nz)
    w = fomo.ricker_wavelet(15.0, 1e-3, nt)
    res = fomo.acoustic2d(vp, rho, 10.0, 1e-3, nt, 15.0, sx=sx, sz=sz, rx=rx, rz=rz, nbc=nbc, fd_order=fd_order,
                          boundary="sponge", snap_interval=100)
    ref = orc.acoustic2d(vp, rho, 10.0, 1e-3, nt, sx=sx, sz=sz, rx=rx, rz=rz, wavelet=np.stack([w, w]), nbc=nbc,
                         fd_order=fd_order, sponge_factor=0.015, snap_interval=100)
    for k in ("seis_p", "seis_vx", "seis_vz"):
        assert_parity(getattr(res, k), ref[k])
    assert np.array_equal(np.stack(res.snaps), ref["snaps"])
    # a different factor is honoured (and changes the result)
    res2 = fomo.acoustic2d(vp, rho, 10.0, 1e-3, nt, 15.0, sx=sx, sz=sz, rx=rx, rz=rz, nbc=nbc, fd_order=fd_order,
                           boundary="sponge", sponge_factor=0.025)
    ref2 = orc.acoustic2d(vp, rho, 10.0, 1e-3, nt, sx=sx, sz=sz, rx=rx, rz=rz, wavelet=np.stack([w, w]), nbc=nbc,
                          fd_order=fd_order, sponge_factor=0.025)
    assert_parity(res2.seis_p, ref2["seis_p"])
    assert not np.array_equal(res2.seis_p, res.seis_p)


@pytest.mark.parametrize("source_type", ["explosive", "fz"])
def test_sponge_elastic2d_bitwise_vs_oracle(gpu, orc, source_type):
    import fomo
    nx, nz, nt, nbc = 89, 77, 240, 10
    vp, vs, rho = small_model(nx, nz, seed=5)
    sx, sz, rx, rz = geometry(nx, nz)
    w = fomo.ricker_wavelet(20.0, 5e-4, nt)
    res = fomo.elastic2d(vp, vs, rho, 5.0, 5e-4, nt, 20.0, sx=sx, sz=sz, rx=rx, rz=rz, nbc=nbc, fd_order=8,
                         source_type=source_type, boundary="sponge")
    ref = orc.elastic2d(vp, vs, rho, 5.0, 5e-4, nt, sx=sx, sz=sz, rx=rx, rz=rz, wavelet=np.stack([w, w]), nbc=nbc,
                        fd_order=8, source_type=1 if source_type == "fz" else 0, sponge_factor=0.015)
    for k in ("seis_vx", "seis_vz"):
        assert_parity(getattr(res, k), ref[k])


def test_sponge_mirror_symmetry_and_calibration_gpu(gpu):
    """verify_patch.jl:57-73 with boundary=:sponge (causally shielded mirror geometry, < 1e-5) and the
    reference's calibrated sponge residual 2.87e-2 (verify_patch.jl:9-11) on the GPU path."""
    import fomo
    from fomo.learned import edge_extend
    from test_oracle_pins import make_model, rel_asym_vx, rel_sym
    vp, vs, rho = make_model(601, 201)
    r = fomo.acoustic2d(vp, rho, 10.0, 1e-3, 1200, 10.0, sx=[301], sz=[41], rx=np.arange(150, 453),
                        rz=np.full(303, 41), nbc=30, fd_order=4, boundary="sponge")
    assert rel_sym(r.seis_p) < 1e-5 and rel_sym(r.seis_vz) < 1e-5 and rel_asym_vx(r.seis_vx) < 1e-5
    nxr, nzr, ntr, L = 301, 201, 1500, 270
    vp2, _, rho2 = make_model(nxr, nzr)
    rx2 = np.arange(1, nxr + 1)
    h = fomo.acoustic2d(vp2, rho2, 10.0, 1e-3, ntr, 10.0, sx=[151], sz=[41], rx=rx2, rz=np.full(nxr, 41), nbc=20,
                        fd_order=4, boundary="sponge")
    ref = fomo.acoustic2d(edge_extend(vp2, L), edge_extend(rho2, L), 10.0, 1e-3, ntr, 10.0, sx=[151 + L],
                          sz=[41 + L], rx=rx2 + L, rz=np.full(nxr, 41 + L), nbc=20, fd_order=4, boundary="sponge")
    resid = np.linalg.norm(h.seis_p - ref.seis_p) / np.linalg.norm(ref.seis_p)
    assert abs(resid - 2.87e-2) < 0.005e-2, resid
