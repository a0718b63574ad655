"""GPU learned boundary: the MLP kernels against the numpy oracle (forward, gradients, feature
gather on a live wavefield), and the learned plan's basic behaviour."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_mlp_kernels_match_oracle(gpu):
    import mlp
    from fomo.learned import mlp_eval
    rng = np.random.default_rng(5)
    w = rng.normal(0, 0.25, mlp.N_PARAMS).astype(np.float32)
    n = 203  # ragged: not a multiple of the 32-cell chunk
    x = rng.normal(0, 1, (n, mlp.N_IN)).astype(np.float32)
    x[:, 75] = rng.uniform(0, 1, n)
    ga = rng.normal(0, 1, (n, 2)).astype(np.float32)
    alpha, grad = mlp_eval(w, x, ga)
    wb = mlp.bf16_round(w)
    a_ref, cache = mlp.forward(wb, x.astype(np.float64))
    g_ref = mlp.backward(wb, cache, ga.astype(np.float64))
    assert np.allclose(alpha, a_ref, rtol=2e-5, atol=2e-6)
    scale = np.abs(g_ref).max()
    assert np.abs(grad - g_ref).max() <= 2e-5 * scale
    # determinism of the reduction
    alpha2, grad2 = mlp_eval(w, x, ga)
    assert np.array_equal(grad, grad2) and np.array_equal(alpha, alpha2)


def test_learned_inference_on_live_wavefield(gpu):
    """Blend-weight corrections computed by the engine (matrix-core inference) from the state at time n
    equal the oracle's BF16 inference path on features gathered from the same state."""
    import mlp
    import fomo
    import fomo._lib as L
    from fomo.api import Plan
    from fomo.learned import alphas, enable, initial_weights
    nx, nz, nt = 120, 90, 80
    vp = np.full((nx, nz), 2500.0, np.float32)
    vp[:, 45:] = 3000.0
    rho = np.full((nx, nz), 2100.0, np.float32)
    w = fomo.ricker_wavelet(20.0, 1e-3, nt)
    p = Plan(L.ACOUSTIC2D, vp, rho, 10.0, 1e-3, nt, sx=[60], sz=[20], rx=[1], rz=[1], wavelet=w[None], nbc=16,
             fd_order=8, use_graph=False)
    rng = np.random.default_rng(2)
    weights = initial_weights(4)
    weights[-66:] = rng.normal(0, 0.5, 66).astype(np.float32)  # non-zero output layer
    enable(p, weights)
    p.advance(60)
    p.sync()
    state = [p.field(k) for k in range(3)]  # p, vx, vz at time 60
    p.advance(1)
    cells, av, ap = alphas(p)
    x = mlp.features(state[0], state[1], state[2], cells, 19)
    a_ref = mlp.forward_bf16(mlp.bf16_round(weights), x)
    # BF16 activations: a hidden unit that lands on a rounding boundary may round the other way
    # (one BF16 ulp, 2^-8 relative); bound the worst cell and require the typical one to be exact
    # to FP32 accumulation order.
    for got, want in ((av, a_ref[:, 0]), (ap, a_ref[:, 1])):
        err = np.abs(got - want)
        assert err.max() < 2e-3 and np.median(err) < 1e-5, (err.max(), np.median(err))
    a32, _ = mlp.forward(mlp.bf16_round(weights), x)
    assert np.abs(av - a32[:, 0]).max() < 5e-3  # BF16 inference stays close to the FP32 network
    p.close()


def test_zero_correction_is_plain_habc_bitwise(gpu):
    """With the initial weights (zero output layer) the learned boundary reproduces the plain
    HABC bit for bit: clamp(w + 0, 0, 1) = w."""
    import fomo
    import fomo._lib as L
    from fomo.api import Plan
    from fomo.learned import enable, initial_weights
    nx, nz, nt = 200, 140, 300
    vp = np.full((nx, nz), 2000.0, np.float32)
    vp[:, 70:] = 2600.0
    rho = np.full((nx, nz), 2000.0, np.float32)
    w = fomo.ricker_wavelet(15.0, 8e-4, nt)
    rx = np.arange(1, nx + 1, 3)
    out = []
    for learned in (False, True):
        p = Plan(L.ACOUSTIC2D, vp, rho, 10.0, 8e-4, nt, sx=[100], sz=[30], rx=rx, rz=np.full(len(rx), 2),
                 wavelet=w[None], nbc=16, fd_order=8)
        if learned:
            enable(p, initial_weights(7))
        p.advance(nt)
        out.append(p.seis(3) + [p.field(k) for k in range(3)])
        p.close()
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_learned_graph_equals_stream(gpu, storage):
    """The inference launch runs on the plan's stream before the frame tiles that read its blend
    weights, inside the captured graph as in direct launches.  Graph replays and direct launches
    give the same bits, for step counts ending inside and at the end of a replay, and the trained
    network changes the result (so the corrections are read)."""
    import fomo
    import fomo._lib as L
    from fomo.api import Plan
    from fomo.learned import enable, load_weights
    nx, nz = 220, 130
    vp = np.full((nx, nz), 2200.0, np.float32)
    vp[:, 60:] = 3000.0
    rho = np.full((nx, nz), 2000.0, np.float32)
    rx = np.arange(1, nx + 1, 2)
    for nt in (37, 64):
        w = fomo.ricker_wavelet(15.0, 8e-4, nt)
        out = []
        for graph, learned in ((True, True), (False, True), (True, False)):
            p = Plan(L.ACOUSTIC2D, vp, rho, 10.0, 8e-4, nt, sx=[110], sz=[8], rx=rx, rz=np.full(len(rx), 4),
                     wavelet=w[None], nbc=16, fd_order=8, use_graph=graph, storage=storage)
            try:
                if learned:
                    enable(p, load_weights())
                p.advance(nt)
                out.append(p.seis(3) + [p.field(k) for k in range(3)])
            finally:
                p.close()
        for a, b in zip(out[0], out[1]):
            assert np.array_equal(a, b)
        assert any(not np.array_equal(a, b) for a, b in zip(out[0], out[2]))


def test_shipped_network_reduces_reflection(gpu):
    """The shipped weights, in the reference's absorption experiment (test/verify_patch.jl:76-91, its
    model, source and receivers; FD8): at nbc = 16 the residual against the boundary-free reference
    is <= -39.5 dB and at least 2 dB below the plain HABC of the same width (-37.6 dB); the API path
    acoustic2d(boundary="learned") gives the same traces as the plan it is built on."""
    import fomo
    from fomo import learned
    vp, rho, sx, sz, rx, rz = learned.patch_geometry()
    w = learned.load_weights()
    res = learned.reflection_residual(vp, rho, sx=sx, sz=sz, rx=rx, rz=rz, nt=1500, L=270,
                                      boundaries=[("habc16", 16, None), ("learned16", 16, w)], judge="none")
    print({k: round(v["reflection_db"], 2) for k, v in res.items()})
    assert res["learned16"]["reflection_db"] <= -39.5
    assert res["learned16"]["reflection_db"] <= res["habc16"]["reflection_db"] - 2.0
    r = fomo.acoustic2d(vp, rho, 10.0, 1e-3, 300, 10.0, sx=[sx], sz=[sz], rx=rx, rz=rz, nbc=16, boundary="learned")
    from fomo.api import Plan
    import fomo._lib as L
    p = Plan(L.ACOUSTIC2D, vp, rho, 10.0, 1e-3, 300, sx=[sx], sz=[sz], rx=rx, rz=rz,
             wavelet=fomo.ricker_wavelet(10.0, 1e-3, 300)[None], nbc=16)
    learned.enable(p, w)
    p.advance(300)
    seis = p.seis(3)
    p.close()
    assert np.array_equal(r.seis_p, seis[0]) and np.array_equal(r.seis_vz, seis[2])
