"""CPU pins of the learned-boundary MLP oracle (finite differences) and of the host-side
initialisation (a fresh network outputs zero correction to the distance-ramp Higdon weights)."""
import numpy as np


def test_mlp_oracle_gradient_matches_finite_differences():
    import mlp
    rng = np.random.default_rng(0)
    w = rng.normal(0, 0.3, mlp.N_PARAMS)
    x = rng.normal(0, 1, (7, mlp.N_IN))
    ga = rng.normal(0, 1, (7, 2))
    alpha, cache = mlp.forward(w, x)
    g = mlp.backward(w, cache, ga)

    def L(wv):
        return float((mlp.forward(wv, x)[0] * ga).sum())

    idx = rng.choice(mlp.N_PARAMS, 60, replace=False)
    for k in idx:
        e = np.zeros_like(w)
        e[k] = 1e-6
        fd = (L(w + e) - L(w - e)) / 2e-6
        assert abs(fd - g[k]) <= 1e-6 + 1e-5 * abs(fd), (k, fd, g[k])


def test_features_are_scale_invariant_and_clamped():
    import mlp
    rng = np.random.default_rng(1)
    nx, nz = 12, 9
    p, vx, vz = (rng.normal(size=(nx, nz)) for _ in range(3))
    cells = np.array([0, 1 + 1 * nx, 5 + 4 * nx, nx * nz - 1])
    f1 = mlp.features(p, vx, vz, cells, 3)
    f2 = mlp.features(8.0 * p, 0.125 * vx, 0.125 * vz, cells, 3)
    assert np.allclose(f1, f2)
    assert np.abs(f1[:, :75]).max() <= 1.0
    # corner cell: clamped patch repeats the corner value
    assert np.allclose(f1[0, :25].reshape(5, 5)[:3, :3], f1[0, 12])
    assert f1[2, 75] == min(5, nx - 1 - 5, 4, nz - 1 - 4) / 3


def test_initial_network_outputs_zero_correction():
    """The initial weights give corrections of exactly 0 (the plain distance-ramp Higdon boundary,
    habc.jl:30-34); the corrections stay within +-0.25 for any weights."""
    import mlp
    from fomo.learned import initial_weights
    rng = np.random.default_rng(2)
    x = np.concatenate([rng.normal(size=(50, 75)), rng.uniform(0, 1, (50, 1))], axis=1)
    corr, _ = mlp.forward(mlp.bf16_round(initial_weights(3)), x)
    assert np.all(corr == 0.0)
    corr, _ = mlp.forward(rng.normal(0, 3, mlp.N_PARAMS), x)
    assert np.abs(corr).max() <= mlp.CORR_MAX
