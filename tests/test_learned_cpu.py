"""CPU pins of the learned-boundary MLP oracle (closed forms, finite differences) and of the host-side
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


def _zero_net():
    import mlp
    return {name: np.zeros(shape) for name, shape in mlp.SIZES}


def test_mlp_forward_closed_form_depth_route():
    """Weights that carry the depth feature (input 75) through hidden unit 3 of layer 1 and unit 6 of
    layer 2 into both outputs with gains +a and -a: corr = 0.25 * tanh(+-a * depth) exactly.  The
    route uses off-diagonal entries (W2[6][3]), so a transposed W2 sends nothing and fails."""
    import mlp
    a = 1.7
    d = _zero_net()
    d["W1"][3, 75] = 1.0
    d["W2"][6, 3] = 1.0
    d["W3"][0, 6], d["W3"][1, 6] = a, -a
    rng = np.random.default_rng(5)
    x = np.concatenate([rng.normal(size=(40, 75)), rng.uniform(0, 1, (40, 1))], axis=1)
    corr, _ = mlp.forward(mlp.pack(d), x)
    want = 0.25 * np.tanh(a * x[:, 75])
    assert np.array_equal(corr[:, 0], want) and np.array_equal(corr[:, 1], -want)
    dt = _zero_net()
    dt.update(W1=d["W1"], W2=d["W2"].T.copy(), W3=d["W3"])
    assert not np.any(mlp.forward(mlp.pack(dt), x)[0])


def test_mlp_forward_closed_form_biases_and_relu():
    """Zero weight matrices: the output is 0.25 * tanh(b3 + W3 @ relu(b2 + W2 @ relu(b1))), a constant
    over the inputs, computed here by hand for chosen biases: a negative b1 unit is cut by the ReLU,
    a positive one passes."""
    import mlp
    d = _zero_net()
    d["b1"][2], d["b1"][9] = 0.5, -4.0
    d["W2"][1, 2], d["W2"][1, 9] = 2.0, 3.0     # unit 9 is negative before the ReLU: contributes 0
    d["b2"][1], d["b2"][4] = -0.25, -1.0         # h2[1] = relu(2 * 0.5 - 0.25) = 0.75, h2[4] = 0
    d["W3"][0, 1], d["W3"][1, 4] = 0.8, 5.0
    d["b3"][:] = (0.1, -0.3)
    x = np.random.default_rng(6).normal(size=(9, mlp.N_IN))
    corr, _ = mlp.forward(mlp.pack(d), x)
    want = 0.25 * np.tanh(np.array([0.1 + 0.8 * 0.75, -0.3]))
    assert np.allclose(corr, want[None, :], rtol=1e-15, atol=0)


def test_features_patch_layout_from_a_spike():
    """A unit spike of p at cell (i0, j0): the patch of cell (i, j) holds it at position
    5 * (j0 - j + 2) + (i0 - i + 2) (x fastest within the patch), scaled to exactly 1; vx and vz
    spikes land at the same position in features 25..49 and 50..74."""
    import mlp
    nx, nz, i0, j0 = 11, 10, 5, 4
    p = np.zeros((nx, nz))
    vx = np.zeros((nx, nz))
    vz = np.zeros((nx, nz))
    p[i0, j0] = -3.0
    vx[i0, j0] = 2.0
    vz[i0, j0] = 0.5
    for di in range(-2, 3):
        for dj in range(-2, 3):
            i, j = i0 + di, j0 + dj
            f = mlp.features(p, vx, vz, np.array([i + j * nx]), 4)[0]
            k = 5 * (j0 - j + 2) + (i0 - i + 2)
            assert f[k] == -1.0 and np.count_nonzero(f[:25]) == 1
            assert f[25 + k] == 1.0 and np.count_nonzero(f[25:50]) == 1
            assert f[50 + k] == 0.25 and np.count_nonzero(f[50:75]) == 1
            assert f[75] == min(i, nx - 1 - i, j, nz - 1 - j) / 4


def test_bf16_round_against_nearest_search():
    """Round-to-nearest-even to BFloat16 checked against a search over the two BFloat16 neighbours
    of each value (exact in FP64), on ties, carries into the exponent and random values."""
    import mlp
    one = np.float32(1.0)
    ties = np.array([1 + 2.0 ** -8, 1 + 3 * 2.0 ** -8, 2 - 2.0 ** -8, -(1 + 2.0 ** -8), 1 + 2.0 ** -8 + 2.0 ** -20],
                    np.float32)
    assert list(mlp.bf16_round(ties)) == [1.0, 1 + 2.0 ** -6, 2.0, -1.0, 1 + 2.0 ** -7]
    assert mlp.bf16_round(one) == 1.0
    rng = np.random.default_rng(7)
    v = np.concatenate([rng.normal(0, 10, 2000), rng.uniform(-1e-3, 1e-3, 2000)]).astype(np.float32)
    bits = v.view(np.uint32)
    lo = (bits & np.uint32(0xFFFF0000)).view(np.float32).astype(np.float64)       # toward zero
    hi = ((bits & np.uint32(0xFFFF0000)) + np.uint32(0x10000)).view(np.float32).astype(np.float64)
    x = v.astype(np.float64)
    dlo, dhi = np.abs(x - lo), np.abs(hi - x)
    lo_even = ((bits >> 16) & 1) == 0
    want = np.where(dlo < dhi, lo, np.where(dhi < dlo, hi, np.where(lo_even, lo, hi)))
    assert np.array_equal(mlp.bf16_round(v), want)


def test_forward_bf16_rounds_features_and_hidden_units():
    """With BFloat16-exact weights, forward_bf16 equals forward on BFloat16-exact inputs, and on the
    depth route above it sees depth 1 + 2^-8 as 1 (a tie, rounded to even) and 1 + 3 * 2^-8 as
    1 + 2^-6."""
    import mlp
    d = _zero_net()
    d["W1"][3, 75] = 1.0
    d["W2"][6, 3] = 1.0
    d["W3"][0, 6], d["W3"][1, 6] = 1.5, -0.5
    w = mlp.pack(d)
    x = np.zeros((3, mlp.N_IN))
    x[:, 75] = (0.75, 1 + 2.0 ** -8, 1 + 3 * 2.0 ** -8)
    got = mlp.forward_bf16(w, x)
    dep = np.array([0.75, 1.0, 1 + 2.0 ** -6])
    assert np.array_equal(got, 0.25 * np.tanh(np.stack([1.5 * dep, -0.5 * dep], 1)))
    assert np.array_equal(got[0], mlp.forward(w, x)[0][0])
