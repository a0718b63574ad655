"""Closed-form pins of the oracle's coupled P-S potential solver (coupled2d), independent of the GPU.

The oracle restates src/equations/coupled2d/{medium,update_fields,coupled2d}.jl and
src/utils/FD_centered.jl.  The reference holds no output of this solver (its examples only plot),
so each part is pinned against what the mathematics fixes:

  * the centred weights by their moment conditions (exactness on polynomials);
  * the smoothing by its impulse response: the normalised Gaussian exp(-k^2 / (2 sigma^2)) on
    |k| <= ceil(3 sigma), zero beyond (medium.jl:118-131);
  * the medium derivatives on a model whose vp, vs are linear (alpha, beta quadratic), where the
    centred differences are exact;
  * the "velocity" half-step against the coupled equations the reference states in README.md:363-371
    (Li et al. 2018), written here in vector form -- gradients, divergence-free curl of
    S = (0, S_y, 0), cross products -- and evaluated on oblique plane waves, whose centred
    differences have closed forms: sum_l d1_l (f(x+l) - f(x-l)) = 2 sum_l d1_l sin(l k) A cos(k.x+phi)
    and c0 f(x) + sum_l d2_l (f(x+l) + f(x-l)) = (c0 + 2 sum_l d2_l cos(l k)) f(x);
  * a whole homogeneous run against the leapfrog's dispersion relation (P travels with alpha,
    S with beta);
  * the solver's physics statement (README.md:372-375): with grad beta = 0 no S arises.

Planted mistakes (swapped alpha / beta, transposed or negated beta gradients, swapped Laplacians,
a wrong smoothing width) enter through the probes' inputs, where each would enter, and must fail.
"""
import numpy as np
import pytest

RNG_SEED = 20261205


def _moments_d1(d1):
    N = len(d1)
    return [sum(2.0 * float(d1[l]) * (l + 1) ** (2 * k - 1) for l in range(N)) for k in range(1, N + 1)]


@pytest.mark.parametrize("order", [2, 4, 6, 8, 10])
def test_centred_weights_are_exact_on_polynomials(orc, order):
    d1 = orc.centered_d1(order).astype(np.float64)
    c0, d2 = orc.centered_d2(order)
    d2 = d2.astype(np.float64)
    N = order // 2
    assert len(d1) == len(d2) == N
    # first derivative: sum_l d1_l ((l)^(2k-1) - (-l)^(2k-1)) = [k == 1]
    for k, m in enumerate(_moments_d1(d1), start=1):
        assert abs(m - (1.0 if k == 1 else 0.0)) < 2e-7 * max(1.0, (N ** (2 * k - 1))), (k, m)
    # second derivative: c0 + 2 sum d2 = 0; 2 sum d2_l l^(2k) = [k == 1] * 2
    assert abs(c0 + 2 * d2.sum()) < 1e-6
    for k in range(1, N + 1):
        m = 2.0 * sum(d2[l] * (l + 1) ** (2 * k) for l in range(N))
        assert abs(m - (2.0 if k == 1 else 0.0)) < 2e-6 * max(1.0, N ** (2 * k)), (k, m)


def test_product_library_weights_equal_the_oracle(orc):
    import fomo
    for order in (2, 4, 6, 8, 10):
        assert np.array_equal(fomo.get_centered_d1(order), orc.centered_d1(order))
        c0, d2 = fomo.get_centered_d2(order)
        oc0, od2 = orc.centered_d2(order)
        assert float(c0) == oc0 and np.array_equal(d2, od2)


def _fd_x(f, d1, h, N):
    """Centred first difference along axis 0, interior only (NaN elsewhere)."""
    g = np.full(f.shape, np.nan)
    n = f.shape[0]
    acc = np.zeros((n - 2 * N,) + f.shape[1:])
    for l in range(1, N + 1):
        acc += d1[l - 1] * (f[N + l:n - N + l] - f[N - l:n - N - l])
    g[N:n - N] = acc / h
    return g


@pytest.mark.parametrize("sigma", [1.0, 2.5, 3.0, 4.2])
def test_smoothing_impulse_response_is_the_gaussian(orc, sigma):
    """alpha = 1 with one cell of 4 (vp = 2 there): the smoothed alpha is 1 + 3 g(i) g(j), g the
    normalised taps, so its x-derivative is 3 g(j) (d/dx g)(i).  The oracle's derivative output is
    compared with that, built from the closed-form taps and a centred difference (pinned above)."""
    nbc, order, dh = 20, 8, 10.0
    nxi, nzi = 61, 57
    vp = np.ones((nxi, nzi), np.float32)
    ci, cj = 30, 28
    vp[ci, cj] = 2.0
    vs = np.ones_like(vp)
    med = orc.coupled_medium(vp, vs, dh, nbc, order, smooth_sigma=sigma, precision="f64")
    pad = nbc + order // 2
    R = int(np.ceil(3 * sigma))
    k = np.arange(-R, R + 1)
    g = np.exp(-0.5 * (k / sigma) ** 2)
    g /= g.sum()
    nx, nz = nxi + 2 * pad, nzi + 2 * pad
    sm = np.ones((nx, nz))
    I, J = ci + pad, cj + pad
    sm[I - R:I + R + 1, J - R:J + R + 1] += 3.0 * np.outer(g, g)
    d1 = orc.centered_d1(order).astype(np.float64)
    want = _fd_x(sm, d1, dh, order // 2)
    got = med["dalpha_dx"]
    box = (slice(I - R - 6, I + R + 7), slice(J - R - 6, J + R + 7))
    assert np.nanmax(np.abs(got[box] - want[box])) < 1e-12 * np.nanmax(np.abs(want[box]))
    # beta is constant: its derivatives vanish, its Laplacian up to the Float32 weights' c0 + 2 sum d2
    # (about 1e-7, over h^2)
    for name in ("dbeta_dx", "dbeta_dz"):
        assert np.abs(med[name]).max() < 1e-12
    assert np.abs(med["lap_beta"]).max() < 1e-6 / dh ** 2
    # a planted width (radius 2 sigma instead of 3 sigma) is visible at this tolerance
    R2 = int(np.ceil(2 * sigma))
    k2 = np.arange(-R2, R2 + 1)
    g2 = np.exp(-0.5 * (k2 / sigma) ** 2)
    g2 /= g2.sum()
    sm2 = np.ones((nx, nz))
    sm2[I - R2:I + R2 + 1, J - R2:J + R2 + 1] += 3.0 * np.outer(g2, g2)
    bad = _fd_x(sm2, d1, dh, order // 2)
    assert np.nanmax(np.abs(got[box] - bad[box])) > 1e-6 * np.nanmax(np.abs(want[box]))


@pytest.mark.parametrize("order", [2, 4, 8, 10])
def test_medium_derivatives_of_quadratics_are_exact(orc, order):
    """vp = 2000 + 3i + 5j and vs = 1000 - 2i + 7j (cell indices) make alpha, beta quadratic, so
    d alpha/dx = 6 vp / h, d alpha/dz = 10 vp / h, lap alpha = (18 + 50) / h^2 (and likewise for
    beta, up to the weights' zeroth moment below) wherever the stencils stay inside the unpadded
    model.  Unsmoothed medium, FP64 arithmetic; the weights are the reference's Float32 values,
    exact on polynomials to about 1e-7."""
    nbc, dh = 6, 5.0
    nxi, nzi = 40, 36
    i, j = np.meshgrid(np.arange(nxi), np.arange(nzi), indexing="ij")
    vp = (2000.0 + 3 * i + 5 * j).astype(np.float32)
    vs = (1000.0 - 2 * i + 7 * j).astype(np.float32)
    med = orc.coupled_medium(vp, vs, dh, nbc, order, smooth_sigma=0.0, precision="f64")
    pad, N = nbc + order // 2, order // 2
    inner = (slice(pad + N, pad + nxi - N), slice(pad + N, pad + nzi - N))
    V, S = vp.astype(np.float64)[N:nxi - N, N:nzi - N], vs.astype(np.float64)[N:nxi - N, N:nzi - N]
    # per direction, c0 f + sum_l d2_l (f(i+l) + f(i-l)) = (c0 + 2 sum d2) f + (sum_l d2_l l^2) f'' for a
    # quadratic f; with the Float32 weights c0 + 2 sum d2 is ~1e-7, not 0, and alpha is ~1e7
    c0, d2 = orc.centered_d2(order)
    d2 = d2.astype(np.float64)
    s0 = float(c0) + 2.0 * d2.sum()
    s2 = sum(d2[l] * (l + 1) ** 2 for l in range(len(d2)))
    lap = lambda f, fxx, fzz: (2.0 * s0 * f + s2 * (fxx + fzz)) / dh ** 2  # noqa: E731
    checks = {"alpha": V ** 2, "beta": S ** 2, "dalpha_dx": 6 * V / dh, "dalpha_dz": 10 * V / dh,
              "dbeta_dx": -4 * S / dh, "dbeta_dz": 14 * S / dh,
              "lap_alpha": lap(V ** 2, 18.0, 50.0), "lap_beta": lap(S ** 2, 8.0, 98.0)}
    for name, want in checks.items():
        got = med[name][inner]
        assert np.allclose(got, want, rtol=2e-7, atol=2e-7 * np.abs(want).max()), name


# ---------------------------------------------------------------------------------------------
# the velocity half-step against the coupled equations in vector form

def _curl(J):
    """curl of a vector field from its Jacobian J[i][j] = d F_i / d x_j (x, y, z)."""
    return np.stack([J[2][1] - J[1][2], J[0][2] - J[2][0], J[1][0] - J[0][1]])


def _wave(shape, k, A, phi):
    i, j = np.meshgrid(np.arange(shape[0], dtype=np.float64), np.arange(shape[1], dtype=np.float64),
                       indexing="ij")
    return A * np.sin(k[0] * i + k[1] * j + phi), A * np.cos(k[0] * i + k[1] * j + phi)


def _symbols(order, k, h, orc):
    """Exact responses of the centred differences to a sinusoid of wavenumber k (per cell)."""
    d1 = orc.centered_d1(order).astype(np.float64)
    c0, d2 = orc.centered_d2(order)
    D1 = lambda kk: 2.0 * sum(d1[l] * np.sin((l + 1) * kk) for l in range(len(d1))) / h  # noqa: E731
    D2 = lambda kk: (float(c0) + 2.0 * sum(float(d2[l]) * np.cos((l + 1) * kk) for l in range(len(d2)))) / h ** 2  # noqa: E731
    return D1(k[0]), D1(k[1]), D2(k[0]) + D2(k[1])


def _pde_rhs(P, S, grad_P, grad_S, lap_P, lap_S, med):
    """README.md:363-371 in 2D (x, y, z with d/dy = 0): the P and S_y accelerations."""
    zero = np.zeros_like(P)
    ga = np.stack([med["dalpha_dx"], zero, med["dalpha_dz"]])
    gb = np.stack([med["dbeta_dx"], zero, med["dbeta_dz"]])
    gP = np.stack([grad_P[0], zero, grad_P[1]])
    # S = (0, S_y, 0): Jacobian rows are the components' gradients
    JS = [[zero, zero, zero], [grad_S[0], zero, grad_S[1]], [zero, zero, zero]]
    curlS = _curl(JS)
    dot = lambda a, b: (a * b).sum(axis=0)  # noqa: E731
    rhs_P = (P * med["lap_alpha"] + 2 * dot(ga, gP) - 2 * P * med["lap_beta"] - 2 * dot(gb, curlS)
             + med["alpha"] * lap_P)
    # (grad beta . grad) S, - grad beta x curl S, 2 grad beta x grad P, beta lap S: y components
    adv = dot(gb, np.stack([grad_S[0], zero, grad_S[1]]))
    rhs_S = adv - np.cross(gb, curlS, axis=0)[1] + 2 * np.cross(gb, gP, axis=0)[1] + med["beta"] * lap_S
    return rhs_P, rhs_S


def _manufactured(orc, order, rng, shape=(41, 37), h=7.0):
    kP, kS = (0.31, 0.47), (0.52, -0.23)
    P, Pc = _wave(shape, kP, 1.3, 0.4)
    S, Sc = _wave(shape, kS, 0.8, 1.9)
    DxP, DzP, LP = _symbols(order, kP, h, orc)
    DxS, DzS, LS = _symbols(order, kS, h, orc)
    med = {name: rng.uniform(0.5, 2.0, shape) * (1 if name in ("alpha", "beta") else rng.choice([-1, 1], shape))
           for name in orc.COUPLED_MEDIUM}
    med["alpha"] *= 9e6
    med["beta"] *= 2e6
    for name in ("dalpha_dx", "dalpha_dz", "dbeta_dx", "dbeta_dz"):
        med[name] *= 1e4
    for name in ("lap_alpha", "lap_beta"):
        med[name] *= 1e2
    grads = dict(grad_P=(DxP * Pc, DzP * Pc), grad_S=(DxS * Sc, DzS * Sc), lap_P=LP * P, lap_S=LS * S)
    return P, S, med, grads


def _half_step_error(orc, order, P, S, med, grads, med_used=None, dt=2.0 ** -13, h=7.0):
    N = order // 2
    rng = np.random.default_rng(1)
    dP0 = rng.normal(size=P.shape)
    dS0 = rng.normal(size=P.shape)
    dP, dS = orc.probe_coupled_velocity(dP0, dS0, P, S, med if med_used is None else med_used, order, dt, h)
    rP, rS = _pde_rhs(P, S, grads["grad_P"], grads["grad_S"], grads["lap_P"], grads["lap_S"], med)
    inner = (slice(N, P.shape[0] - N), slice(N, P.shape[1] - N))
    eP = np.abs(dP - (dP0 + dt * rP))[inner].max() / np.abs(dt * rP[inner]).max()
    eS = np.abs(dS - (dS0 + dt * rS))[inner].max() / np.abs(dt * rS[inner]).max()
    # outside the update region nothing changes
    edge = np.ones(P.shape, bool)
    edge[inner] = False
    assert np.array_equal(dP[edge], dP0[edge]) and np.array_equal(dS[edge], dS0[edge])
    return eP, eS


@pytest.mark.parametrize("order", [2, 4, 6, 8, 10])
def test_velocity_half_step_is_the_coupled_pde(orc, order):
    rng = np.random.default_rng(RNG_SEED + order)
    P, S, med, grads = _manufactured(orc, order, rng)
    eP, eS = _half_step_error(orc, order, P, S, med, grads)
    assert eP < 1e-9 and eS < 1e-9, (eP, eS)


@pytest.mark.parametrize("mistake", ["alpha<->beta", "dbeta x<->z", "-dbeta", "lap alpha<->lap beta",
                                     "dalpha<->dbeta"])
def test_planted_medium_mistakes_fail_the_pde(orc, mistake):
    rng = np.random.default_rng(RNG_SEED)
    P, S, med, grads = _manufactured(orc, 8, rng)
    bad = dict(med)
    if mistake == "alpha<->beta":
        bad["alpha"], bad["beta"] = med["beta"], med["alpha"]
    elif mistake == "dbeta x<->z":
        bad["dbeta_dx"], bad["dbeta_dz"] = med["dbeta_dz"], med["dbeta_dx"]
    elif mistake == "-dbeta":
        bad["dbeta_dx"], bad["dbeta_dz"] = -med["dbeta_dx"], -med["dbeta_dz"]
    elif mistake == "lap alpha<->lap beta":
        bad["lap_alpha"], bad["lap_beta"] = med["lap_beta"], med["lap_alpha"]
    else:
        bad["dalpha_dx"], bad["dbeta_dx"] = med["dbeta_dx"], med["dalpha_dx"]
        bad["dalpha_dz"], bad["dbeta_dz"] = med["dbeta_dz"], med["dalpha_dz"]
    eP, eS = _half_step_error(orc, 8, P, S, med, grads, med_used=bad)
    assert max(eP, eS) > 1e-3, (mistake, eP, eS)


def test_position_half_step(orc):
    rng = np.random.default_rng(3)
    shape, N, dt = (30, 26), 4, 2.0 ** -12
    P, S, dP, dS = (rng.normal(size=shape) for _ in range(4))
    Pn, Sn = orc.probe_coupled_position(P, S, dP, dS, 8, dt)
    inner = (slice(N, shape[0] - N), slice(N, shape[1] - N))
    assert np.allclose(Pn[inner], (P + dt * dP)[inner], rtol=0, atol=1e-15)
    assert np.allclose(Sn[inner], (S + dt * dS)[inner], rtol=0, atol=1e-15)
    edge = np.ones(shape, bool)
    edge[inner] = False
    assert np.array_equal(Pn[edge], P[edge]) and np.array_equal(Sn[edge], S[edge])


@pytest.mark.parametrize("order", [4, 8])
def test_homogeneous_modes_follow_the_leapfrog_dispersion(orc, order):
    """alpha, beta constant, their derivatives zero: a P mode cos(k.x) with dP/dt = 0 evolves as
    P^n = cos((n + 1/2) theta) / cos(theta / 2) P^0 with cos theta = 1 - dt^2 alpha kappa^2 / 2 and
    -kappa^2 the Laplacian's symbol; the S mode with beta.  Swapping alpha and beta fails."""
    N, h, dt, n = order // 2, 10.0, 2.0 ** -10, 25
    m = 2 * N * n + 24
    shape = (m, m)
    alpha, beta = 3000.0 ** 2, 1700.0 ** 2
    zero = np.zeros(shape)
    med = dict(alpha=np.full(shape, alpha), beta=np.full(shape, beta), dalpha_dx=zero, dalpha_dz=zero,
               dbeta_dx=zero, dbeta_dz=zero, lap_alpha=zero, lap_beta=zero)
    kP, kS = (0.41, 0.23), (0.17, 0.62)
    P0 = _wave(shape, kP, 1.0, 0.3)[0]
    S0 = _wave(shape, kS, 1.0, 1.1)[0]

    def run(md):
        P, S, dP, dS = P0.copy(), S0.copy(), zero.copy(), zero.copy()
        for _ in range(n):
            dP, dS = orc.probe_coupled_velocity(dP, dS, P, S, md, order, dt, h)
            P, S = orc.probe_coupled_position(P, S, dP, dS, order, dt)
        return P, S

    def amp(c, k):
        lap = _symbols(order, k, h, orc)[2]
        th = np.arccos(1.0 - 0.5 * dt * dt * c * (-lap))
        return np.cos((n + 0.5) * th) / np.cos(0.5 * th)

    P, S = run(med)
    core = (slice(N * n + 2, m - N * n - 2),) * 2
    assert np.abs(P - amp(alpha, kP) * P0)[core].max() < 1e-9
    assert np.abs(S - amp(beta, kS) * S0)[core].max() < 1e-9
    Pb, _ = run(dict(med, alpha=med["beta"], beta=med["alpha"]))
    assert np.abs(Pb - amp(alpha, kP) * P0)[core].max() > 1e-2


def _layered(nx, nz, vs_interface):
    vp = np.full((nx, nz), 2500.0, np.float32)
    vs = np.full((nx, nz), 1200.0, np.float32)
    vp[:, nz // 2:] = 3400.0
    if vs_interface:
        vs[:, nz // 2:] = 1900.0
    return vp, vs


def test_no_s_without_a_beta_gradient(orc):
    """README.md:372-375: mode conversion occurs only at vs contrasts.  A vp-only interface leaves
    S at rounding level (FP64 oracle, smoothing on); a vs interface converts P to S."""
    nx, nz, nt = 90, 80, 420
    w = orc.ricker(15.0, 1e-3, nt)
    rx = np.arange(1, nx + 1, 3)
    kw = dict(sx=[45], sz=[10], rx=rx, rz=np.full(len(rx), 8), wavelet=w[None], nbc=20, precision="f64")
    r0 = orc.coupled2d(*_layered(nx, nz, False), 10.0, 1e-3, nt, **kw)
    r1 = orc.coupled2d(*_layered(nx, nz, True), 10.0, 1e-3, nt, **kw)
    pmax = np.abs(r0["seis_P"]).max()
    assert np.abs(r0["seis_S"]).max() < 1e-10 * pmax
    assert np.abs(r1["seis_S"]).max() > 1e-4 * np.abs(r1["seis_P"]).max()


def test_fp32_oracle_tracks_fp64(orc):
    """The FP32 restatement (the GPU's bit-exact twin) stays within FP32 rounding growth of FP64."""
    nx, nz, nt = 80, 70, 300
    vp, vs = _layered(nx, nz, True)
    w = orc.ricker(15.0, 1e-3, nt)
    rx = np.arange(1, nx + 1, 2)
    kw = dict(sx=[40], sz=[10], rx=rx, rz=np.full(len(rx), 6), wavelet=w[None], nbc=20)
    a = orc.coupled2d(vp, vs, 10.0, 1e-3, nt, **kw)
    b = orc.coupled2d(vp, vs, 10.0, 1e-3, nt, precision="f64", **kw)
    for key in ("seis_P", "seis_S"):
        err = np.linalg.norm(a[key] - b[key]) / np.linalg.norm(b[key])
        assert err < 1e-4, (key, err)


def test_p_field_boundary_absorbs_against_a_boundary_free_domain(orc):
    """The Higdon blends with the P coefficients (v_ref_p) absorb the P potential: a 100 x 100 run
    with nbc 40 matches the same receivers in a domain 700 cells wide and 450 deep (no boundary
    reaches them within 0.7 s) to 0.5 % (measured 0.16 %); giving the P fields the S fields'
    reference velocity instead (a swapped coefficient set) is ten times worse."""
    nt = 700
    w = orc.ricker(15.0, 1e-3, nt)

    def run(nx, nz, off, **kw):
        vp = np.full((nx, nz), 3000.0, np.float32)
        vs = np.full((nx, nz), 1500.0, np.float32)
        vs[:, 50:] = 2000.0
        rx = np.arange(1, 101) + off
        return orc.coupled2d(vp, vs, 10.0, 1e-3, nt, sx=[50 + off], sz=[10], rx=rx, rz=np.full(100, 10),
                             wavelet=w[None], nbc=40, **kw)["seis_P"]

    big = run(700, 450, 300)
    res = lambda a: float(np.linalg.norm(a - big) / np.linalg.norm(big))  # noqa: E731
    good = res(run(100, 100, 0))
    swapped = res(run(100, 100, 0, v_ref_p=1500.0, v_ref_s=3000.0))
    assert good < 5e-3, good
    assert swapped > 5 * good, (good, swapped)
