"""Closed-form pins of the oracle's elastic (and 3D) update equations, independent of the GPU.

The elastic sweeps restate src/equations/elastic2d/fused_kernels.jl:21-124 and
src/equations/elastic3d/update_velocity.jl:17-84, update_stress.jl:21-94.  No reference output
pins them (the committed elastic report, example/benchmark/output/report_elastic.txt, was written
by an older engine; the oracle's own figure for it is recorded, not used, in the design notes), so
they are pinned here against physics the reference states in its comments:

  * Newton's law  rho dv_i/dt = d sigma_ij / dx_j  and isotropic Hooke's law
    d sigma_ij/dt = lam delta_ij div v + mu (dv_i/dx_j + dv_j/dx_i)
    (update_stress.jl:3-16, elastic2d/fused_kernels.jl:105-121);
  * the staggered positions the reference documents: elastic2d.jl:38-40 (txx/tzz (i,j), vx
    (i-1/2,j), vz (i,j+1/2), txz (i-1/2,j+1/2)) and elastic3d/medium.jl:68-71 plus the buoyancy
    stagger of pad_array_3d.jl (vx (i-1/2), vy (j-1/2), vz (k+1/2), txy (i-1/2,j-1/2,k), txz
    (i-1/2,j,k+1/2), tyz (i,j-1/2,k+1/2));
  * the exact response of a staggered difference to a sinusoid: for f = A sin(k.x + phi) sampled
    at its staggered points, sum_l a_l (f(y + (2l-1)/2 e_d) - f(y - (2l-1)/2 e_d)) equals
    A K(k_d) cos(k.y + phi) with K(k) = 2 sum_l a_l sin((2l-1) k / 2), exactly, at the midpoint y.

One sweep of the oracle on a state made of oblique plane waves (a different amplitude, phase and
position per field, per-cell random materials) must equal that closed form to rounding.  A
swapped lam / lam+2mu, mu replaced by lam, a flipped sign in a shear sum, or a stencil one cell
off all fail it (planted below through the probe's inputs, which is where each mistake enters).
The whole elastic2d solver is then pinned by its wave speeds: P and S arrivals travel at vp and vs.
"""
import numpy as np
import pytest

RNG_SEED = 20261204


def K(a, k):
    """Symbol of the staggered first difference with coefficients a (cell units)."""
    return 2.0 * sum(float(a[l]) * np.sin((2 * l + 1) * k / 2.0) for l in range(len(a)))


def grids(shape):
    return np.meshgrid(*[np.arange(1, n + 1, dtype=np.float64) for n in shape], indexing="ij")


def wave(shape, off, k, A, phi):
    g = grids(shape)
    return A * np.sin(sum(kd * (gd + od) for kd, gd, od in zip(k, g, off)) + phi)


def dwave(shape, off, k, A, phi, d, a):
    """Staggered d-derivative of wave(...) evaluated at the points `off` (the target field's)."""
    g = grids(shape)
    return A * K(a, k[d]) * np.cos(sum(kd * (gd + od) for kd, gd, od in zip(k, g, off)) + phi)


def inner(x, M):
    return x[tuple(slice(M, n - M) for n in x.shape)]


def rel(got, want, M):
    g, w = inner(got, M), inner(want, M)
    return float(np.max(np.abs(g - w)) / np.max(np.abs(w)))


# ---------------------------------------------------------------------------------------------
# 2D elastic

POS2 = {"txx": (0.0, 0.0), "tzz": (0.0, 0.0), "vx": (-0.5, 0.0), "vz": (0.0, 0.5), "txz": (-0.5, 0.5)}


def state2(shape, k, rng):
    amp = {f: rng.uniform(0.5, 2.0) for f in POS2}
    ph = {f: rng.uniform(0, 2 * np.pi) for f in POS2}
    F = {f: wave(shape, POS2[f], k, amp[f], ph[f]) for f in POS2}
    return F, amp, ph


@pytest.mark.parametrize("order", [2, 4, 6, 8, 10])
def test_elastic2d_sweeps_are_newton_and_hooke(orc, order):
    rng = np.random.default_rng(RNG_SEED + order)
    shape = (37, 29)
    k = (0.37, 0.53)
    a = orc.fd_coeffs(order).astype(np.float64)
    M = order // 2
    dtx = 0.1
    F, A, P = state2(shape, k, rng)
    bx, bz, lam, l2m, mxz = (rng.uniform(0.3, 3.0, shape) for _ in range(5))

    def D(f, target, d):
        return dwave(shape, POS2[target], k, A[f], P[f], d, a)

    vx, vz = orc.probe_elastic_velocity(F["vx"], F["vz"], F["txx"], F["tzz"], F["txz"], bx, bz, order, dtx)
    # rho dvx/dt = dtxx/dx + dtxz/dz ; rho dvz/dt = dtxz/dx + dtzz/dz
    assert rel(vx - F["vx"], bx * dtx * (D("txx", "vx", 0) + D("txz", "vx", 1)), M) < 1e-12
    assert rel(vz - F["vz"], bz * dtx * (D("txz", "vz", 0) + D("tzz", "vz", 1)), M) < 1e-12
    # outside the update region nothing moves
    assert np.array_equal(vx[:M], F["vx"][:M]) and np.array_equal(vz[:, -M:], F["vz"][:, -M:])

    txx, tzz, txz = orc.probe_elastic_stress(F["txx"], F["tzz"], F["txz"], F["vx"], F["vz"], lam, l2m, mxz, order,
                                             dtx)
    exx, ezz = D("vx", "txx", 0), D("vz", "txx", 1)
    want_xx = dtx * (l2m * exx + lam * ezz)
    want_zz = dtx * (lam * exx + l2m * ezz)
    want_xz = dtx * mxz * (D("vx", "txz", 1) + D("vz", "txz", 0))
    assert rel(txx - F["txx"], want_xx, M) < 1e-12
    assert rel(tzz - F["tzz"], want_zz, M) < 1e-12
    assert rel(txz - F["txz"], want_xz, M) < 1e-12

    # planted mistakes, each entering where the oracle would make it
    sw = orc.probe_elastic_stress(F["txx"], F["tzz"], F["txz"], F["vx"], F["vz"], l2m, lam, mxz, order, dtx)
    assert rel(sw[0] - F["txx"], want_xx, M) > 1e-2 and rel(sw[1] - F["tzz"], want_zz, M) > 1e-2
    ml = orc.probe_elastic_stress(F["txx"], F["tzz"], F["txz"], F["vx"], F["vz"], lam, l2m, lam, order, dtx)
    assert rel(ml[2] - F["txz"], want_xz, M) > 1e-2
    fl = orc.probe_elastic_stress(F["txx"], F["tzz"], F["txz"], F["vx"], -F["vz"], lam, l2m, mxz, order, dtx)
    assert rel(fl[2] - F["txz"], want_xz, M) > 1e-2  # mu (dvx/dz - dvz/dx)
    for name, ax in (("txx", 0), ("txz", 1)):
        G = dict(F)
        G[name] = np.roll(F[name], 1, axis=ax)
        sh = orc.probe_elastic_velocity(G["vx"], G["vz"], G["txx"], G["tzz"], G["txz"], bx, bz, order, dtx)
        assert rel(sh[0] - F["vx"], bx * dtx * (D("txx", "vx", 0) + D("txz", "vx", 1)), M + 1) > 1e-2


# ---------------------------------------------------------------------------------------------
# 3D elastic and acoustic

POS3 = {"vx": (-0.5, 0, 0), "vy": (0, -0.5, 0), "vz": (0, 0, 0.5), "txx": (0, 0, 0), "tyy": (0, 0, 0),
        "tzz": (0, 0, 0), "txy": (-0.5, -0.5, 0), "txz": (-0.5, 0, 0.5), "tyz": (0, -0.5, 0.5)}
ORDER9 = ("vx", "vy", "vz", "txx", "tyy", "tzz", "txy", "txz", "tyz")


@pytest.mark.parametrize("order", [2, 4, 8])
def test_elastic3d_sweeps_are_newton_and_hooke(orc, order):
    rng = np.random.default_rng(RNG_SEED + 10 + order)
    shape = (19, 17, 21)
    k = (0.41, 0.29, 0.57)
    a = orc.fd_coeffs(order).astype(np.float64)
    M = order // 2
    dtdh = 0.07
    A = {f: rng.uniform(0.5, 2.0) for f in ORDER9}
    P = {f: rng.uniform(0, 2 * np.pi) for f in ORDER9}
    F = [wave(shape, POS3[f], k, A[f], P[f]) for f in ORDER9]
    bx, by, bz, lam, l2m, mxy, mxz, myz = (rng.uniform(0.3, 3.0, shape) for _ in range(8))

    def D(f, target, d):
        return dwave(shape, POS3[target], k, A[f], P[f], d, a)

    out = orc.probe_elastic3d_velocity(F, bx, by, bz, order, dtdh)
    want_v = want = {"vx": bx * dtdh * ((D("txx", "vx", 0) + D("txy", "vx", 1)) + D("txz", "vx", 2)),
            "vy": by * dtdh * ((D("txy", "vy", 0) + D("tyy", "vy", 1)) + D("tyz", "vy", 2)),
            "vz": bz * dtdh * ((D("txz", "vz", 0) + D("tyz", "vz", 1)) + D("tzz", "vz", 2))}
    for m, f in enumerate(ORDER9[:3]):
        assert rel(out[m] - F[m], want[f], M) < 1e-12, f
    for m in range(3, 9):
        assert np.array_equal(out[m], F[m])

    out = orc.probe_elastic3d_stress(F, lam, l2m, mxy, mxz, myz, order, dtdh)
    e = [D("vx", "txx", 0), D("vy", "txx", 1), D("vz", "txx", 2)]
    want = {"txx": dtdh * (l2m * e[0] + lam * e[1] + lam * e[2]),
            "tyy": dtdh * (lam * e[0] + l2m * e[1] + lam * e[2]),
            "tzz": dtdh * (lam * e[0] + lam * e[1] + l2m * e[2]),
            "txy": dtdh * mxy * (D("vx", "txy", 1) + D("vy", "txy", 0)),
            "txz": dtdh * mxz * (D("vx", "txz", 2) + D("vz", "txz", 0)),
            "tyz": dtdh * myz * (D("vy", "tyz", 2) + D("vz", "tyz", 1))}
    for m, f in enumerate(ORDER9):
        if m < 3:
            assert np.array_equal(out[m], F[m])
        else:
            assert rel(out[m] - F[m], want[f], M) < 1e-12, f

    sw = orc.probe_elastic3d_stress(F, l2m, lam, mxy, mxz, myz, order, dtdh)
    assert rel(sw[4] - F[4], want["tyy"], M) > 1e-2
    ml = orc.probe_elastic3d_stress(F, lam, l2m, lam, mxz, myz, order, dtdh)
    assert rel(ml[6] - F[6], want["txy"], M) > 1e-2
    Fn = list(F)
    Fn[1] = -F[1]
    fl = orc.probe_elastic3d_stress(Fn, lam, l2m, mxy, mxz, myz, order, dtdh)
    assert rel(fl[8] - F[8], want["tyz"], M) > 1e-2
    Fs = list(F)
    Fs[6] = np.roll(F[6], 1, axis=1)  # txy one cell off along y
    sh = orc.probe_elastic3d_velocity(Fs, bx, by, bz, order, dtdh)
    assert rel(sh[0] - F[0], want_v["vx"], M + 1) > 1e-2


@pytest.mark.parametrize("order", [2, 4, 8])
def test_acoustic3d_sweeps_plane_wave(orc, order):
    """update_velocity.jl:15-52 / update_pressure.jl:13-48: v += b dt/dh grad p at the staggered
    velocity points (vx i-1/2, vy j-1/2, vz k+1/2), p += kappa dt/dh div v at cell centres."""
    rng = np.random.default_rng(RNG_SEED + 20 + order)
    shape = (19, 17, 21)
    k = (0.33, 0.61, 0.45)
    a = orc.fd_coeffs(order).astype(np.float64)
    M = order // 2
    dtdh = 0.09
    pos = {"p": (0, 0, 0), "vx": (-0.5, 0, 0), "vy": (0, -0.5, 0), "vz": (0, 0, 0.5)}
    A = {f: rng.uniform(0.5, 2.0) for f in pos}
    P = {f: rng.uniform(0, 2 * np.pi) for f in pos}
    F = {f: wave(shape, pos[f], k, A[f], P[f]) for f in pos}
    bx, by, bz, kap = (rng.uniform(0.3, 3.0, shape) for _ in range(4))

    def D(f, target, d):
        return dwave(shape, pos[target], k, A[f], P[f], d, a)

    vx, vy, vz = orc.probe_acoustic3d_velocity(F["vx"], F["vy"], F["vz"], F["p"], bx, by, bz, order, dtdh)
    assert rel(vx - F["vx"], bx * dtdh * D("p", "vx", 0), M) < 1e-12
    assert rel(vy - F["vy"], by * dtdh * D("p", "vy", 1), M) < 1e-12
    assert rel(vz - F["vz"], bz * dtdh * D("p", "vz", 2), M) < 1e-12
    p = orc.probe_acoustic3d_pressure(F["p"], F["vx"], F["vy"], F["vz"], kap, order, dtdh)
    want = kap * dtdh * ((D("vx", "p", 0) + D("vy", "p", 1)) + D("vz", "p", 2))
    assert rel(p - F["p"], want, M) < 1e-12
    # vy on the other side of its cell (j+1/2) is a different, self-consistent staggering that the
    # pressure alone would not reveal; the velocity sweep's position check does
    wrong = dwave(shape, (0, 0.5, 0), k, A["p"], P["p"], 1, a)
    assert rel(vy - F["vy"], by * dtdh * wrong, M) > 1e-2


def test_acoustic2d_sweeps_plane_wave(orc):
    """acoustic2d/fused_kernels.jl:27-114 at its documented stagger (vx i-1/2, vz j+1/2)."""
    rng = np.random.default_rng(RNG_SEED + 30)
    shape, k, order, dtx = (41, 33), (0.43, 0.31), 8, 0.1
    a = orc.fd_coeffs(order).astype(np.float64)
    pos = {"p": (0, 0), "vx": (-0.5, 0), "vz": (0, 0.5)}
    A = {f: rng.uniform(0.5, 2.0) for f in pos}
    P = {f: rng.uniform(0, 2 * np.pi) for f in pos}
    F = {f: wave(shape, pos[f], k, A[f], P[f]) for f in pos}
    bx, bz, kap = (rng.uniform(0.3, 3.0, shape) for _ in range(3))
    vx, vz = orc.probe_acoustic_velocity(F["vx"], F["vz"], F["p"], bx, bz, order, dtx, "f64")
    assert rel(vx - F["vx"], bx * dtx * dwave(shape, pos["vx"], k, A["p"], P["p"], 0, a), 4) < 1e-12
    assert rel(vz - F["vz"], bz * dtx * dwave(shape, pos["vz"], k, A["p"], P["p"], 1, a), 4) < 1e-12
    p = orc.probe_acoustic_pressure(F["p"], F["vx"], F["vz"], kap, order, dtx, "f64")
    want = kap * dtx * (dwave(shape, pos["p"], k, A["vx"], P["vx"], 0, a) +
                        dwave(shape, pos["p"], k, A["vz"], P["vz"], 1, a))
    assert rel(p - F["p"], want, 4) < 1e-12


# ---------------------------------------------------------------------------------------------
# whole elastic2d solver: wave speeds

def _lag(a, b):
    """Delay of b behind a in samples (cross-correlation peak, parabolic refinement)."""
    a = a - a.mean()
    b = b - b.mean()
    cc = np.correlate(b, a, "full")
    m = int(np.argmax(cc))
    y0, y1, y2 = cc[m - 1], cc[m], cc[m + 1]
    return m - (len(a) - 1) + 0.5 * (y0 - y2) / (y0 - 2 * y1 + y2)


def test_elastic2d_p_and_s_speeds(orc):
    """Homogeneous vp 3000, vs 1700: an explosive source radiates P only (the radial vx on a line
    through the source moves at vp); a vertical force radiates S broadside (vz on the horizontal
    line moves at vs) and P along its axis (vz on the vertical line moves at vp).  Speeds from the
    delay between receivers 50 and 150 cells away, within 1 % (FD8 at >= 13 points per
    wavelength; the measured errors are 0.1-0.4 %).  Replacing mu by lam in txz would give an S
    speed of sqrt(vp^2 - 2 vs^2) = 1794 (+5.5 %), lam+mu in place of lam+2mu a P speed of 2472."""
    n, nt, dh, dt, f0 = 361, 1400, 5.0, 5e-4, 12.0
    vp = np.full((n, n), 3000.0, np.float32)
    vs = np.full((n, n), 1700.0, np.float32)
    rho = np.full((n, n), 2000.0, np.float32)
    c = 181
    offs = np.array([50, 150])
    w = orc.ricker(f0, dt, nt)[None]
    kw = dict(wavelet=w, nbc=30, fd_order=8, precision="f64")

    def speed(tr):
        return (offs[1] - offs[0]) * dh / (_lag(tr[0], tr[1]) * dt)

    e = orc.elastic2d(vp, vs, rho, dh, dt, nt, sx=[c], sz=[c], rx=c + offs, rz=np.full(2, c), **kw)
    assert abs(speed(e["seis_vx"]) / 3000.0 - 1) < 0.01
    e = orc.elastic2d(vp, vs, rho, dh, dt, nt, sx=[c], sz=[c], rx=np.r_[c + offs, np.full(2, c)],
                      rz=np.r_[np.full(2, c), c + offs], source_type=1, **kw)
    assert abs(speed(e["seis_vz"][:2]) / 1700.0 - 1) < 0.01
    assert abs(speed(e["seis_vz"][2:]) / 3000.0 - 1) < 0.01
