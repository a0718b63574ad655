"""Learned thin absorbing boundary for acoustic2d: initialisation, training and evaluation.

The network (76 -> 32 -> 32 -> 2) predicts, for every frame cell and step, bounded corrections
(0.25 tanh) to the Higdon blend weights of the velocity fields and of p, from max-abs-scaled 5x5
patches of p, vx, vz and the depth in the layer; the blend weight is clamp(ramp + correction, 0, 1).
Training data come from the simulator itself: a *student* run on the model with a
thin boundary (nbc = 16) and a *teacher* run on the same model edge-extended by `extend` cells with
a wide boundary (nbc = 200) advance together; at every step the one-step loss at the student's
frame cells against the teacher's field at the same physical cells is differentiated through the
MLP by the engine's own forward/backward kernels (fomo_learned_train_step), and Adam updates the
weights on the host.  The plain distance-ramp HABC stays the default, the initial state of the
network and the yardstick of reflection error.
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

from . import _lib
from .api import Plan

N_IN, N_H, N_OUT = 76, 32, 2
N_PARAMS = N_H * N_IN + N_H + N_H * N_H + N_H + N_OUT * N_H + N_OUT
WEIGHTS_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "learned_weights.npy")


def _split(w):
    o = 0
    out = []
    for shape in ((N_H, N_IN), (N_H,), (N_H, N_H), (N_H,), (N_OUT, N_H), (N_OUT,)):
        n = int(np.prod(shape))
        out.append(w[o:o + n].reshape(shape))
        o += n
    return out


def initial_weights(seed: int = 0) -> np.ndarray:
    """Random hidden layers and zero output layer: the corrections start at exactly 0, so the
    learned boundary starts as the plain distance-ramp Higdon boundary (habc.jl:30-34)."""
    rng = np.random.default_rng(seed)
    W1 = rng.normal(0, 1.0 / math.sqrt(N_IN), (N_H, N_IN))
    b1 = np.full(N_H, 0.05)
    W2 = rng.normal(0, 1.0 / math.sqrt(N_H), (N_H, N_H))
    b2 = np.full(N_H, 0.05)
    W3 = np.zeros((N_OUT, N_H))
    b3 = np.zeros(N_OUT)
    return np.concatenate([W1.ravel(), b1, W2.ravel(), b2, W3.ravel(), b3]).astype(np.float32)


def load_weights() -> np.ndarray:
    """The trained weights shipped with the package."""
    if not os.path.exists(WEIGHTS_PATH):
        raise FileNotFoundError(f"{WEIGHTS_PATH} is missing: train with fomo.learned.train() first")
    w = np.load(WEIGHTS_PATH, allow_pickle=False).astype(np.float32)
    if w.shape != (N_PARAMS,):
        raise ValueError(f"learned weights have shape {w.shape}, expected ({N_PARAMS},)")
    return w


def enable(plan: Plan, weights, train: bool = False) -> None:
    w = np.ascontiguousarray(np.asarray(weights, np.float32))
    _lib.check(_lib.lib().fomo_learned_enable(plan._h, w.ctypes.data, 1 if train else 0))


def set_weights(plan: Plan, weights) -> None:
    w = np.ascontiguousarray(np.asarray(weights, np.float32))
    _lib.check(_lib.lib().fomo_learned_set_weights(plan._h, w.ctypes.data))


def alphas(plan: Plan):
    n = C.c_int32()
    _lib.check(_lib.lib().fomo_learned_alpha(plan._h, C.byref(n), None, None, None))
    cells = np.zeros(n.value, np.int32)
    av = np.zeros(n.value, np.float32)
    ap = np.zeros(n.value, np.float32)
    _lib.check(_lib.lib().fomo_learned_alpha(plan._h, C.byref(n), cells.ctypes.data, av.ctypes.data,
                                             ap.ctypes.data))
    return cells, av, ap


def mlp_eval(weights, feats, galpha):
    """GPU forward/backward of the MLP on given features: (alpha (n,2), summed gradient)."""
    w = np.ascontiguousarray(np.asarray(weights, np.float32))
    f = np.ascontiguousarray(np.asarray(feats, np.float32))
    g = np.ascontiguousarray(np.asarray(galpha, np.float32))
    n = f.shape[0]
    alpha = np.zeros((n, 2), np.float32)
    grad = np.zeros(N_PARAMS, np.float32)
    _lib.check(_lib.lib().fomo_mlp_eval(w.ctypes.data, n, f.ctypes.data, g.ctypes.data, alpha.ctypes.data,
                                        grad.ctypes.data))
    return alpha, grad


def edge_extend(m, L):
    ix = np.clip(np.arange(1, m.shape[0] + 2 * L + 1) - L, 1, m.shape[0]) - 1
    iz = np.clip(np.arange(1, m.shape[1] + 2 * L + 1) - L, 1, m.shape[1]) - 1
    return np.asfortranarray(m[np.ix_(ix, iz)])


class Adam:
    def __init__(self, w, lr=2e-3, b1=0.9, b2=0.999, eps=1e-8):
        self.w = np.asarray(w, np.float64).copy()
        self.m = np.zeros_like(self.w)
        self.v = np.zeros_like(self.w)
        self.t = 0
        self.lr, self.b1, self.b2, self.eps = lr, b1, b2, eps

    def step(self, g):
        self.t += 1
        self.m = self.b1 * self.m + (1 - self.b1) * g
        self.v = self.b2 * self.v + (1 - self.b2) * g * g
        mh = self.m / (1 - self.b1 ** self.t)
        vh = self.v / (1 - self.b2 ** self.t)
        self.w -= self.lr * mh / (np.sqrt(vh) + self.eps)
        return self.w.astype(np.float32)


def training_model(seed: int, nx: int, nz: int):
    from .models import layered_2d
    vp, vs, rho = layered_2d(nx, nz, seed=seed, n_interfaces=4, jitter=0.08 * nz, vacuum_top=False)
    return vp, rho


def train(n_models: int = 512, nx: int = 240, nz: int = 160, nt: int = 500, nbc: int = 16,
          teacher_nbc: int = 200, extend: int = 60, update_every: int = 25, lr: float = 5e-4, seed: int = 1,
          weights=None, log=None, clip: float = 1.0):
    """Train the learned boundary on `n_models` seeded random layered models.  Returns (weights,
    history of per-model mean losses)."""
    dh, dt, f0 = 10.0, 8e-4, 15.0
    w = initial_weights() if weights is None else np.asarray(weights, np.float32)
    opt = Adam(w, lr=lr)
    from .setup import ricker_wavelet
    wav = ricker_wavelet(f0, dt, nt)
    scale_p = float(np.abs(wav).max())
    history = []
    g = np.zeros(N_PARAMS, np.float32)
    for m in range(n_models):
        vp, rho = training_model(seed * 100003 + m, nx, nz)
        rng = np.random.default_rng(seed * 7919 + m)
        sx = int(rng.integers(nx // 5, 4 * nx // 5))
        sz = int(rng.integers(nz // 5, 4 * nz // 5))
        z_min = float((rho * vp).min())
        scale_v = scale_p / z_min
        rx = np.arange(1, nx + 1)
        rz = np.full(nx, 1)
        stu = Plan(_lib.ACOUSTIC2D, vp, rho, dh, dt, nt, sx=[sx], sz=[sz], rx=rx, rz=rz, wavelet=wav[None],
                   nbc=nbc, fd_order=8, use_graph=False, tile_cfg=2)
        tea = Plan(_lib.ACOUSTIC2D, edge_extend(vp, extend), edge_extend(rho, extend), dh, dt, nt,
                   sx=[sx + extend], sz=[sz + extend], rx=rx + extend, rz=rz + extend, wavelet=wav[None],
                   nbc=teacher_nbc, fd_order=8, use_graph=False, v_ref=float(vp.min()), tile_cfg=2)
        enable(stu, w, train=True)
        off = extend + tea.pad - stu.pad
        losses = []
        for t in range(nt):
            _lib.check(_lib.lib().fomo_learned_train_step(stu._h, tea._h, off, off, scale_p, scale_v))
            if (t + 1) % update_every == 0:
                loss = C.c_double(0.0)
                _lib.check(_lib.lib().fomo_learned_grad(stu._h, g.ctypes.data, C.byref(loss)))
                gn = g.astype(np.float64) / update_every
                norm = float(np.linalg.norm(gn))
                if norm > clip:
                    gn *= clip / norm
                w = opt.step(gn)
                set_weights(stu, w)
                losses.append(loss.value / update_every)
        stu.close()
        tea.close()
        history.append(float(np.mean(losses)) if losses else 0.0)
        if log and (m % max(1, n_models // 16) == 0 or m == n_models - 1):
            log(f"model {m}: mean one-step loss {history[-1]:.4e}")
    return w, history


def reflection_residual(vp, rho, *, sx, sz, rx, rz, nt, boundaries, dh=10.0, dt=1e-3, f0=10.0, fd_order=8,
                        L=None, judge="habc_nbc100"):
    """The reference's absorption measure (test/verify_patch.jl:76-91): resid = ||u - u_ref|| / ||u_ref||
    over seis_p at all receivers, with u_ref from the model edge-extended by L cells (reflections of its
    own boundary arrive after the record ends), also given as 20 log10(resid) dB.  `boundaries` is a
    list of (name, nbc, weights-or-None[, storage]).  Also reports per-trace relative L2 against u_ref and against
    the boundary named `judge`, and the time of the step loop.  The norms are FP64 sums formed on the
    GPU (Plan.trace_misfit); the traces are not downloaded.  A 16-bit run whose stored values
    overflowed (saturated, so its traces are invalid) is reported as {"storage_overflow": True,
    "ms": ...} without residuals."""
    from .setup import ricker_wavelet
    wav = ricker_wavelet(f0, dt, nt)
    if L is None:
        L = int(math.ceil(float(vp.max()) * nt * dt / dh / 2.0)) + 20
    rx, rz = np.asarray(rx), np.asarray(rz)

    def run(v, r, nbc, sx_, sz_, rx_, rz_, weights=None, v_ref=None, storage="f32"):
        p = Plan(_lib.ACOUSTIC2D, v, r, dh, dt, nt, sx=[sx_], sz=[sz_], rx=rx_, rz=rz_, wavelet=wav[None], nbc=nbc,
                 fd_order=fd_order, v_ref=v_ref, storage=storage)
        try:
            if weights is not None:
                enable(p, weights)
            ms, overflowed = p.timed_advance_status(nt)
            return p, ms, overflowed
        except Exception:
            p.close()
            raise

    def trace_errors(num, den):
        ok = den > 1e-6 * den.max()  # traces with more than 1e-3 of the strongest trace's norm
        tr = np.sqrt(num[ok] / den[ok])
        return float(np.median(tr)), float(tr.max())

    plans = {}
    ref = None
    out = {}
    try:
        ref, _, _ = run(edge_extend(vp, L), edge_extend(rho, L), 20, sx + L, sz + L, rx + L, rz + L,
                        v_ref=float(vp.min()))
        for b in boundaries:
            plans[b[0]] = run(vp, rho, b[1], sx, sz, rx, rz, b[2], storage=b[3] if len(b) > 3 else "f32")
        for name, (p, ms, overflowed) in plans.items():
            if overflowed:
                out[name] = {"storage_overflow": True, "ms": ms, "padded_cells": p.padded_cells}
                continue
            # seis_p (slot 0) against the edge-extended run, summed in FP64 on the GPU
            num, den = p.trace_misfit(ref)
            resid = float(math.sqrt(num.sum() / den.sum()))
            med, mx = trace_errors(num, den)
            rec = {"resid": resid, "reflection_db": 20.0 * math.log10(max(resid, 1e-300)),
                   "trace_rel_l2_median": med, "trace_rel_l2_max": mx, "ms": ms,
                   "padded_cells": p.padded_cells, "shots_per_s": 1000.0 / ms}
            if judge in plans and name != judge and not plans[judge][2]:
                mj, xj = trace_errors(*p.trace_misfit(plans[judge][0]))
                rec["vs_" + judge] = {"trace_rel_l2_median": mj, "trace_rel_l2_max": xj}
            out[name] = rec
    finally:
        for p, _, _ in plans.values():
            p.close()
        if ref is not None:
            ref.close()
    return out


def patch_geometry(nx: int = 301, nz: int = 201):
    """test/verify_patch.jl:76-91: the model, the source at (151, 41) and a receiver on every column at
    depth 41 (for other sizes: the middle column, depth 41)."""
    from .models import patch_model
    vp, _, rho = patch_model(nx, nz)
    return vp, rho, (nx + 1) // 2, 41, np.arange(1, nx + 1), np.full(nx, 41)
