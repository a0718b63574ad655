"""GPU training step of the learned boundary (fomo_learned_train_step): repeatability, the
teacher/student stream ordering, and the accumulated gradient against finite differences of the
one-step loss the step reports.

The step advances a thin-boundary student and a wide-boundary teacher (the same model, edge-extended)
one step each and accumulates d loss / d weights, where the loss is the sum over the student's frame
cells of ((f_student - f_teacher) / scale)^2 after the step and the student's blend weights come
from the network evaluated on the state before it.  With the state before the step held fixed, the
loss is a function of the weights alone, so the accumulated gradient must predict the change of the
loss under a small weight perturbation."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NX, NZ, NT, K, EXT = 160, 120, 200, 100, 40


def _plans(weights, storage="f32"):
    import fomo
    import fomo._lib as L
    from fomo.api import Plan
    from fomo.learned import edge_extend, enable
    vp = np.full((NX, NZ), 2400.0, np.float32)
    vp[:, 55:] = 3100.0
    rho = np.full((NX, NZ), 2000.0, np.float32)
    rho[:, 55:] = 2300.0
    w = fomo.ricker_wavelet(15.0, 8e-4, NT)
    rx = np.arange(1, NX + 1)
    rz = np.full(NX, 1)
    # the source 12 cells from the left edge: the wave is in the frame well before step K
    stu = Plan(L.ACOUSTIC2D, vp, rho, 10.0, 8e-4, NT, sx=[12], sz=[40], rx=rx, rz=rz, wavelet=w[None], nbc=16,
               fd_order=8, use_graph=False, storage=storage)
    tea = Plan(L.ACOUSTIC2D, edge_extend(vp, EXT), edge_extend(rho, EXT), 10.0, 8e-4, NT, sx=[12 + EXT],
               sz=[40 + EXT], rx=rx + EXT, rz=rz + EXT, wavelet=w[None], nbc=60, fd_order=8, use_graph=False,
               v_ref=float(vp.min()))
    enable(stu, weights, train=True)
    scale_p = float(np.abs(w).max())
    scale_v = scale_p / float((rho * vp).min())
    return stu, tea, EXT + tea.pad - stu.pad, scale_p, scale_v


def _train(weights, n_steps, perturbed=None):
    """Advance student and teacher K steps with `weights`, optionally switch the student to
    `perturbed`, then run n_steps training steps; returns (gradient, loss)."""
    import fomo._lib as L
    from fomo.learned import N_PARAMS, set_weights
    stu, tea, off, sp, sv = _plans(weights)
    try:
        stu.advance(K)
        tea.advance(K)
        if perturbed is not None:
            set_weights(stu, perturbed)
        g = np.zeros(N_PARAMS, np.float32)
        loss = C.c_double(0.0)
        L.check(L.lib().fomo_learned_grad(stu._h, g.ctypes.data, C.byref(loss)))  # clear
        for _ in range(n_steps):
            L.check(L.lib().fomo_learned_train_step(stu._h, tea._h, off, off, sp, sv))
        L.check(L.lib().fomo_learned_grad(stu._h, g.ctypes.data, C.byref(loss)))
        return g, loss.value
    finally:
        stu.close()
        tea.close()


def test_train_steps_repeat_bitwise(gpu):
    """Two consecutive training steps (the teacher's second step overwrites the buffer the first
    training launch read, so its stream must wait for that launch) give the same gradient and loss
    bits on every run, and a non-zero gradient."""
    from fomo.learned import load_weights
    w = load_weights()
    g1, l1 = _train(w, 2)
    g2, l2 = _train(w, 2)
    assert l1 > 0 and np.abs(g1).max() > 0
    assert np.array_equal(g1, g2) and l1 == l2


def _fd_rows(w, g, d, epss):
    import mlp
    rows = []
    for eps in epss:
        wp = (w + eps * d).astype(np.float32)
        wm = (w - eps * d).astype(np.float32)
        _, lp = _train(w, 1, perturbed=wp)
        _, lm = _train(w, 1, perturbed=wm)
        step = mlp.bf16_round(wp) - mlp.bf16_round(wm)  # the weights as the engine stores them
        rows.append((eps, lp - lm, float(np.dot(g.astype(np.float64), step))))
    return rows


def test_train_step_gradient_matches_finite_differences(gpu):
    """Central differences of the reported one-step loss against the accumulated gradient, along
    random directions (the weights as stored, BFloat16).  The perturbations move the loss by only
    ~1e-5 of its value (most of it comes from frame cells the weights barely change); the loss is a
    Float32 sum of per-block partials formed in a fixed order, whose rounding (~1e-23 here against
    differences of ~1e-21) and the O(eps^2) term of the central difference leave a few per cent, so
    the bar is 10 % for the output layer.  Hidden layers: the step's inference also rounds
    activations to BFloat16, which quantises their change; 20 %.  A wrong sign, factor or missing
    term in the composition of the sensitivities fails by far more."""
    from fomo.learned import N_H, N_IN, load_weights
    w = load_weights().astype(np.float32)
    g, l0 = _train(w, 1)
    rng = np.random.default_rng(11)
    n1 = N_H * N_IN + N_H
    n2 = n1 + N_H * N_H + N_H
    d_out = np.zeros(w.size)
    d_out[n2:] = rng.choice([-1.0, 1.0], size=w.size - n2)
    d_hid = np.zeros(w.size)
    d_hid[:n2] = rng.choice([-1.0, 1.0], size=n2) * (np.abs(w[:n2]) + 1e-2)
    out_rows = _fd_rows(w, g, d_out, (0.01, 0.03))
    hid_rows = _fd_rows(w, g, d_hid, (0.03, 0.06))
    print({"loss": l0, "output layer": out_rows, "hidden layers": hid_rows})
    for rows, tol in ((out_rows, 0.10), (hid_rows, 0.20)):
        for _, act, pred in rows:
            assert pred != 0.0, rows
            assert abs(act - pred) <= tol * abs(pred), rows


def test_training_refuses_16bit_materials(gpu):
    """Training needs Float32 fields and materials: a student with Float16 materials (:mat16) is
    refused by the training step."""
    import fomo._lib as L
    from fomo.learned import initial_weights
    stu, tea, off, sp, sv = _plans(initial_weights(3), storage="mat16")
    try:
        with pytest.raises(L.ArgumentError, match="Float32 storage"):
            L.check(L.lib().fomo_learned_train_step(stu._h, tea._h, off, off, sp, sv))
    finally:
        stu.close()
        tea.close()
