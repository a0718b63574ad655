"""CPU restatement of the learned-boundary MLP (features, forward, backward) in numpy.

TEST INFRASTRUCTURE ONLY: the checker for the HIP kernels in fomo.jl_amd/csrc/learned2d.hip.
The reference (Fomo.jl) has no learned boundary; this oracle is pinned by closed forms (chosen
routes through the network, a unit spike for the patch layout, a nearest-neighbour search for the
BFloat16 rounding) and finite differences in tests/test_learned_cpu.py, not by reference outputs.

Parameter layout (3586 values): W1 [32][76], b1 [32], W2 [32][32], b2 [32], W3 [2][32], b3 [2].
Outputs are blend-weight corrections corr = CORR_MAX * tanh(z).
"""
from __future__ import annotations

import numpy as np

N_IN, N_H, N_OUT = 76, 32, 2
SIZES = [("W1", (N_H, N_IN)), ("b1", (N_H,)), ("W2", (N_H, N_H)), ("b2", (N_H,)), ("W3", (N_OUT, N_H)),
         ("b3", (N_OUT,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in SIZES)
CORR_MAX = 0.25


def unpack(w):
    w = np.asarray(w, np.float64)
    out, o = {}, 0
    for name, shape in SIZES:
        n = int(np.prod(shape))
        out[name] = w[o:o + n].reshape(shape)
        o += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], np.float64).ravel() for name, _ in SIZES])


def bf16_round(w):
    """Round Float32 values to BFloat16 (round to nearest even), returned as float64."""
    f = np.asarray(w, np.float32).view(np.uint32).astype(np.uint64)
    lsb = (f >> 16) & 1
    r = ((f + 0x7FFF + lsb) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def features(p, vx, vz, cells, H):
    """5x5 patches (clamped) of p, vx, vz at padded cells (x-fastest index); p divided by its patch
    max-abs (+1e-30), vx and vz by theirs; then dist/H.  Fields are (nx, nz) arrays; the scaling is
    evaluated in Float32 like the kernel."""
    nx, nz = p.shape
    cells = np.asarray(cells)
    i, j = cells % nx, cells // nx
    off = np.arange(-2, 3)
    ii = np.clip(i[:, None, None] + off[None, None, :], 0, nx - 1)   # (n, 1, 5) along x
    jj = np.clip(j[:, None, None] + off[None, :, None], 0, nz - 1)   # (n, 5, 1) along z
    P = p[ii, jj].reshape(len(cells), 25).astype(np.float32)
    X = vx[ii, jj].reshape(len(cells), 25).astype(np.float32)
    Z = vz[ii, jj].reshape(len(cells), 25).astype(np.float32)
    sp = (np.abs(P).max(1) + np.float32(1e-30)).astype(np.float32)
    sv = (np.maximum(np.abs(X).max(1), np.abs(Z).max(1)) + np.float32(1e-30)).astype(np.float32)
    dist = np.minimum(np.minimum(i, nx - 1 - i), np.minimum(j, nz - 1 - j))
    return np.concatenate([P / sp[:, None], X / sv[:, None], Z / sv[:, None], (dist / H)[:, None]],
                          axis=1).astype(np.float64)


def forward(w, x):
    m = unpack(w)
    a1 = x @ m["W1"].T + m["b1"]
    h1 = np.maximum(a1, 0.0)
    a2 = h1 @ m["W2"].T + m["b2"]
    h2 = np.maximum(a2, 0.0)
    z = h2 @ m["W3"].T + m["b3"]
    corr = CORR_MAX * np.tanh(z)
    return corr, (x, h1, h2, corr)


def forward_bf16(w, x):
    """The inference path of the sweeps (matrix cores): features and the hidden activations are
    rounded to BFloat16 before each layer, every layer accumulates in floating point (here FP64) from
    BFloat16 weights; returns the corrections (n x 2)."""
    m = unpack(w)
    h = bf16_round(np.asarray(x, np.float64))
    h1 = bf16_round(np.maximum(h @ m["W1"].T + m["b1"], 0.0))
    h2 = bf16_round(np.maximum(h1 @ m["W2"].T + m["b2"], 0.0))
    z = h2 @ m["W3"].T + m["b3"]
    return CORR_MAX * np.tanh(z)


def backward(w, cache, galpha):
    """Summed weight gradients for dL/dcorr = galpha (n x 2)."""
    m = unpack(w)
    x, h1, h2, corr = cache
    t = corr / CORR_MAX
    gz = galpha * CORR_MAX * (1.0 - t * t)
    gW3 = gz.T @ h2
    gb3 = gz.sum(0)
    gh2 = (gz @ m["W3"]) * (h2 > 0)
    gW2 = gh2.T @ h1
    gb2 = gh2.sum(0)
    gh1 = (gh2 @ m["W2"]) * (h1 > 0)
    gW1 = gh1.T @ x
    gb1 = gh1.sum(0)
    return pack({"W1": gW1, "b1": gb1, "W2": gW2, "b2": gb2, "W3": gW3, "b3": gb3})
