"""Seeded synthetic earth models for the benchmark workloads and tests.

The generator is a project-owned SplitMix64 stream so that the same seed gives the same model
in any host language.  Layered models follow the recipe of the reference's test model
(test/verify_patch.jl:174-187: cosine-jittered interfaces, vs = vp/1.73, Gardner rho = 310 vp^0.25).
"""
from __future__ import annotations

import hashlib
import math

import numpy as np

_MASK = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed: int):
        self.s = seed & _MASK

    def next(self) -> int:
        self.s = (self.s + 0x9E3779B97F4A7C15) & _MASK
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK
        return z ^ (z >> 31)

    def uniform(self, lo: float = 0.0, hi: float = 1.0) -> float:
        return lo + (hi - lo) * ((self.next() >> 11) * (1.0 / (1 << 53)))

    def integer(self, lo: int, hi: int) -> int:
        """Uniform integer in [lo, hi]."""
        return lo + int(self.uniform() * (hi - lo + 1))


def gardner_rho(vp: np.ndarray) -> np.ndarray:
    """rho = 310 * vp^0.25 evaluated in Float64, stored as Float32 (verify_patch.jl:185)."""
    return (310.0 * np.asarray(vp, np.float64) ** 0.25).astype(np.float32)


def layered_2d(nx: int, nz: int, *, seed: int, n_interfaces: int = 12, jitter: float = 20.0,
               vmin: float = 1800.0, vmax: float = 4500.0, vacuum_top: bool = True):
    """Random layered (vp, vs, rho), each (nx, nz), x first.

    Interface k has mean depth U(1, nz), is displaced by jitter*sin(2 pi x / L_k + phi_k) with
    L_k ~ U(nx/4, 2 nx) and phi_k ~ U(0, 2 pi); each of the n+1 layers gets vp ~ U(vmin, vmax).
    vs = vp/1.73, rho = Gardner.  With vacuum_top, row z=1 is vacuum (vp = vs = rho = 0)."""
    rng = SplitMix64(seed)
    depth = np.array([rng.uniform(1.0, float(nz)) for _ in range(n_interfaces)])
    wl = np.array([rng.uniform(nx / 4.0, 2.0 * nx) for _ in range(n_interfaces)])
    ph = np.array([rng.uniform(0.0, 2.0 * math.pi) for _ in range(n_interfaces)])
    vlay = np.array([rng.uniform(vmin, vmax) for _ in range(n_interfaces + 1)], np.float32)
    x = np.arange(1, nx + 1, dtype=np.float64)[:, None]
    z = np.arange(1, nz + 1, dtype=np.float64)[None, :]
    zk = depth[None, :] + jitter * np.sin(2.0 * math.pi * x / wl[None, :] + ph[None, :])  # (nx, K)
    layer = (z[:, :, None] >= zk[:, None, :]).sum(axis=2)  # (nx, nz)
    vp = vlay[layer].astype(np.float32)
    vs = (vp / np.float32(1.73)).astype(np.float32)
    rho = gardner_rho(vp)
    if vacuum_top:
        vp[:, 0] = 0.0
        vs[:, 0] = 0.0
        rho[:, 0] = 0.0
    return np.asfortranarray(vp), np.asfortranarray(vs), np.asfortranarray(rho)


def layered_3d(nx: int, ny: int, nz: int, *, seed: int, n_interfaces: int = 12, jitter: float = 20.0,
               vmin: float = 1800.0, vmax: float = 4500.0, n_blobs: int = 8, sigma: float = 30.0,
               blob_amp: float = 0.05):
    """3D layered vp with Gaussian perturbations (+-blob_amp, width sigma cells) and Gardner rho.

    Returns (vp, rho), each (nx, ny, nz), Float32, x fastest in memory."""
    rng = SplitMix64(seed)
    K = n_interfaces
    depth = np.array([rng.uniform(1.0, float(nz)) for _ in range(K)])
    wlx = np.array([rng.uniform(nx / 4.0, 2.0 * nx) for _ in range(K)])
    wly = np.array([rng.uniform(ny / 4.0, 2.0 * ny) for _ in range(K)])
    phx = np.array([rng.uniform(0.0, 2.0 * math.pi) for _ in range(K)])
    phy = np.array([rng.uniform(0.0, 2.0 * math.pi) for _ in range(K)])
    vlay = np.array([rng.uniform(vmin, vmax) for _ in range(K + 1)], np.float32)
    blobs = [(rng.uniform(1, nx), rng.uniform(1, ny), rng.uniform(1, nz), rng.uniform(-blob_amp, blob_amp))
             for _ in range(n_blobs)]
    x = np.arange(1, nx + 1, dtype=np.float64)
    y = np.arange(1, ny + 1, dtype=np.float64)
    vp = np.empty((nx, ny, nz), np.float32, order="F")
    pert_xy = [np.exp(-((x[:, None] - bx) ** 2 + (y[None, :] - by) ** 2) / (2 * sigma * sigma))
               for (bx, by, _, _) in blobs]
    zk = np.stack([depth[k] + 0.5 * jitter * (np.sin(2 * math.pi * x[:, None] / wlx[k] + phx[k])
                                              + np.sin(2 * math.pi * y[None, :] / wly[k] + phy[k]))
                   for k in range(K)], axis=-1)  # (nx, ny, K)
    for kz in range(nz):
        zz = kz + 1.0
        layer = (zz >= zk).sum(axis=-1)
        v = vlay[layer].astype(np.float64)
        fac = np.ones_like(v)
        for (bxyz, pxy) in zip(blobs, pert_xy):
            fac += bxyz[3] * pxy * math.exp(-((zz - bxyz[2]) ** 2) / (2 * sigma * sigma))
        vp[:, :, kz] = (v * fac).astype(np.float32)
    rho = np.asfortranarray(gardner_rho(vp))
    return vp, rho


def patch_model(nx: int, nz: int):
    """The reference's absorption-test model (test/verify_patch.jl:16-30): 2000 m/s over 3000 m/s
    below a cosine interface at depth 100 +- 20, plus a 400 m/s Gaussian anomaly at depth 60;
    vs = vp/1.73, Gardner rho.  Returns (vp, vs, rho) as Float32 (nx, nz) arrays."""
    xc = (nx + 1) / 2
    vp = np.full((nx, nz), 2000.0, np.float32)
    for i in range(1, nx + 1):
        zint = 100 + int(np.rint(20 * math.cos(2 * math.pi * (i - xc) / nx)))
        vp[i - 1, zint - 1:] = 3000.0
    xg = (np.arange(1, nx + 1) - xc) / 30
    zg = (np.arange(1, nz + 1) - 60) / 15
    g = np.exp(-(xg[:, None] ** 2 + zg[None, :] ** 2)).astype(np.float32)  # Float32.(exp.(...))
    vp = (vp + np.float32(400.0) * g).astype(np.float32)
    vs = (vp / np.float32(1.73)).astype(np.float32)
    rho = (310.0 * vp.astype(np.float64) ** 0.25).astype(np.float32)
    return vp, vs, rho


def model_digest(*arrays) -> str:
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(np.asarray(a, np.float32).ravel(order="F")).tobytes())
    return h.hexdigest()[:16]
