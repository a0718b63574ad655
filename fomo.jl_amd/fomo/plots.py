"""Figures of returned seismograms and wavefield snapshots (host only, headless).

Mirrors the reference's plotting extension (ext/FomoPlotsExt.jl): `plot_shot` (:36-70) saves a
shot record as a heatmap, `plot_wavefield_video` (:79-110) renders snapshots into an animation.
The inputs are the host arrays the solvers return: a seismogram `(n_rec, nt)` and a list of
`(nx, nz)` snapshots.  matplotlib is imported when a function is called (the reference loads its
extension only with Plots), with the non-interactive Agg backend, so nothing opens a window.
"""
from __future__ import annotations

import logging
import os
import shutil
import warnings

import numpy as np

_log = logging.getLogger(__name__)


def _pyplot():
    import matplotlib

    matplotlib.use("Agg", force=False)
    import matplotlib.pyplot as plt

    return plt


def plot_shot(shot, save_path):
    """Saves the shot record `shot` (receivers x time steps) as an 800x600 heatmap at `save_path`:
    receivers along x, time steps down, the `seismic` colour map over +-(98th percentile of |shot|)
    (the smallest positive number of the element type when that is zero).  NaN entries are drawn
    as gaps, with a warning and a title that says so; a record that holds only NaN is not drawn
    (an error is logged and nothing is written), as in the reference."""
    a = np.asarray(shot)
    if a.ndim != 2:
        raise ValueError(f"plot_shot expects a matrix, got {a.ndim} dimensions")
    valid = a[~np.isnan(a)] if np.issubdtype(a.dtype, np.floating) else a.ravel()
    has_nan = valid.size != a.size
    if has_nan:
        warnings.warn("NaN values detected in the shot record. Plotting with available data to identify issues...",
                      RuntimeWarning, stacklevel=2)
    if valid.size == 0:
        _log.error("The shot record contains only NaNs. Plotting aborted.")
        return None
    ft = a.dtype if np.issubdtype(a.dtype, np.floating) else np.dtype(np.float64)
    scale = ft.type(np.quantile(np.abs(valid).astype(ft), 0.98))
    if scale == 0:
        scale = np.finfo(ft).eps
    plt = _pyplot()
    fig, ax = plt.subplots(figsize=(8, 6), dpi=100)
    try:
        im = ax.imshow(a.T, cmap="seismic", vmin=-float(scale), vmax=float(scale), aspect="auto", origin="upper",
                       interpolation="nearest")
        ax.set_title("Shot Record (NaNs Detected)" if has_nan else "Shot Record")
        ax.set_xlabel("Receiver Number")
        ax.set_ylabel("Time Step")
        fig.colorbar(im, ax=ax)
        fig.savefig(save_path, dpi=100)
    finally:
        plt.close(fig)
    _log.info("Shot record successfully saved to: %s", save_path)
    return None


def plot_wavefield_video(snaps, snapshot_interval, save_path="wavefield.mp4", *, fps=10, adaptive_clims=False):
    """Renders the snapshots (each `(nx, nz)`: x along the horizontal, z down, equal aspect) into
    an animation at `save_path`, frame i titled with its step i * snapshot_interval (i from 1).
    Colour limits are +-half the largest |value| over all frames, or, with `adaptive_clims`, over
    each frame (1e-10 when that is below 1e-10).  The writer follows the extension: `.gif`
    (Pillow), `.html`, and any other (the reference's `.mp4`) through ffmpeg, which must then be
    on the PATH.  An empty list prints a message and writes nothing."""
    snaps = [np.asarray(s) for s in snaps] if snaps is not None else []
    if not snaps:
        print("No snapshots available for video creation.")
        return None
    if any(s.ndim != 2 for s in snaps):
        raise ValueError("plot_wavefield_video expects a list of matrices")
    ext = os.path.splitext(str(save_path))[1].lower()
    if ext not in (".gif", ".html") and shutil.which("ffmpeg") is None:
        raise RuntimeError(f"writing {save_path!r} needs ffmpeg, which is not on the PATH; "
                           "save as .gif or .html instead")

    def clims_of(frames):
        m = max(float(np.max(np.abs(f))) for f in frames)
        s = 1e-10 if m < 1e-10 else m * 0.5
        return -s, s

    global_clims = None if adaptive_clims else clims_of(snaps)
    plt = _pyplot()
    from matplotlib import animation

    fig, ax = plt.subplots(figsize=(8, 6), dpi=100)
    try:
        lo, hi = global_clims if global_clims else clims_of(snaps[:1])
        im = ax.imshow(snaps[0].T, cmap="seismic", vmin=lo, vmax=hi, aspect="equal", origin="upper",
                       interpolation="nearest")
        ax.set_xlabel("X (grid)")
        ax.set_ylabel("Z (grid)")
        fig.colorbar(im, ax=ax)

        def frame(i):
            s = snaps[i]
            im.set_data(s.T)
            im.set_clim(*(global_clims if global_clims else clims_of([s])))
            ax.set_title(f"VZ Wavefield (Step {(i + 1) * int(snapshot_interval)})")
            return (im,)

        anim = animation.FuncAnimation(fig, frame, frames=len(snaps), blit=False)
        writer = {".gif": "pillow", ".html": "html"}.get(ext, "ffmpeg")
        anim.save(str(save_path), writer=writer, fps=fps)
    finally:
        plt.close(fig)
    print(f"Wavefield video saved as {save_path}")
    return None
