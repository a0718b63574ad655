"""Post-processing of returned seismograms on the host.

`trace_norm` mirrors src/utils/trace_norm.jl:23-32, which the reference applies to the host
matrices its solvers return (example/acoustic2d_demo.jl:43).  It is elementwise work on a few MB
of traces after a run, so it stays in numpy, in the traces' own element type.  Misfits and
energies of traces still on the device are reduced there (`Plan.trace_misfit`).
"""
from __future__ import annotations

import numpy as np

from ._lib import ArgumentError


def trace_norm(data, dims=2, epsilon=1e-10):
    """Divides every trace by its maximum absolute value.

    `dims` counts from 1 as in Julia: for a seismogram `(n_rec, nt)`, `dims=2` normalises each
    receiver's trace over time and `dims=1` each time sample over the receivers.  A tuple reduces
    over several dimensions; a dimension past the second has length 1, so every element is then
    its own maximum.  Maxima at or below `epsilon` (converted to the element type first) are
    replaced by 1, so near-silent traces are left as they are.  NaN propagates, as in Julia's
    `maximum`.  The result has the input's element type."""
    a = np.asarray(data)
    if a.ndim != 2:
        raise ArgumentError(f"trace_norm expects a matrix, got {a.ndim} dimensions")
    if not np.issubdtype(a.dtype, np.floating):
        raise ArgumentError(f"trace_norm expects floating-point data, got {a.dtype}")
    dl = (dims,) if np.isscalar(dims) else tuple(dims)
    if not dl or any(isinstance(d, (bool, np.bool_)) or not isinstance(d, (int, np.integer)) or d < 1 for d in dl):
        raise ArgumentError(f"dimensions must be integers >= 1, got {dims}")
    axes = tuple(sorted({int(d) - 1 for d in dl if d <= 2}))
    t = a.dtype.type
    if axes:
        if any(a.shape[ax] == 0 for ax in axes):
            raise ArgumentError("trace_norm over an empty dimension")
        m = np.max(np.abs(a), axis=axes, keepdims=True)
    else:
        m = np.abs(a)
    m = np.where(m <= t(epsilon), t(1.0), m).astype(a.dtype)
    return (a / m).astype(a.dtype)
