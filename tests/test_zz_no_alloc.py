"""fomo_trace_misfit_noalloc (the trace misfit with the library's static device scratch, for callers
that cannot allocate) against fomo_trace_misfit: the same bits on two calls in a row, the second
smaller, and a clear error beyond the scratch's size."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _plan(nx, nz, nt, n_rec, sx, seed):
    import fomo
    import fomo._lib as L
    from fomo.api import Plan
    from fomo.models import layered_2d
    vp, _, rho = layered_2d(nx, nz, seed=seed, n_interfaces=3, jitter=2.0, vacuum_top=False)
    rng = np.random.default_rng(seed)
    rx = rng.integers(1, nx + 1, n_rec)
    rz = rng.integers(1, nz + 1, n_rec)
    w = fomo.ricker_wavelet(15.0, 1e-3, nt)
    return Plan(L.ACOUSTIC2D, vp, rho, 10.0, 1e-3, nt, sx=[sx], sz=[nz // 2], rx=rx, rz=rz, wavelet=w[None], nbc=10,
                fd_order=8)


def _misfit(name, a, b, slot, t0, t1):
    import fomo._lib as L
    n = a.n_rec[0]
    num = np.full(n, np.nan)
    den = np.full(n, np.nan)
    L.check(getattr(L.lib(), name)(a._h, 0, b._h if b is not None else None, 0, slot, t0, t1, num.ctypes.data,
                                   den.ctypes.data))
    return num, den


def test_noalloc_misfit_equals_allocating_entry(gpu):
    # first call: 1500 samples (6 time chunks) x 900 receivers, two plans
    a = _plan(120, 90, 1500, 900, 40, seed=1)
    b = _plan(120, 90, 1500, 900, 41, seed=1)
    try:
        a.advance(1500)
        b.advance(1500)
        for slot in (0, 2):
            want = _misfit("fomo_trace_misfit", a, b, slot, 0, 1500)
            got = _misfit("fomo_trace_misfit_noalloc", a, b, slot, 0, 1500)
            assert np.all(want[1] > 0)
            assert np.array_equal(want[0], got[0]) and np.array_equal(want[1], got[1])
    finally:
        a.close()
        b.close()
    # second call, smaller: 37 receivers, a window of 300 samples (2 chunks), one plan (energy only)
    c = _plan(70, 50, 400, 37, 30, seed=2)
    try:
        c.advance(400)
        want = _misfit("fomo_trace_misfit", c, None, 0, 100, 400)
        got = _misfit("fomo_trace_misfit_noalloc", c, None, 0, 100, 400)
        assert np.any(want[0] > 0)
        assert np.array_equal(want[0], got[0]) and np.array_equal(want[1], got[1])
        # an empty window and the repeat of the first call still agree
        assert np.array_equal(_misfit("fomo_trace_misfit_noalloc", c, None, 0, 50, 50)[0], np.zeros(37))
        again = _misfit("fomo_trace_misfit_noalloc", c, None, 0, 100, 400)
        assert np.array_equal(again[0], got[0])
    finally:
        c.close()


def test_noalloc_misfit_rejects_sizes_beyond_its_scratch(gpu):
    """4096 samples are 16 time chunks: (16 + 1) x 2 x 61,700 receivers = 2,097,800 doubles, just
    over the 2^21 of the scratch; the allocating entry point still serves the call."""
    import fomo._lib as L
    p = _plan(40, 30, 4096, 61_700, 20, seed=3)
    try:
        with pytest.raises(L.ArgumentError, match="static scratch"):
            _misfit("fomo_trace_misfit_noalloc", p, None, 0, 0, 4096)
        num, den = _misfit("fomo_trace_misfit", p, None, 0, 0, 4096)
        assert np.all(num == 0) and np.all(den == 0)
    finally:
        p.close()
