"""trace_norm (src/utils/trace_norm.jl:23-32) on the host: closed forms on hand-made traces, a
per-trace loop in the element type, and Julia's argument rules (dims from 1, a dimension past
the matrix's own has length 1, maxima at or below epsilon replaced by 1, NaN propagating)."""
import numpy as np
import pytest


def loop_norm(a, eps=1e-10):
    """One trace (row) at a time: max |x| by a scalar loop, then x / max in the element type."""
    t = a.dtype.type
    out = np.empty_like(a)
    for r in range(a.shape[0]):
        m = t(0)
        for x in a[r]:
            m = max(m, abs(x))
        if m <= t(eps):
            m = t(1)
        for c in range(a.shape[1]):
            out[r, c] = a[r, c] / m
    return out


def test_closed_forms():
    import fomo
    a = np.array([[1, -4, 2], [0, 0, 0], [1e-12, -2e-12, 0], [-3, 1.5, 3]], np.float32)
    got = fomo.trace_norm(a)
    assert got.dtype == np.float32
    assert np.array_equal(got[0], np.array([0.25, -1, 0.5], np.float32))
    assert np.array_equal(got[1], a[1])   # silent trace: maximum 0 -> divided by 1
    assert np.array_equal(got[2], a[2])   # maximum 2e-12 <= 1e-10 -> left as it is
    assert np.array_equal(got[3], np.array([-1, 0.5, 1], np.float32))
    col = fomo.trace_norm(a, dims=1)
    assert np.array_equal(col[:, 0], a[:, 0] / np.float32(3))
    assert np.array_equal(col[:, 1], a[:, 1] / np.float32(4))
    assert np.array_equal(fomo.trace_norm(a, dims=(1, 2)), a / np.float32(4))
    # epsilon in the element type: 2e-12 is above 1e-12 (as Float32), so that trace is scaled
    assert np.array_equal(fomo.trace_norm(a, epsilon=1e-12)[2], a[2] / np.float32(2e-12))


def test_matches_a_per_trace_loop_bitwise():
    import fomo
    rng = np.random.default_rng(3)
    for dt in (np.float32, np.float64):
        a = (rng.normal(size=(17, 64)) * rng.uniform(1e-3, 1e3, (17, 1))).astype(dt)
        a[4] = 0
        a[9] *= dt(1e-13)
        got = fomo.trace_norm(a, dims=2)
        assert got.dtype == dt
        assert np.array_equal(got, loop_norm(a))
        assert np.array_equal(fomo.trace_norm(a.T, dims=1), loop_norm(a).T)
        big = np.abs(a).max(1) > 1e-10
        assert np.all(np.abs(got[big]).max(1) == 1)   # the maximum divides itself exactly


def test_julia_argument_rules():
    import fomo
    a = np.array([[2, -0.5], [np.nan, 1], [0, -0.0]], np.float32)
    got = fomo.trace_norm(a)
    assert np.all(np.isnan(got[1]))                  # NaN propagates through the maximum
    assert np.array_equal(fomo.trace_norm(a, dims=3)[0], np.array([1, -1], np.float32))  # own maximum
    assert np.array_equal(fomo.trace_norm(a, dims=3)[2], a[2])
    for bad in (0, -1, 2.0, True, ()):
        with pytest.raises(fomo.ArgumentError):
            fomo.trace_norm(a, dims=bad)
    with pytest.raises(fomo.ArgumentError):
        fomo.trace_norm(np.ones(5, np.float32))
    with pytest.raises(fomo.ArgumentError):
        fomo.trace_norm(np.ones((2, 2), np.int32))
    with pytest.raises(fomo.ArgumentError):
        fomo.trace_norm(np.ones((2, 0), np.float32))
