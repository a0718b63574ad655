"""CPU-side checks of the product library: it loads, exports every declared symbol, and its
host-side setup helpers agree with the oracle.  No GPU compute is called here."""
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "fomo_hip.h")).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(fomo_\w+)\s*\(", text, flags=re.M)))


def test_library_exports_every_declared_symbol():
    import fomo._lib as L
    lib = L.lib()
    names = declared_symbols()
    assert len(names) >= 18
    for n in names:
        assert hasattr(lib, n), n
        assert n in L.SIGNATURES, f"{n} has no ctypes signature"
    assert lib.fomo_version().decode().startswith("fomo-mi355x")


def test_fd_coefficients_match_oracle(orc):
    import fomo
    for order in (2, 4, 6, 8, 10):
        assert np.array_equal(fomo.get_fd_coefficients(order), orc.fd_coeffs(order))
    with pytest.raises(ValueError):
        fomo.get_fd_coefficients(12)


@pytest.mark.parametrize("f0,dt,nt", [(15.0, 1e-3, 2000), (20.0, 4e-4, 4000), (12.0, 8e-4, 1500),
                                      (10.0, 1e-3, 1500), (7.3, 1.23e-3, 777)])
def test_ricker_matches_oracle_bitwise(orc, f0, dt, nt):
    import fomo
    assert np.array_equal(fomo.ricker_wavelet(f0, dt, nt), orc.ricker(f0, dt, nt))


def test_habc_coefficients_match_oracle(orc):
    import fomo
    for args in [(2500, 1e-3, 10), (1800, 4e-4, 5), (2000, 8e-4, 10)]:
        assert np.array_equal(np.float32(fomo.habc_coefficients(*args)), orc.habc_coeffs(*args, "f32"))


def test_validation_errors_raise_argument_error():
    """verify_patch.jl:263-267: CFL violation and out-of-range geometry throw ArgumentError
    before any device work."""
    import fomo
    vp = np.full((60, 40), 2000.0, np.float32)
    rho = np.full((60, 40), 2000.0, np.float32)
    kw = dict(sx=[30], sz=[10], rx=[5, 6], rz=[5, 5], nbc=10, fd_order=4, verbose=False)
    with pytest.raises(fomo.ArgumentError):
        fomo.acoustic2d(vp, rho, 10.0, 9e-3, 100, 10.0, **kw)
    with pytest.raises(fomo.ArgumentError):
        fomo.elastic2d(vp, vp / 1.73, rho, 10.0, 9e-3, 100, 10.0, **kw)
    with pytest.raises(fomo.ArgumentError):
        fomo.acoustic2d(vp, rho, 10.0, 1e-3, 100, 10.0, **{**kw, "sx": [65]})
    with pytest.raises(fomo.ArgumentError):
        fomo.acoustic2d(vp, rho, 10.0, 1e-3, 100, 10.0, **{**kw, "rz": [0, 5]})
    with pytest.raises(fomo.ArgumentError):
        fomo.acoustic2d(vp, rho, 10.0, 1e-3, 100, 10.0, **kw, boundary="pml")
    with pytest.raises(fomo.ArgumentError):
        fomo.acoustic2d(vp, rho, 10.0, 1e-3, 100, 10.0, **kw, wavelet=np.zeros(99))
    with pytest.raises(fomo.ArgumentError):
        fomo.acoustic2d(vp, rho, 10.0, 1e-3, 100, 10.0, **kw, storage="f8")
    with pytest.raises(fomo.ArgumentError):
        fomo.elastic2d(vp, vp / 1.73, rho, 10.0, 1e-3, 100, 10.0, **kw, storage=7)


def test_trace_misfit_rejects_null_arguments():
    """The C-ABI's argument checks run before any device work."""
    import ctypes as C

    import fomo._lib as L
    num, den = (C.c_double * 4)(), (C.c_double * 4)()
    assert L.lib().fomo_trace_misfit(None, 0, None, 0, 0, 0, 1, num, den) == L.FOMO_EINVAL
    assert "NULL" in L.lib().fomo_last_error().decode()


def test_stream_probe_rejects_bad_shapes():
    """The calibration entry points check their arguments before touching a device."""
    import ctypes as C

    import fomo._lib as L
    ms = C.c_double(0.0)
    for nr, nw, nb, it in ((0, 1, 1024, 1), (9, 1, 1024, 1), (2, 3, 1024, 1), (5, 0, 1024, 1), (5, 5, 1024, 1),
                           (5, 2, 2, 1), (5, 2, 1024, 0)):
        assert L.lib().fomo_stream_probe(nr, nw, nb, it, C.byref(ms)) == L.FOMO_EINVAL, (nr, nw, nb, it)
    assert L.lib().fomo_stream_probe(5, 2, 1024, 1, None) == L.FOMO_EINVAL
    assert L.lib().fomo_stream_copy(8, 1, C.byref(ms)) == L.FOMO_EINVAL


def test_grids_beyond_32bit_cell_offsets_are_rejected():
    """A 2D field (or a 3D plane) above 2^30 padded cells is refused before any device work: the
    sweeps address cells with 32-bit offsets."""
    import ctypes as C

    import fomo._lib as L
    from fomo.api import _desc
    h = C.c_void_p()
    for eq, nx, ny, nz in ((L.ACOUSTIC2D, 33000, 1, 33000), (L.ELASTIC2D, 70000, 1, 16000),
                           (L.ACOUSTIC3D, 33000, 33000, 8)):
        d = _desc(eq, nx, nz, 10, 8, 10, 10.0, 1e-3, 2000.0, 0, 0, 0, True, ny=ny)
        rc = L.lib().fomo_plan_create(C.byref(d), None, None, None, None, None, None, None, None, None, None,
                                      C.byref(h))
        assert rc == L.FOMO_EUNSUP, (eq, rc)
        assert "2^30" in L.lib().fomo_last_error().decode()


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU; --gpus N > 1 exits with a message before touching a device."""
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2", "--steps", "1"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in r.stderr


def test_cfl_bound_matches_reference_formula():
    """dt_max = dh / (vmax sqrt(2) sum|a|): just below passes, just above raises."""
    from fomo.checks import check_numerics
    import fomo
    a = fomo.get_fd_coefficients(8).astype(np.float64)
    dt_max = 10.0 / (3000.0 * np.sqrt(2.0) * np.abs(a).sum())
    check_numerics(3000.0, 3000.0, 10.0, dt_max * 0.999, 5.0, 8)
    with pytest.raises(fomo.ArgumentError):
        check_numerics(3000.0, 3000.0, 10.0, dt_max * 1.001, 5.0, 8)


def test_seeded_models_are_reproducible():
    from fomo.models import layered_2d, model_digest
    a = layered_2d(200, 80, seed=7)
    b = layered_2d(200, 80, seed=7)
    c = layered_2d(200, 80, seed=8)
    assert model_digest(*a) == model_digest(*b) != model_digest(*c)
    vp, vs, rho = a
    assert np.all(vp[:, 0] == 0) and np.all(rho[:, 0] == 0)
    assert 1800 <= vp[:, 1:].min() and vp.max() <= 4500


def test_storage_names_map_to_the_descriptor_codes():
    """storage = :f32 / :f16 / :bf16 / :mat16 (and their long names) map to FOMO_STORE_* codes 0-3;
    anything else is an ArgumentError before any device work."""
    import fomo
    from fomo import _lib
    from fomo.api import storage_code
    want = {"f32": 0, "float32": 0, ":f16": 1, "float16": 1, "bf16": 2, "bfloat16": 2, "mat16": 3, "f32_mat16": 3}
    for name, code in want.items():
        assert storage_code(name) == code, name
    assert _lib.STORE_F32_MAT16 == 3
    for bad in ("f8", "half", 4, -1):
        with pytest.raises(fomo.ArgumentError):
            storage_code(bad)
