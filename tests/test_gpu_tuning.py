"""The autotuners' choices (fomo_plan_tuning): timed by the first plan of a problem, reused by later
plans of the process and, through the tuning file, by a process that reads it; every layout the
tuners can pick gives the same bits.  FOMO_TUNE_CACHE names the file; switching it inside the
process starts from that file's records, which stands in for a later process here."""
import os
import shutil

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _plan2d(**kw):
    import fomo
    from fomo import _lib
    nx, nz, nt = 180, 120, 40
    vp = np.full((nx, nz), 3000.0, np.float32)
    vp[:, 60:] = 3500.0
    vs = (vp / np.float32(1.73)).astype(np.float32)
    rho = np.full((nx, nz), 2100.0, np.float32)
    w = fomo.ricker_wavelet(15.0, 1e-3, nt)
    rx = np.arange(1, nx + 1, 2)
    p = fomo.Plan(_lib.ELASTIC2D, vp, rho, 10.0, 1e-3, nt, vs=vs, sx=[90], sz=[10], rx=rx, rz=np.full(len(rx), 4),
                  wavelet=w[None], nbc=20, **kw)
    p.advance(nt)
    out = p.seis(2)
    t = p.tuning()
    p.close()
    return t, out


def _plan3d(**kw):
    import fomo
    from fomo import _lib
    n, nt = 40, 24
    vp = np.full((n, n, n), 2500.0, np.float32)
    vp[:, :, n // 2:] = 3000.0
    rho = np.full((n, n, n), 2000.0, np.float32)
    w = fomo.ricker_wavelet(15.0, 1e-3, nt)
    r = np.arange(1, n + 1, 4)
    gx, gy = np.meshgrid(r, r, indexing="ij")
    p = fomo.Plan(_lib.ACOUSTIC3D, vp, rho, 10.0, 1e-3, nt, sx=[20], sy=[20], sz=[5], rx=gx.ravel(), ry=gy.ravel(),
                  rz=np.full(gx.size, 3), wavelet=w[None], nbc=10, **kw)
    p.advance(nt)
    out = p.seis(3)
    t = p.tuning()
    p.close()
    return t, out


def test_choice_is_timed_then_reused_from_the_process_and_the_file(gpu, tmp_path, monkeypatch):
    a = tmp_path / "a.txt"
    monkeypatch.setenv("FOMO_TUNE_CACHE", str(a))
    t1, s1 = _plan2d()
    u1, q1 = _plan3d()
    assert t1["source"] == "timed" and u1["source"] == "timed", (t1, u1)
    t2, s2 = _plan2d()
    u2, q2 = _plan3d()
    assert t2["source"] == "process" and u2["source"] == "process"
    assert (t2["family"], t2["shape"]) == (t1["family"], t1["shape"])
    assert (u2["family"], u2["shape"], u2["extent"]) == (u1["family"], u1["shape"], u1["extent"])
    lines = [ln for ln in a.read_text().splitlines() if ln.strip()]
    assert len(lines) == 2, lines
    assert any(" 2d eq=2 " in ln for ln in lines) and any(" 3d eq=3 " in ln for ln in lines)
    # a process starting from this file: the records come from it
    b = tmp_path / "b.txt"
    shutil.copy(a, b)
    with open(b, "a") as f:
        f.write("a line that is no record\n")
    monkeypatch.setenv("FOMO_TUNE_CACHE", str(b))
    t3, s3 = _plan2d()
    u3, q3 = _plan3d()
    assert t3["source"] == "file" and u3["source"] == "file", (t3, u3)
    assert (t3["family"], t3["shape"]) == (t1["family"], t1["shape"])
    assert (u3["family"], u3["shape"], u3["extent"]) == (u1["family"], u1["shape"], u1["extent"])
    for x, y, z in zip(s1, s2, s3):
        assert np.array_equal(x, y) and np.array_equal(x, z)
    for x, y, z in zip(q1, q2, q3):
        assert np.array_equal(x, y) and np.array_equal(x, z)


def test_records_of_unknown_shapes_are_timed_again(gpu, tmp_path, monkeypatch):
    a = tmp_path / "a.txt"
    monkeypatch.setenv("FOMO_TUNE_CACHE", str(a))
    t1, _ = _plan3d()
    assert t1["source"] == "timed"
    key = a.read_text().splitlines()[0].split(" = ")[0]
    c = tmp_path / "c.txt"
    c.write_text(f"{key} = 1 9 9\n")
    monkeypatch.setenv("FOMO_TUNE_CACHE", str(c))
    t2, _ = _plan3d()
    assert t2["source"] == "timed"


def test_off_keeps_no_file_and_fixed_shapes_are_not_tuned(gpu, tmp_path, monkeypatch):
    monkeypatch.setenv("FOMO_TUNE_CACHE", "off")
    t, _ = _plan2d()
    assert t["source"] in ("timed", "process")
    assert not any(tmp_path.iterdir())
    # a forced elastic2d shape is one of the fused interior's strip lengths, or of the two sweeps'
    t, _ = _plan2d(tile_cfg=2)
    assert (t["source"], t["family"], t["shape"], t["extent"]) == ("fixed", "fused", 1, 32), t
    monkeypatch.setenv("FOMO_2D_UNFUSED", "1")
    t, _ = _plan2d(tile_cfg=2)
    assert (t["source"], t["family"], t["shape"], t["extent"]) == ("fixed", "sweeps", 1, 32), t
    monkeypatch.delenv("FOMO_2D_UNFUSED")
    u, _ = _plan3d(tile_cfg=3)
    assert u["source"] == "fixed" and u["shape"] == 2


@pytest.mark.parametrize("tile", [1, 2, 3])
def test_every_3d_layout_gives_the_same_bits(gpu, monkeypatch, tile):
    """Segment lengths and families the 3D tuner chooses between change no bits."""
    monkeypatch.setenv("FOMO_TUNE_CACHE", "off")
    _, ref = _plan3d()
    _, q = _plan3d(tile_cfg=tile)
    monkeypatch.setenv("FOMO_3D_UNFUSED", "1")
    _, r = _plan3d(tile_cfg=tile)
    for a, b, c in zip(ref, q, r):
        assert np.array_equal(a, b) and np.array_equal(a, c)
