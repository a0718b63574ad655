"""Host figures (ext/FomoPlotsExt.jl:36-110): the shot record's PNG, its NaN handling, and the
wavefield animation's frames.  CPU only."""
import os
import shutil

import numpy as np
import pytest

pytest.importorskip("matplotlib")
PIL = pytest.importorskip("PIL.Image")


def test_plot_shot_writes_an_800x600_png(tmp_path):
    import fomo
    rng = np.random.default_rng(3)
    seis = rng.normal(0, 1, (40, 300)).astype(np.float32)  # (n_rec, nt), as the solvers return
    out = tmp_path / "shot.png"
    assert fomo.plot_shot(fomo.trace_norm(seis, dims=2), str(out)) is None
    with PIL.open(out) as im:
        assert im.size == (800, 600)


def test_plot_shot_nan_handling(tmp_path):
    import fomo
    seis = np.ones((10, 50), np.float32)
    seis[3, 7] = np.nan
    out = tmp_path / "nan.png"
    with pytest.warns(RuntimeWarning, match="NaN"):
        fomo.plot_shot(seis, str(out))
    assert out.exists() and out.stat().st_size > 0
    all_nan = tmp_path / "all_nan.png"
    with pytest.warns(RuntimeWarning):
        fomo.plot_shot(np.full((4, 5), np.nan, np.float32), str(all_nan))
    assert not all_nan.exists()  # nothing to draw: logged, not raised, no file
    zero = tmp_path / "zero.png"
    fomo.plot_shot(np.zeros((4, 5), np.float32), str(zero))  # zero scale falls back to eps
    assert zero.exists()
    with pytest.raises(ValueError):
        fomo.plot_shot(np.zeros(5), str(tmp_path / "vec.png"))


@pytest.mark.parametrize("adaptive", [False, True])
def test_plot_wavefield_video_frames(tmp_path, adaptive):
    import fomo
    snaps = [np.sin(np.linspace(0, 3, 30 * 20)).reshape(30, 20).astype(np.float32) * (k + 1) for k in range(4)]
    out = tmp_path / "wave.gif"
    assert fomo.plot_wavefield_video(snaps, 50, str(out), fps=5, adaptive_clims=adaptive) is None
    with PIL.open(out) as im:
        assert im.n_frames == 4


def test_plot_wavefield_video_empty_and_mp4(tmp_path, capsys):
    import fomo
    assert fomo.plot_wavefield_video([], 10, str(tmp_path / "none.gif")) is None
    assert "No snapshots" in capsys.readouterr().out
    assert not os.path.exists(tmp_path / "none.gif")
    if shutil.which("ffmpeg") is None:
        with pytest.raises(RuntimeError, match="ffmpeg"):
            fomo.plot_wavefield_video([np.zeros((4, 4), np.float32)], 10, str(tmp_path / "v.mp4"))
