"""The scene loader of the reference's converted dataset layout (xdn.data, scripts/utils.py)."""
import os

import numpy as np
import pytest
import torch


def _scene(tmp_path, frames=3, H=8, W=12):
    from xdn import data, synth
    spec = synth.StreamSpec(H, W, seed=2, pan=(1.0, 0.0))
    fr = [synth.make_frame(spec, t, composition=True) for t in range(frames)]
    data.save_npy_scene(str(tmp_path), fr, first=5)
    return fr


def test_npy_scene_roundtrip_cpu(tmp_path):
    from xdn import data
    fr = _scene(tmp_path)
    sc = data.NpyScene(str(tmp_path), device="cpu")
    assert len(sc) == 3 and sc.start_frame == 5
    idx, f = sc[4]  # wraps around: index % num_frames
    assert idx == 6
    for t in data.READ_TYPES:
        assert torch.equal(f[t], fr[1][t][..., :1] if t == "opacity" else fr[1][t]), t
    assert f["opacity"].shape == (1, 8, 12, 1)
    # max_num_frames drops the later frames of every type
    assert len(data.NpyScene(str(tmp_path), max_num_frames=7, device="cpu")) == 2


def test_npy_scene_cleans_and_checks_cpu(tmp_path):
    from xdn import data
    _scene(tmp_path)
    a = np.load(tmp_path / "color_0005.npy")
    a[0, 0, 0] = [np.nan, np.inf, -np.inf]
    np.save(tmp_path / "color_0005.npy", a)
    f = data.NpyScene(str(tmp_path), device="cpu")[0][1]
    assert torch.equal(f["color"][0, 0, 0], torch.zeros(3))  # NaN and +-inf -> 0 (utils.py:61-70)
    os.remove(tmp_path / "albedo_0006.npy")
    with pytest.raises(ValueError, match="different number of files"):
        data.NpyScene(str(tmp_path), device="cpu")
    for i in (5, 6, 7):
        p = tmp_path / f"normal_{i:04d}.npy"
        if p.exists():
            os.remove(p)
    with pytest.raises(FileNotFoundError, match="normal"):
        data.NpyScene(str(tmp_path), device="cpu")


@pytest.mark.gpu
def test_driver_on_a_saved_scene_equals_direct_frames(dev, tmp_path):
    """The driver fed from the .npy layout gives the same outputs and errors as the same frames
    passed directly."""
    from xdn import data, driver
    fr = _scene(tmp_path, frames=3, H=48, W=64)
    a = driver.FrameDriver(48, 64, device=dev, seed=3)
    b = driver.FrameDriver(48, 64, device=dev, seed=3)
    ea = a.run([{k: v.to(dev) for k, v in f.items()} for f in fr], log=None)
    eb = b.run((f for _, f in data.NpyScene(str(tmp_path), device=dev)), log=None)
    torch.cuda.synchronize()
    assert ea == eb
    assert torch.equal(a.shown, b.shown)


@pytest.mark.gpu
def test_driver_cli_on_an_exr_scene_writes_the_reference_images(dev, tmp_path, capsys):
    """`python -m xdn.driver --scene <.exr dir> --out-dir` (main.py --save_img): the same errors as the
    frames passed directly, one log line per frame, and olsavg/out/ref/path_{frame}.exr whose "out"
    is the displayed frame of the step."""
    from xdn import data, driver, exr
    fr = _scene(tmp_path / "npy", frames=2, H=32, W=48)
    data.save_exr_scene(str(tmp_path / "exr"), fr)
    direct = driver.FrameDriver(32, 48, device=dev, seed=1234)
    errs = direct.run([{k: v.to(dev) for k, v in f.items()} for f in fr], log=None)
    shown = direct.shown.cpu().numpy()[0]
    driver.main(["--scene", str(tmp_path / "exr"), "--frames", "2", "--out-dir", str(tmp_path / "out")])
    lines = [l for l in capsys.readouterr().out.splitlines() if l.startswith("[Frame")]
    assert len(lines) == 2 and [float(l.split()[3]) for l in lines] == [float(f"{e:.4f}") for e in errs]
    names = sorted(os.listdir(tmp_path / "out"))
    assert names == sorted(f"{n}_{i:04d}.exr" for n in ("olsavg", "out", "ref", "path") for i in range(2))
    np.testing.assert_array_equal(exr.read(str(tmp_path / "out" / "out_0001.exr")), np.maximum(shown, 0))
