"""The reference's frame-driver loop (scripts/main.py:318-359) on the device.

Per frame: the online training step (xdn.pipeline.Denoiser: frame-0 zero history of
main.py:323-330, the history handoff of :338-341), then the report of :344-347 -- the displayed
frame out + emissive + envLight, the reference ref + ref_emissive + ref_envLight and their RelL2
(main.py:191-196) -- in one fused pass (xdn_frame_error), and the reference's log line
``[Frame %3d] relL2: %.4f (...)``. With an output directory, the images of main.py:350-359
(olsavg, out, ref, path) are collected per frame and written at the end as
``{name}_{frame:04d}.exr`` (ZIP-compressed FLOAT R, G, B; out/ols/ref clamped at zero,
main.py:150-167).

    python -m xdn.driver --height 1080 --width 1920 --frames 16 [--network unet] [--out-dir DIR]
    python -m xdn.driver --scene <dir of {type}_{frame:04d}.npy or .exr> [--frames 101]
"""
from __future__ import annotations

import os
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import torch

from . import pipeline
from ._lib import check, lib, ptr


def frame_error(out, ref, emissive=None, env_light=None, ref_emissive=None, ref_env_light=None, composed=None):
    """RelL2 per image of out + emissive + env_light against ref + ref_emissive + ref_env_light
    ([B,H,W,3] float32 device tensors; absent terms are zero). Returns err [B] (device) and writes
    the displayed frame into `composed` when given."""
    ts = [out, emissive, env_light, ref, ref_emissive, ref_env_light, composed]
    for t in ts:
        if t is not None and (t.dtype != torch.float32 or not t.is_contiguous() or t.shape != out.shape):
            raise ValueError("frame_error: contiguous float32 tensors of the same [B,H,W,3] shape expected")
    B, H, W, _ = out.shape
    err = torch.empty(B, device=out.device, dtype=torch.float32)
    ws_bytes = int(lib.xdn_frame_error_workspace_size(B, H, W))
    ws = torch.empty(max(ws_bytes, 4), device=out.device, dtype=torch.uint8)
    check(lib.xdn_frame_error(*[ptr(t) for t in ts], ptr(err), ptr(ws), ws_bytes, B, H, W,
                              torch.cuda.current_stream(out.device).cuda_stream), "frame_error")
    return err


class FrameDriver:
    """One stream through the reference's main loop: train step, history handoff, displayed frame
    and its RelL2 against the reference."""

    def __init__(self, H: int, W: int, device="cuda", network: str = "cnn5", seed: int = 1234,
                 lr: float = pipeline.LEARNING_RATE, use_graphs: bool = False):
        self.den = pipeline.Denoiser(H, W, device=device, seed=seed, lr=lr, use_graphs=use_graphs, network=network)
        self.shown = torch.empty((1, H, W, 3), device=self.den.device, dtype=torch.float32)
        self.err = torch.zeros(1, device=self.den.device, dtype=torch.float32)

    def step(self, frame: dict) -> dict:
        """One frame: the step's outputs plus the displayed frame ("shown") and its RelL2 ("relL2",
        a device tensor [1]). Frames without emissive / envLight terms show the filtered output."""
        out = self.den.step(frame)
        g = frame.get
        e = frame_error(out["out"], frame["ref"], g("emissive"), g("envLight"), g("ref_emissive"), g("ref_envLight"),
                        composed=self.shown)
        self.err.copy_(e)
        return {**out, "shown": self.shown, "relL2": self.err}

    def run(self, frames, log=print, first_index: int = 0):
        """Runs the frames in order, logging the reference's per-frame line; returns the errors."""
        errs = []
        for i, f in enumerate(frames):
            t0 = time.perf_counter()
            r = self.step(f)
            e = float(r["relL2"][0])  # synchronizes: the line reports the frame's wall time
            ms = (time.perf_counter() - t0) * 1e3
            errs.append(e)
            if log is not None:
                log(f"[Frame {first_index + i:3d}] relL2: {e:.4f} ({ms:.2f} ms)")
        return errs


def frame_images(frame: dict, result: dict) -> dict:
    """The images main.py:350-359 saves for one frame, as [H,W,3] float32 host arrays: olsavg (the
    two cross-regression outputs back from log space, averaged, plus emissive and envLight), out (the
    displayed frame), ref (with its emissive and envLight) and path (the mean of the two half-sample
    colours plus emissive and envLight)."""
    zero = torch.zeros_like(frame["color"])
    e, l = frame.get("emissive", zero), frame.get("envLight", zero)
    ols1 = torch.clamp(torch.expm1(result["ols1"]), min=0.0)
    ols2 = torch.clamp(torch.expm1(result["ols2"]), min=0.0)
    imgs = {"olsavg": (ols1 + ols2) * 0.5 + e + l, "out": result["shown"],
            "ref": frame["ref"] + frame.get("ref_emissive", zero) + frame.get("ref_envLight", zero),
            "path": (frame["color"] + frame["color2"]) * 0.5 + e + l}
    return {k: v[0].detach().float().cpu().numpy() for k, v in imgs.items()}


def write_image(path: str, img: np.ndarray) -> None:
    """main.py:150-167: out/ols/ref images clamped at zero; R, G, B names up to three channels, ch{i}
    beyond; ZIP compression, FLOAT."""
    from . import exr
    name = os.path.basename(path)
    if any(k in name for k in ("out", "ols", "ref")):
        img = np.maximum(0.0, img)
    names = None if img.shape[-1] <= 3 else [f"ch{i}" for i in range(img.shape[-1])]
    exr.write(path, img, channel_names=names, compression=exr.ZIP_COMPRESSION)


def write_images(out_dir: str, images, workers: int = 8) -> int:
    """Writes [(name, frame, image)] as {out_dir}/{name}_{frame:04d}.exr, highest frame first
    (main.py:169-184); returns the number of files."""
    os.makedirs(out_dir, exist_ok=True)
    jobs = sorted(((os.path.join(out_dir, f"{n}_{i:04d}.exr"), img) for n, i, img in images),
                  key=lambda j: int(j[0].rsplit("_", 1)[1].split(".")[0]), reverse=True)
    with ThreadPoolExecutor(max_workers=workers) as pool:
        list(pool.map(lambda j: write_image(*j), jobs))
    return len(jobs)


def main(argv=None):
    import argparse

    from . import synth
    ap = argparse.ArgumentParser(description="per-frame online denoising of a synthetic stream (main.py loop)")
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--frames", type=int, default=16)
    ap.add_argument("--network", choices=("cnn5", "unet"), default="cnn5")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--graphs", action="store_true", help="replay one captured step graph per frame")
    ap.add_argument("--scene", type=str, default=None,
                    help="directory of {type}_{frame:04d}.npy buffers (the reference's converted dataset), or of "
                         "the .exr files, instead of the synthetic stream")
    ap.add_argument("--out-dir", type=str, default=None,
                    help="write olsavg/out/ref/path_{frame:04d}.exr of every frame there (main.py --save_img)")
    a = ap.parse_args(argv)
    dev = torch.device("cuda:0")
    images = []
    if a.scene:
        from . import data
        try:
            scene = data.NpyScene(a.scene, max_num_frames=a.frames, device=dev)
        except FileNotFoundError:
            scene = data.ExrScene(a.scene, max_num_frames=a.frames, device=dev)
        first_id, first = scene[0]
        H, W = first["color"].shape[1:3]
        drv = FrameDriver(H, W, device=dev, network=a.network, use_graphs=a.graphs)
        frames = ((fid, f) for fid, f in scene)
    else:
        spec = synth.StreamSpec(a.height, a.width, seed=a.seed, pan=None, pan_max=3.0, disocclusion=0.10)
        drv = FrameDriver(a.height, a.width, device=dev, network=a.network, use_graphs=a.graphs)
        frames = ((t, synth.make_frame(spec, t, device=dev, composition=True)) for t in range(a.frames))
    for fid, f in frames:
        t0 = time.perf_counter()
        r = drv.step(f)
        e = float(r["relL2"][0])
        ms = (time.perf_counter() - t0) * 1e3
        print(f"[Frame {fid:3d}] relL2: {e:.4f} ({ms:.2f} ms)")
        if a.out_dir:
            images += [(k, fid, v) for k, v in frame_images(f, r).items()]
    if a.out_dir:
        print("Writing images...", end="", flush=True)
        write_images(a.out_dir, images)
        print("Done!")


if __name__ == "__main__":
    main()
