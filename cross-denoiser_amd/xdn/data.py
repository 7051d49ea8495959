"""Scene loaders for the reference's dataset layout (scripts/utils.py:146-305).

The reference reads one EXR per buffer type and frame, ``{type}_{frame:04d}.exr`` (its "default"
channel group, scripts/exr.py), and converts each to ``{type}_{frame:04d}.npy`` ([1,H,W,C] float32,
NaN and +-inf replaced by 0, utils.py:55-70) for faster loading. ``NpyScene`` reads the converted
layout (as ``DataLoader`` does unless ``--no_npy``), ``ExrScene`` the EXR files themselves (through
``xdn.exr``: scanline files with NO, RLE, ZIPS or ZIP compression), ``convert_exr_scene`` writes the
``.npy`` next to each ``.exr`` as utils.py:55-70 does. Both loaders make the same checks: every
buffer type present, equal file counts and frame indices per type (utils.py:115-139), frames below
``max_num_frames``, frame ``index % num_frames`` for an index past the end (utils.py:281-285).

Frames come back as dicts of [1,H,W,C] float32 device tensors, the shape ``xdn.pipeline`` takes;
``opacity`` is reduced to its first channel ([1,H,W,1]).
"""
from __future__ import annotations

import os
import re

import numpy as np
import torch

from . import exr

# buffer types the reference's driver reads (scripts/main.py:59-65)
READ_TYPES = ("ref", "color", "color2", "albedo", "normal", "mvec", "linearZ", "pnFwidth", "opacity",
              "emissive", "envLight", "ref_emissive", "ref_envLight")

_NAME = re.compile(r"^(?P<type>.+)_(?P<frame>\d+)\.(?P<ext>npy|exr)$")


def _clean(a: np.ndarray) -> np.ndarray:
    """[1,H,W,C] float32 with NaN and +-inf set to 0 (utils.py:55-70)."""
    a = np.asarray(a).astype(np.float32, copy=False)
    if a.ndim == 3:
        a = a[None]
    return np.nan_to_num(a, nan=0.0, posinf=0.0, neginf=0.0)


class NpyScene:
    """A directory of ``{type}_{frame:04d}.npy`` buffers."""

    EXT = "npy"

    def __init__(self, directory: str, types=READ_TYPES, max_num_frames: int | None = 101, device="cuda"):
        self.directory = directory
        self.types = tuple(types)
        self.device = torch.device(device)
        frames = {t: set() for t in self.types}
        for name in sorted(os.listdir(directory)):
            m = _NAME.match(name)
            if not m or m.group("ext") != self.EXT or m.group("type") not in frames:
                continue
            f = int(m.group("frame"))
            if max_num_frames is None or f < max_num_frames:
                frames[m.group("type")].add(f)
        missing = [t for t in self.types if not frames[t]]
        if missing:
            raise FileNotFoundError(f"{sorted(missing)} not given in directory: {directory}")
        counts = {t: len(v) for t, v in frames.items()}
        if len(set(counts.values())) != 1:
            raise ValueError(f"different number of files per type: {counts}")
        ref = frames[self.types[0]]
        for t, v in frames.items():
            if v != ref:
                raise ValueError(f"{t} has different frame indices than {self.types[0]}")
        self.frame_ids = sorted(ref)
        self.start_frame = self.frame_ids[0]
        self.num_frames = len(self.frame_ids)

    def __len__(self):
        return self.num_frames

    def path(self, t: str, frame: int) -> str:
        return os.path.join(self.directory, f"{t}_{frame:04d}.{self.EXT}")

    def read(self, t: str, frame: int) -> np.ndarray:
        return np.load(self.path(t, frame), allow_pickle=False)

    def load(self, t: str, frame: int) -> torch.Tensor:
        a = _clean(self.read(t, frame))
        if t == "opacity":
            a = a[..., :1]
        return torch.from_numpy(np.ascontiguousarray(a)).to(self.device)

    def __getitem__(self, index: int):
        """(frame number, frame dict) of the index-th frame (wrapping around like the reference)."""
        frame = self.frame_ids[index % self.num_frames]
        return frame, {t: self.load(t, frame) for t in self.types}

    def __iter__(self):
        for i in range(self.num_frames):
            yield self[i]


class ExrScene(NpyScene):
    """A directory of ``{type}_{frame:04d}.exr`` buffers, each read as its "default" channel group
    (utils.py:47-52)."""

    EXT = "exr"

    def read(self, t: str, frame: int) -> np.ndarray:
        return exr.read(self.path(t, frame), "default")


def convert_exr_scene(directory: str) -> int:
    """Writes ``{name}.npy`` next to every ``{name}.exr`` that has none (utils.py:55-70: the "default"
    group with a batch axis, NaN and +-inf set to 0); returns the number written."""
    n = 0
    for name in sorted(os.listdir(directory)):
        if not name.endswith(".exr"):
            continue
        npy = os.path.join(directory, name[:-4] + ".npy")
        if os.path.exists(npy):
            continue
        np.save(npy, _clean(exr.read(os.path.join(directory, name), "default")))
        n += 1
    return n


def save_exr_scene(directory: str, frames, first: int = 0, compression="ZIP_COMPRESSION"):
    """Writes frames (dicts of [1,H,W,C] tensors) as ``{type}_{frame:04d}.exr`` (FLOAT channels)."""
    os.makedirs(directory, exist_ok=True)
    for i, f in enumerate(frames):
        for t, v in f.items():
            a = v.detach().cpu().numpy().astype(np.float32)[0]
            exr.write(os.path.join(directory, f"{t}_{first + i:04d}.exr"), a,
                      channel_names=None if a.shape[-1] <= 4 else [f"ch{k}" for k in range(a.shape[-1])],
                      compression=compression)


def save_npy_scene(directory: str, frames, first: int = 0):
    """Writes frames (dicts of [1,H,W,C] tensors) in the converted-dataset layout."""
    os.makedirs(directory, exist_ok=True)
    for i, f in enumerate(frames):
        for t, v in f.items():
            np.save(os.path.join(directory, f"{t}_{first + i:04d}.npy"), v.detach().cpu().numpy().astype(np.float32))
