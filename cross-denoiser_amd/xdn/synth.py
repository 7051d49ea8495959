"""Seeded synthetic render streams for the per-frame filtering path.

The reference runs on a 61 GB EXR dataset that is not available here, so every workload uses this
generator instead (shapes and channel meanings of scripts/main.py:60-66 and utils.py:260-272):

    ref      sum of 8 Gaussian blobs (centers U[0,W]xU[0,H], sigma U[W/16,W/4], RGB amplitude
             U[0,0.5] each so radiance is in [0,4]) moving with a global pan
    color    ref x log-normal(0, sigma) noise, independently per pixel and channel
    color2   an independent second half-buffer of the same kind
    albedo   U[0,1]^3 per pixel
    normal   unit vectors of a smooth field moving with the pan
    linearZ  (Z, fwidth Z, 0) with Z in [1,50] from a smooth field moving with the pan;
             a random `disocclusion` fraction of the pixels gets Z x 2 in each frame
    pnFwidth U[0,0.05]^3;  mvec = (-pan_x/W, -pan_y/H, 0);  opacity = 1
    with composition=True also the terms the driver adds to the displayed frame and to the
    reference (main.py:344-345): emissive = one noise-free disc light (radiance 4, radius W/32)
    moving with the pan, envLight = a smooth sky term where the depth field is far (Z > 45);
    ref_emissive / ref_envLight equal them (both are noise-free)

Seeds are (seed, stream, frame): scene constants depend on (seed, stream) only, per-frame noise
on all three. Generation runs on any torch device; CPU generation feeds the oracle tests.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import torch


@dataclass
class StreamSpec:
    height: int
    width: int
    seed: int = 0
    stream: int = 0
    pan: tuple = (0.0, 0.0)          # pixels per frame; None -> drawn U[-pan_max, pan_max]
    pan_max: float = 0.0
    disocclusion: float = 0.0
    noise_sigma: float = 1.0
    blobs: int = 8
    _scene: dict = field(default=None, repr=False)


def _gen(device, *key) -> torch.Generator:
    h = 1469598103934665603
    for k in key:
        h = ((h ^ (int(k) & 0xFFFFFFFF)) * 1099511628211) & 0x7FFFFFFFFFFFFFFF
    g = torch.Generator(device=device)
    g.manual_seed(h)
    return g


def _scene(spec: StreamSpec, device) -> dict:
    g = _gen("cpu", spec.seed, spec.stream, 0xC0FFEE)
    H, W = spec.height, spec.width
    u = lambda *s: torch.rand(*s, generator=g, dtype=torch.float64)
    n = spec.blobs
    sc = {
        "cx": u(n) * W, "cy": u(n) * H,
        "sig": W / 16 + u(n) * (W / 4 - W / 16),
        "amp": u(n, 3) * 0.5,
        "lam": W / 4 + u(6) * (W - W / 4),
        "phi": u(6) * 2 * math.pi,
    }
    if spec.pan is None:
        sc["pan"] = ((u(2) * 2 - 1) * spec.pan_max).tolist()
    else:
        sc["pan"] = list(spec.pan)
    return sc


def scene_pan(spec: StreamSpec) -> tuple:
    return tuple(_scene(spec, "cpu")["pan"])


def make_frame(spec: StreamSpec, t: int, device="cpu", composition: bool = False) -> dict:
    """Frame t of the stream as a dict of [1,H,W,C] float32 tensors on `device` (with the
    emissive / envLight terms of the displayed frame when `composition` is set)."""
    sc = _scene(spec, device)
    H, W = spec.height, spec.width
    dev = torch.device(device)
    f64 = dict(device=dev, dtype=torch.float64)
    px, py = sc["pan"]
    ys = torch.arange(H, **f64).view(H, 1).expand(H, W)
    xs = torch.arange(W, **f64).view(1, W).expand(H, W)
    X, Y = xs - t * px, ys - t * py  # scene coordinates of each pixel
    ref = torch.zeros(H, W, 3, **f64)
    for k in range(len(sc["cx"])):
        d2 = (X - float(sc["cx"][k])) ** 2 + (Y - float(sc["cy"][k])) ** 2
        gk = torch.exp(-d2 / (2 * float(sc["sig"][k]) ** 2))
        ref += gk.unsqueeze(-1) * sc["amp"][k].to(dev).view(1, 1, 3)
    lam, phi = [float(v) for v in sc["lam"]], [float(v) for v in sc["phi"]]
    two_pi = 2 * math.pi
    nx = 0.5 * torch.sin(two_pi * X / lam[0] + phi[0])
    ny = 0.5 * torch.cos(two_pi * Y / lam[1] + phi[1])
    nz = torch.ones_like(nx)
    nrm = torch.stack([nx, ny, nz], -1)
    nrm = nrm / nrm.norm(dim=-1, keepdim=True)
    ax, ay = two_pi * X / lam[2] + phi[2], two_pi * Y / lam[3] + phi[3]
    Z = 25.5 + 24.5 * torch.sin(ax) * torch.cos(ay)
    dZx = 24.5 * torch.cos(ax) * torch.cos(ay) * two_pi / lam[2]
    dZy = -24.5 * torch.sin(ax) * torch.sin(ay) * two_pi / lam[3]
    fwZ = torch.maximum(dZx.abs(), dZy.abs())

    g = _gen(dev, spec.seed, spec.stream, t + 1)
    f32 = dict(device=dev, dtype=torch.float32, generator=g)
    noise1 = torch.exp(spec.noise_sigma * torch.randn(H, W, 3, **f32))
    noise2 = torch.exp(spec.noise_sigma * torch.randn(H, W, 3, **f32))
    albedo = torch.rand(H, W, 3, **f32)
    pnfw = torch.rand(H, W, 3, **f32) * 0.05
    if spec.disocclusion > 0:
        dis = torch.rand(H, W, **f32) < spec.disocclusion
        Z = torch.where(dis.to(torch.bool), Z * 2, Z)
    ref32 = ref.float()
    linz = torch.stack([Z.float(), fwZ.float(), torch.zeros(H, W, device=dev)], -1)
    mvec = torch.zeros(H, W, 3, device=dev)
    mvec[..., 0] = -px / W
    mvec[..., 1] = -py / H
    b = lambda a: a.unsqueeze(0).contiguous()
    extra = {}
    if composition:
        r2 = (X - float(sc["cx"][0])) ** 2 + (Y - float(sc["cy"][0])) ** 2
        emissive = (r2 <= (W / 32.0) ** 2).to(torch.float64).unsqueeze(-1).expand(H, W, 3) * 4.0
        sky = (Z > 45.0).to(torch.float64) * (0.5 + 0.25 * torch.sin(two_pi * X / W))
        env = sky.unsqueeze(-1) * torch.tensor([0.6, 0.8, 1.0], **f64).view(1, 1, 3)
        extra = {"emissive": b(emissive.float()), "envLight": b(env.float()),
                 "ref_emissive": b(emissive.float()), "ref_envLight": b(env.float())}
    return {
        **extra,
        "ref": b(ref32),
        "color": b(ref32 * noise1),
        "color2": b(ref32 * noise2),
        "albedo": b(albedo),
        "normal": b(nrm.float()),
        "mvec": b(mvec),
        "linearZ": b(linz),
        "pnFwidth": b(pnfw),
        "opacity": torch.ones(1, H, W, 1, device=dev),
    }


def zero_history(frame: dict) -> dict:
    """History of frame 0 (scripts/main.py:323-330): zeros for everything reprojected."""
    z = torch.zeros_like(frame["color"])
    return {"out_log": z, "ols1": z.clone(), "ols2": z.clone(), "linearZ": z.clone(), "normal": z.clone()}


def log_uniform_band(shape, lo=0.1, hi=10.0, seed=0, device="cpu") -> torch.Tensor:
    g = _gen(device, seed, 0xBA4D)
    u = torch.rand(*shape, generator=g, device=device)
    return torch.exp(math.log(lo) + u * (math.log(hi) - math.log(lo))).float()
