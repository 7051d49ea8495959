"""Bandwidth network of the online training step.

``cnn5`` is the network of the benchmark configuration: five 3x3 convolutions
15 -> 32 -> 32 -> 32 -> 32 -> 6 with ReLU between them, BF16 MFMA operands with FP32 accumulation,
FP32 master weights, Keras glorot-uniform initialisation (the reference's network uses Keras
defaults, scripts/main.py:79-118) and Keras Adam (xdn_adam). Forward, data and weight gradients
are the hand-written gfx950 kernels of libxdn.so (xdn_cnn5_*).

The head of the network (band = 1/(x^2 + 1e-4), alpha = sigmoid(x), main.py:110-113) is fused into
the filter kernel, so the network returns the raw 6-channel output.

``unet`` is the reference's own network (scripts/main.py:83-118, 31,414 parameters) on the same
calling convention (xdn_unet_*): 3x3 convolutions on BF16 MFMA, max pools, linear transposed 1x1
stride-2 convolutions and skip concatenations, the 1x1 8->6 head in FP32.
"""
from __future__ import annotations

import math

import torch

from ._lib import check, lib, ptr

CNN5_CHANNELS = (15, 32, 32, 32, 32, 6)


def glorot_uniform(shape_hwio, gen) -> torch.Tensor:
    kh, kw, ci, co = shape_hwio
    limit = math.sqrt(6.0 / (kh * kw * ci + kh * kw * co))
    return (torch.rand(shape_hwio, generator=gen, dtype=torch.float64) * 2 - 1) * limit


def init_params(channels=CNN5_CHANNELS, seed=1234):
    """[(w [3,3,cin,cout] HWIO float32, b [cout] float32)] on the CPU, deterministic in `seed`."""
    g = torch.Generator().manual_seed(seed)
    return [(glorot_uniform((3, 3, ci, co), g).float(), torch.zeros(co)) for ci, co in zip(channels[:-1], channels[1:])]


def flatten_params(layers) -> torch.Tensor:
    return torch.cat([t.reshape(-1) for wb in layers for t in wb])


def unflatten_params(flat: torch.Tensor, channels=CNN5_CHANNELS):
    out, o = [], 0
    for ci, co in zip(channels[:-1], channels[1:]):
        w = flat[o:o + 9 * ci * co].reshape(3, 3, ci, co)
        o += 9 * ci * co
        b = flat[o:o + co]
        o += co
        out.append((w, b))
    return out


# (name, type, ci, co) of the U-Net layers in Keras order (scripts/main.py:83-108); kernel shapes
# conv [3,3,ci,co], convT [co,ci] (Keras 1x1xCOxCI), head [ci,co]
UNET_LAYERS = (("c1_1", "conv", 15, 8), ("c1_2", "conv", 8, 8), ("c2_1", "conv", 8, 16),
               ("c2_2", "conv", 16, 16), ("c3_1", "conv", 16, 32), ("c3_2", "conv", 32, 32),
               ("c3_3", "convT", 32, 32), ("c4_1", "conv", 48, 16), ("c4_2", "conv", 16, 16),
               ("c4_3", "convT", 16, 16), ("c5_1", "conv", 24, 8), ("head", "head", 8, 6))


def _unet_shape(kind, ci, co):
    return {"conv": (3, 3, ci, co), "convT": (co, ci), "head": (ci, co)}[kind]


def unet_init_params(seed=1234):
    """[(w, b)] of the U-Net (float32, CPU): glorot-uniform kernels (Keras default; a 1x1 kernel's
    fans are its two channel counts), zero biases."""
    g = torch.Generator().manual_seed(seed)
    out = []
    for _, kind, ci, co in UNET_LAYERS:
        shp = _unet_shape(kind, ci, co)
        rf = 9 if kind == "conv" else 1
        limit = math.sqrt(6.0 / (rf * ci + rf * co))
        w = ((torch.rand(shp, generator=g, dtype=torch.float64) * 2 - 1) * limit).float()
        out.append((w, torch.zeros(co)))
    return out


def unet_unflatten(flat: torch.Tensor):
    out, o = [], 0
    for _, kind, ci, co in UNET_LAYERS:
        shp = _unet_shape(kind, ci, co)
        n = math.prod(shp)
        out.append((flat[o:o + n].reshape(shp), flat[o + n:o + n + co]))
        o += n + co
    return out


class _HipNet:
    """A bandwidth network on libxdn.so (xdn_<name>_*): parameters, Adam state, packed
    fragments and activations resident on the device."""

    name = ""

    def __init__(self, device, flat, shape):
        self.device = torch.device(device)
        n = int(self._fn("param_count")())
        assert flat.numel() == n, (flat.numel(), n)
        self.params = flat.to(self.device)
        self.grad = torch.zeros_like(self.params)
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.packed = torch.zeros(int(self._fn("packed_bytes")()), device=self.device, dtype=torch.uint8)
        self.step_count = 0
        self.resize(shape)
        self.pack()

    def _fn(self, what):
        return getattr(lib, f"xdn_{self.name}_{what}")

    def resize(self, shape):
        B, H, W = shape
        self.shape = (B, H, W)
        self.ws_bytes = int(self._fn("workspace_size")(B, H, W))
        self.ws = torch.empty(self.ws_bytes, device=self.device, dtype=torch.uint8)
        self.out = torch.empty((B, H, W, 6), device=self.device, dtype=torch.float32)
        self._inputs = None

    def _stream(self):
        return torch.cuda.current_stream(self.device).cuda_stream

    def pack(self):
        check(self._fn("pack")(ptr(self.params), ptr(self.packed), self._stream()), f"{self.name} pack")

    def forward(self, ols1, ols2, prev_out, albedo, normal):
        B, H, W = self.shape
        if tuple(ols1.shape[:3]) != (B, H, W):
            raise ValueError(f"{self.name} built for {self.shape}, got {tuple(ols1.shape)}")
        self._inputs = (ols1, ols2, prev_out, albedo, normal)
        check(self._fn("forward")(*[ptr(t) for t in self._inputs], ptr(self.packed), ptr(self.out), ptr(self.ws),
                                  self.ws_bytes, B, H, W, self._stream()), f"{self.name} forward")
        return self.out

    def backward(self, dparam):
        B, H, W = self.shape
        if self._inputs is None:
            raise RuntimeError(f"{self.name} backward without a forward")
        check(self._fn("backward")(*[ptr(t) for t in self._inputs], ptr(dparam), ptr(self.packed), ptr(self.grad),
                                   ptr(self.ws), self.ws_bytes, B, H, W, self._stream()), f"{self.name} backward")

    def adam_step(self, lr, beta1=0.9, beta2=0.999, eps=1e-7):
        self.step_count += 1
        check(lib.xdn_adam(ptr(self.params), ptr(self.grad), ptr(self.m), ptr(self.v), self.params.numel(),
                           self.step_count, lr, beta1, beta2, eps, self._stream()), "adam")
        self.pack()


class HipCNN5(_HipNet):
    """cnn5 on libxdn.so."""

    name = "cnn5"

    def __init__(self, device, seed=1234, shape=(1, 64, 64)):
        super().__init__(device, flatten_params(init_params(seed=seed)), shape)

    def layers(self):
        return unflatten_params(self.params)


class HipUNet(_HipNet):
    """The reference's U-Net on libxdn.so; height and width must be multiples of 4."""

    name = "unet"

    def __init__(self, device, seed=1234, shape=(1, 64, 64)):
        super().__init__(device, flatten_params(unet_init_params(seed=seed)), shape)

    def layers(self):
        return unet_unflatten(self.params)

    def workspace_buffer(self, index):
        """Workspace buffer `index` (0-12 activations c1_1 .. c5_1, 13-25 backward gradients) as a
        raw byte view."""
        import ctypes
        B, H, W = self.shape
        off, nb = ctypes.c_size_t(), ctypes.c_size_t()
        check(lib.xdn_unet_workspace_layout(B, H, W, index, ctypes.byref(off), ctypes.byref(nb)),
              "unet workspace layout")
        return self.ws[off.value:off.value + nb.value]


def make(kind: str, device, seed=1234, shape=(1, 64, 64)):
    if kind == "cnn5":
        return HipCNN5(device, seed=seed, shape=shape)
    if kind == "unet":
        return HipUNet(device, seed=seed, shape=shape)
    raise ValueError(f"unknown network {kind!r}")
