"""Bandwidth network of the online training step.

``cnn5`` is the network of the benchmark configuration: five 3x3 convolutions
15 -> 32 -> 32 -> 32 -> 32 -> 6 with ReLU between them, BF16 MFMA operands with FP32 accumulation,
FP32 master weights, Keras glorot-uniform initialisation (the reference's network uses Keras
defaults, scripts/main.py:79-118) and Keras Adam (xdn_adam). Forward, data and weight gradients
are the hand-written gfx950 kernels of libxdn.so (xdn_cnn5_*).

The head of the network (band = 1/(x^2 + 1e-4), alpha = sigmoid(x), main.py:110-113) is fused into
the filter kernel, so the network returns the raw 6-channel output.
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


class HipCNN5:
    """cnn5 on libxdn.so: parameters, Adam state, packed bf16 fragments and activations resident."""

    def __init__(self, device, seed=1234, shape=(1, 64, 64)):
        self.device = torch.device(device)
        n = int(lib.xdn_cnn5_param_count())
        flat = flatten_params(init_params(seed=seed))
        assert flat.numel() == n
        self.params = flat.to(self.device)
        self.grad = torch.zeros_like(self.params)
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.packed = torch.empty(int(lib.xdn_cnn5_packed_bytes()), device=self.device, dtype=torch.uint8)
        self.step_count = 0
        self.resize(shape)
        self.pack()

    def resize(self, shape):
        B, H, W = shape
        self.shape = (B, H, W)
        self.ws_bytes = int(lib.xdn_cnn5_workspace_size(B, H, W))
        self.ws = torch.empty(self.ws_bytes, device=self.device, dtype=torch.uint8)
        self.out = torch.empty((B, H, W, 6), device=self.device, dtype=torch.float32)
        self._inputs = None

    def _stream(self):
        return torch.cuda.current_stream(self.device).cuda_stream

    def pack(self):
        check(lib.xdn_cnn5_pack(ptr(self.params), ptr(self.packed), self._stream()), "cnn5 pack")

    def forward(self, ols1, ols2, prev_out, albedo, normal):
        B, H, W = self.shape
        if tuple(ols1.shape[:3]) != (B, H, W):
            raise ValueError(f"cnn5 built for {self.shape}, got {tuple(ols1.shape)}")
        self._inputs = (ols1, ols2, prev_out, albedo, normal)
        check(lib.xdn_cnn5_forward(*[ptr(t) for t in self._inputs], ptr(self.packed), ptr(self.out), ptr(self.ws),
                                   self.ws_bytes, B, H, W, self._stream()), "cnn5 forward")
        return self.out

    def backward(self, dparam):
        B, H, W = self.shape
        if self._inputs is None:
            raise RuntimeError("cnn5 backward without a forward")
        check(lib.xdn_cnn5_backward(*[ptr(t) for t in self._inputs], ptr(dparam), ptr(self.packed), ptr(self.grad),
                                    ptr(self.ws), self.ws_bytes, B, H, W, self._stream()), "cnn5 backward")

    def adam_step(self, lr, beta1=0.9, beta2=0.999, eps=1e-7):
        self.step_count += 1
        check(lib.xdn_adam(ptr(self.params), ptr(self.grad), ptr(self.m), ptr(self.v), self.params.numel(),
                           self.step_count, lr, beta1, beta2, eps, self._stream()), "adam")
        self.pack()

    def layers(self):
        return unflatten_params(self.params)


def make(kind: str, device, seed=1234, shape=(1, 64, 64)):
    if kind == "cnn5":
        return HipCNN5(device, seed=seed, shape=shape)
    raise ValueError(f"unknown network {kind!r}")
