"""Bandwidth network of the online training step.

``cnn5`` is the network of the benchmark configuration: five 3x3 convolutions
15 -> 32 -> 32 -> 32 -> 32 -> 6 with ReLU between them, BF16 operands with FP32 accumulation and
FP32 master weights, Keras glorot-uniform initialisation (scripts/main.py:79-118 uses Keras layers
with the default initialisers) and Keras Adam (libxdn.so xdn_adam).

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


class TorchCNN5:
    """cnn5 through torch convolutions (MIOpen), channels-last BF16."""

    def __init__(self, device, seed=1234):
        self.device = device
        self.params = []
        for w, b in init_params(seed=seed):
            self.params.append(w.permute(3, 2, 0, 1).contiguous().to(device).requires_grad_(True))  # OIHW
            self.params.append(b.to(device).requires_grad_(True))
        self.m = [torch.zeros_like(p) for p in self.params]
        self.v = [torch.zeros_like(p) for p in self.params]
        self.step_count = 0
        self._out = None

    def forward(self, ols1, ols2, prev_out, albedo, normal):
        x = torch.cat([ols1, ols2, prev_out, albedo, normal], -1)  # [B,H,W,15]
        h = x.permute(0, 3, 1, 2).to(torch.bfloat16)              # channels-last NCHW view
        n = len(self.params) // 2
        for i in range(n):
            w, b = self.params[2 * i], self.params[2 * i + 1]
            h = torch.nn.functional.conv2d(h, w.to(torch.bfloat16), b.to(torch.bfloat16), padding=1)
            if i < n - 1:
                h = torch.relu(h)
        out = h.float().permute(0, 2, 3, 1)
        self._out = out
        return out.detach().contiguous()

    def backward(self, dparam):
        self._out.backward(dparam)
        self._out = None

    def adam_step(self, lr, beta1=0.9, beta2=0.999, eps=1e-7):
        self.step_count += 1
        st = torch.cuda.current_stream(self.device).cuda_stream
        for p, m, v in zip(self.params, self.m, self.v):
            check(lib.xdn_adam(ptr(p.data), ptr(p.grad), ptr(m), ptr(v), p.numel(), self.step_count, lr, beta1,
                               beta2, eps, st), "adam")
            p.grad = None


def make(kind: str, device, seed=1234, shape=None):
    if kind == "cnn5":
        return TorchCNN5(device, seed=seed)
    raise ValueError(f"unknown network {kind!r}")
