"""Per-frame online denoising step (train_step, scripts/main.py:250-313) on libxdn.so.

Stage order per frame, all on one HIP stream, all buffers preallocated per stream slot:

  1. cross_regression_from_color  log1p + variance + regression + block recon -> ols1, ols2
  2. Reproject                    history [out_log, ols1, ols2] with depth/normal validity
  3. erode_temporal_screening     3x3 erosion of the mask fused into screening -> prev_out, ...
  4. bandwidth CNN forward        [ols1, ols2, prev_out, albedo, normal] -> raw param (6 ch)
  5. filter_train                 spatial filter + temporal accumulation + loss + backward to param
  6. CNN backward + Adam          (Keras Adam semantics)
  7. history handoff              out_log, ols1, ols2 (+ this frame's linearZ, normal), ping-pong

Labels follow main.py:283-286; frame-0 history is zeros (main.py:323-330).
"""
from __future__ import annotations

import torch

from . import cnn as cnn_mod
from . import ops
from ._lib import check, lib, ptr

REGRESSION_WIDTH = 17   # scripts/main.py:71-74
FILTER_WIDTH = 11
SPARSITY = 4
LEARNING_RATE = 1e-3


def filter_train(ols1, ols2, albedo, normal, param, success, success_prev, prev_out, opacity, prev_ols1, prev_ols2,
                 want_loss=False, win_size=FILTER_WIDTH):
    """Fused spatial filter + aggregation_loss + backward (allocating wrapper of xdn_filter_train).

    Returns (out_log, out, loss or None, avg_loss [B], dparam [B,H,W,6]).
    """
    ts = [t.contiguous() for t in (ols1, ols2, albedo, normal, param, success, success_prev, prev_out, opacity,
                                   prev_ols1, prev_ols2)]
    B, H, W, _ = ts[0].shape
    dev = ts[0].device
    out_log, out = torch.empty_like(ts[0]), torch.empty_like(ts[0])
    dparam = torch.empty((B, H, W, 6), device=dev, dtype=torch.float32)
    loss = torch.empty_like(ts[0]) if want_loss else None
    avg = torch.empty(B, device=dev, dtype=torch.float32)
    ws_bytes = lib.xdn_filter_train_workspace_size(B, H, W)
    ws = torch.empty(max(ws_bytes, 4) // 4, device=dev, dtype=torch.float32)
    check(lib.xdn_filter_train(*[ptr(t) for t in ts], ptr(out_log), ptr(out), ptr(dparam), ptr(loss), ptr(avg),
                               ptr(ws), ws_bytes, B, H, W, win_size, torch.cuda.current_stream(dev).cuda_stream),
          "filter_train")
    return out_log, out, loss, avg, dparam


class FrameBuffers:
    """All per-frame intermediates of one stream slot at one resolution (preallocated)."""

    def __init__(self, B: int, H: int, W: int, device):
        e3 = lambda: torch.empty((B, H, W, 3), device=device, dtype=torch.float32)
        e1 = lambda: torch.empty((B, H, W, 1), device=device, dtype=torch.float32)
        self.shape = (B, H, W)
        self.y1, self.y2, self.var1, self.var2 = e3(), e3(), e1(), e1()
        self.beta_bytes = lib.xdn_cross_regression_workspace_size(B, H, W, SPARSITY)
        self.beta = torch.empty(self.beta_bytes // 4, device=device, dtype=torch.float32)
        self.success = torch.empty((B, H, W, 1), device=device, dtype=torch.bool)
        self.rp = [e3(), e3(), e3()]
        self.prev_out, self.prev_ols1, self.prev_ols2 = e3(), e3(), e3()
        self.success_prev, self.succ_e = e1(), e1()
        self.net_in = torch.empty((B, H, W, 15), device=device, dtype=torch.float32)
        self.out = e3()
        self.dparam = torch.empty((B, H, W, 6), device=device, dtype=torch.float32)
        self.avg_loss = torch.empty(B, device=device, dtype=torch.float32)
        self.ws_bytes = lib.xdn_filter_train_workspace_size(B, H, W)
        self.ws = torch.empty(max(self.ws_bytes, 4) // 4, device=device, dtype=torch.float32)
        self.ones = torch.ones((B, H, W, 1), device=device, dtype=torch.float32)
        # history ping-pong: [out_log, ols1, ols2] x 2
        self.hist = [[e3(), e3(), e3()], [e3(), e3(), e3()]]


class Denoiser:
    """One render stream: resident history, bandwidth CNN and its Adam state."""

    def __init__(self, H: int, W: int, device="cuda", B: int = 1, seed: int = 1234, lr: float = LEARNING_RATE,
                 cnn: str = "cnn5"):
        self.device = torch.device(device)
        self.buf = FrameBuffers(B, H, W, self.device)
        self.net = cnn_mod.make(cnn, self.device, seed=seed, shape=(B, H, W))
        self.lr = lr
        self.t = 0
        self.prev_linz = torch.zeros((B, H, W, 3), device=self.device)
        self.prev_nrm = torch.zeros((B, H, W, 3), device=self.device)
        for h in self.buf.hist[1]:
            h.zero_()

    def history(self):
        """(out_log, ols1, ols2) of the previous frame."""
        return self.buf.hist[(self.t + 1) % 2]

    def step(self, frame: dict) -> dict:
        """One train_step on frame (dict of [B,H,W,C] float32 tensors, see xdn.synth)."""
        b = self.buf
        B, H, W = b.shape
        st = torch.cuda.current_stream(self.device).cuda_stream
        cur = b.hist[self.t % 2]           # written this frame
        prv = b.hist[(self.t + 1) % 2]     # previous frame's
        out_log, ols1, ols2 = cur
        # 1. cross-regression with the log1p prologue
        check(lib.xdn_cross_regression_from_color(
            ptr(frame["color"]), ptr(frame["color2"]), ptr(frame["albedo"]), ptr(frame["normal"]), ptr(b.y1),
            ptr(b.y2), ptr(ols1), ptr(ols2), ptr(b.var1), ptr(b.var2), ptr(b.beta), b.beta_bytes, B, H, W,
            REGRESSION_WIDTH, SPARSITY, st), "cross_regression")
        # 2. reprojection of the history (main.py:266: opacity = prev_opacity = ones)
        import ctypes
        ins = (ctypes.c_void_p * 3)(*[t.data_ptr() for t in prv])
        outs = (ctypes.c_void_p * 3)(*[t.data_ptr() for t in b.rp])
        dims = (ctypes.c_int * 3)(3, 3, 3)
        check(lib.xdn_reproject(ptr(frame["mvec"]), 3, ins, dims, ptr(frame["linearZ"]), ptr(self.prev_linz),
                                ptr(frame["normal"]), ptr(self.prev_nrm), ptr(frame["pnFwidth"]), ptr(b.ones),
                                ptr(b.ones), ptr(b.success), outs, B, H, W, st), "reproject")
        # 3. erosion + temporal screening (win_curr 7, win_prev 1, main.py:280)
        check(lib.xdn_erode_temporal_screening(
            ptr(ols1), ptr(ols2), ptr(b.rp[0]), ptr(b.rp[1]), ptr(b.rp[2]), ptr(b.success), ptr(b.succ_e),
            ptr(b.prev_out), ptr(b.prev_ols1), ptr(b.prev_ols2), ptr(b.success_prev), B, H, W, 7, 1, st),
            "erode_temporal_screening")
        # 4. network forward on [ols1, ols2, prev_out, albedo, normal] (main.py:291)
        param = self.net.forward(ols1, ols2, b.prev_out, frame["albedo"], frame["normal"])
        # 5. fused filter + aggregation + loss + backward to param
        check(lib.xdn_filter_train(
            ptr(ols1), ptr(ols2), ptr(frame["albedo"]), ptr(frame["normal"]), ptr(param), ptr(b.succ_e),
            ptr(b.success_prev), ptr(b.prev_out), ptr(frame["opacity"]), ptr(b.prev_ols1), ptr(b.prev_ols2),
            ptr(out_log), ptr(b.out), ptr(b.dparam), None, ptr(b.avg_loss), ptr(b.ws), b.ws_bytes, B, H, W,
            FILTER_WIDTH, st), "filter_train")
        # 6. network backward + Adam
        self.net.backward(b.dparam)
        self.net.adam_step(self.lr)
        # 7. history handoff (main.py:338-341)
        self.prev_linz.copy_(frame["linearZ"])
        self.prev_nrm.copy_(frame["normal"])
        self.t += 1
        return {"out": b.out, "out_log": out_log, "ols1": ols1, "ols2": ols2, "loss": b.avg_loss}
