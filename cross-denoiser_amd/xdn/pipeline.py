"""Per-frame online denoising step (train_step, scripts/main.py:250-313) on libxdn.so.

A render stream is a *slot* (xdn_slot in include/xdn.h): its history, bandwidth network, Adam
moments and scratch live in HBM for the stream's lifetime. ``xdn_denoise_step`` advances any number
of slots (up to 16, mixed resolutions) by one frame with one launch per stage:

  1. cross-regression with the log1p prologue   -> ols1, ols2
  2. reprojection of [out_log, ols1, ols2]      (opacities one, main.py:266)
  3. 3x3 erosion + temporal screening           (win_curr 7, win_prev 1)
  4. bandwidth network forward on [ols1, ols2, prev_out, albedo, normal]: cnn5 (the benchmark's
     5-layer CNN) or the reference's U-Net (network="unet")
  5. fused spatial filter + temporal accumulation + loss + backward to the network output
  6. network backward, Keras Adam, bf16 fragment repack
  7. history handoff (main.py:338-341): out_log/ols1/ols2 ping-pong, the frame's linearZ/normal

Labels follow main.py:283-286; frame-0 history is zeros (main.py:323-330).
"""
from __future__ import annotations

import ctypes

import torch

from . import cnn as cnn_mod
from ._lib import check, lib, ptr

REGRESSION_WIDTH = 17   # scripts/main.py:71-74
FILTER_WIDTH = 11
SPARSITY = 4
LEARNING_RATE = 1e-3

SLOT_BUFFERS = ("y1", "y2", "var1", "var2", "beta", "success", "rp_out", "rp_ols1", "rp_ols2", "prev_out",
                "prev_ols1", "prev_ols2", "success_prev", "succ_e", "dparam", "param", "ones", "loss_parts",
                "cnn_ws", "adam_alpha")


class XdnSlot(ctypes.Structure):
    """Mirror of xdn_slot (include/xdn.h)."""
    _fields_ = ([("height", ctypes.c_int), ("width", ctypes.c_int)] +
                [(n, ctypes.c_void_p) for n in (
                    "color", "color2", "albedo", "normal", "mvec", "linear_z", "pn_fwidth", "opacity",
                    "prev_out_log", "prev_ols1", "prev_ols2", "prev_linear_z", "prev_normal",
                    "out_log", "ols1", "ols2", "out", "scratch", "params", "grad", "adam_m", "adam_v", "packed")] +
                [("adam_step", ctypes.c_int), ("avg_loss", ctypes.c_void_p), ("adam_t", ctypes.c_void_p),
                 ("network", ctypes.c_int)])

NETWORKS = {"cnn5": 0, "unet": 1}  # XDN_NET_CNN5, XDN_NET_UNET


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


def slot_bytes(H: int, W: int, network: str = "cnn5") -> int:
    """HBM resident per stream slot: scratch, history, outputs and network state."""
    hist = 7 * H * W * 3 * 4  # 2 x (out_log, ols1, ols2) + out
    fn = lambda what: getattr(lib, f"xdn_{network}_{what}")
    net = 4 * int(fn("param_count")()) * 4 + int(fn("packed_bytes")()) + 4
    return int(lib.xdn_slot_scratch_bytes(H, W)) + hist + net


class Denoiser:
    """One render stream: resident history, bandwidth network (cnn5 or unet) and its Adam state."""

    def __init__(self, H: int, W: int, device="cuda", seed: int = 1234, lr: float = LEARNING_RATE,
                 use_graphs: bool = False, network: str = "cnn5"):
        if network not in NETWORKS:
            raise ValueError(f"unknown network {network!r} (one of {sorted(NETWORKS)})")
        self.network = network
        if H % SPARSITY or W % SPARSITY or H <= REGRESSION_WIDTH // 2 or W <= REGRESSION_WIDTH // 2:
            raise ValueError(f"stream size {H}x{W}: height and width must be divisible by {SPARSITY} and > 8")
        self.device = torch.device(device)
        self.H, self.W = H, W
        self.lr = lr
        dev = self.device
        e3 = lambda: torch.zeros((1, H, W, 3), device=dev, dtype=torch.float32)
        self.scratch = torch.empty(int(lib.xdn_slot_scratch_bytes(H, W)), device=dev, dtype=torch.uint8)
        self.hist = [[e3(), e3(), e3()], [e3(), e3(), e3()]]  # [out_log, ols1, ols2] ping-pong
        self.out = e3()
        layers = cnn_mod.unet_init_params(seed=seed) if network == "unet" else cnn_mod.init_params(seed=seed)
        self.params = cnn_mod.flatten_params(layers).to(dev)
        self.grad = torch.zeros_like(self.params)
        self.adam_m = torch.zeros_like(self.params)
        self.adam_v = torch.zeros_like(self.params)
        self.packed = torch.zeros(int(getattr(lib, f"xdn_{network}_packed_bytes")()), device=dev, dtype=torch.uint8)
        self.avg_loss = torch.zeros(1, device=dev, dtype=torch.float32)
        self.adam_t = torch.zeros(1, device=dev, dtype=torch.int32)  # Adam updates done (device counter)
        self.use_graphs = use_graphs
        self._graphs = {}
        self._cap_stream = None
        self.zeros3 = e3()
        self.prev_frame = None  # G-buffers of the previous frame (linearZ, normal)
        self.t = 0
        offs = (ctypes.c_size_t * len(SLOT_BUFFERS))()
        check(lib.xdn_slot_scratch_offsets(H, W, ctypes.addressof(offs)), "slot offsets")
        self._offsets = dict(zip(SLOT_BUFFERS, [int(o) for o in offs]))
        slot = self.make_slot(None)
        check(lib.xdn_slot_init(ctypes.addressof(slot), self._stream()), "slot init")

    def _stream(self):
        return torch.cuda.current_stream(self.device).cuda_stream

    def buffer(self, name: str, channels: int, dtype=torch.float32) -> torch.Tensor:
        """View of an intermediate of the last step inside the slot scratch (see SLOT_BUFFERS)."""
        o = self._offsets[name]
        n = self.H * self.W * channels * torch.empty(0, dtype=dtype).element_size()
        return self.scratch[o:o + n].view(dtype).view(1, self.H, self.W, channels)

    def history(self):
        """(out_log, ols1, ols2) of the previous frame."""
        return self.hist[(self.t + 1) % 2]

    def make_slot(self, frame, t=None) -> XdnSlot:
        """Slot of step `t` (default: the next step) for `frame`. For a step other than the next one
        only the cross-regression stage's fields (frame inputs, ols1/ols2, scratch) are meaningful."""
        t = self.t if t is None else t
        s = XdnSlot()
        s.height, s.width = self.H, self.W
        if frame is not None:
            for k, name in (("color", "color"), ("color2", "color2"), ("albedo", "albedo"), ("normal", "normal"),
                            ("mvec", "mvec"), ("linearZ", "linear_z"), ("pnFwidth", "pn_fwidth"),
                            ("opacity", "opacity")):
                ten = frame[k]
                if ten.device != self.device or ten.dtype != torch.float32 or not ten.is_contiguous():
                    raise ValueError(f"frame[{k!r}] must be a contiguous float32 tensor on {self.device}")
                if tuple(ten.shape[:3]) != (1, self.H, self.W):
                    raise ValueError(f"frame[{k!r}] has shape {tuple(ten.shape)}, stream is {self.H}x{self.W}")
                setattr(s, name, ten.data_ptr())
        prv, cur = self.hist[(t + 1) % 2], self.hist[t % 2]
        s.prev_out_log, s.prev_ols1, s.prev_ols2 = (t.data_ptr() for t in prv)
        pf = self.prev_frame
        s.prev_linear_z = (pf["linearZ"] if pf is not None else self.zeros3).data_ptr()
        s.prev_normal = (pf["normal"] if pf is not None else self.zeros3).data_ptr()
        s.out_log, s.ols1, s.ols2 = (t.data_ptr() for t in cur)
        s.out = self.out.data_ptr()
        s.scratch = self.scratch.data_ptr()
        s.params, s.grad, s.adam_m, s.adam_v = (t.data_ptr() for t in (self.params, self.grad, self.adam_m,
                                                                        self.adam_v))
        s.packed = self.packed.data_ptr()
        s.adam_step = t + 1
        s.avg_loss = self.avg_loss.data_ptr()
        s.adam_t = self.adam_t.data_ptr()
        s.network = NETWORKS[self.network]
        return s

    def repack(self):
        """Rebuild the packed bf16 fragments after the parameters were replaced."""
        check(getattr(lib, f"xdn_{self.network}_pack")(ptr(self.params), ptr(self.packed), self._stream()),
              f"{self.network} pack")

    def set_step(self, t: int):
        """Resume at update t (after the parameters and Adam moments were restored)."""
        self.t = int(t)
        self.adam_t.fill_(self.t)

    def advance(self, frame):
        """Host bookkeeping after the step of `frame` was enqueued."""
        self.prev_frame = frame
        self.t += 1

    def outputs(self) -> dict:
        cur = self.hist[(self.t + 1) % 2]  # written by the last step
        return {"out": self.out, "out_log": cur[0], "ols1": cur[1], "ols2": cur[2], "loss": self.avg_loss}

    def step(self, frame: dict) -> dict:
        """One train_step on `frame` (dict of [1,H,W,C] float32 tensors, see xdn.synth).

        The frame's linearZ and normal are kept (not copied) as the next frame's history: do not
        modify them in place before the next step has run.
        """
        if self.use_graphs:
            self._graph_for(frame).launch(self._stream())
        else:
            slot = self.make_slot(frame)
            check(lib.xdn_denoise_step(ctypes.addressof(slot), 1, self.lr, self._stream()), "denoise_step")
        self.advance(frame)
        return self.outputs()

    def graph_key(self, frame) -> tuple:
        """What the next step's slot depends on besides this stream's own buffers: the history
        parity, the frame's buffers and the previous frame's G-buffers."""
        names = ("color", "color2", "albedo", "normal", "mvec", "linearZ", "pnFwidth", "opacity")
        pf = self.prev_frame
        return (self.t % 2, tuple(frame[k].data_ptr() for k in names),
                None if pf is None else (pf["linearZ"].data_ptr(), pf["normal"].data_ptr()))

    def _graph_for(self, frame) -> "StepGraph":
        """The captured step for this frame's buffers and history parity (captured on first use)."""
        key = self.graph_key(frame)
        g = self._graphs.get(key)
        if g is None:
            if self._cap_stream is None:
                self._cap_stream = torch.cuda.Stream(self.device)
            g = StepGraph([self.make_slot(frame)], self.lr, self._cap_stream.cuda_stream)
            self._graphs[key] = g
        return g

    def close(self):
        for g in self._graphs.values():
            g.destroy()
        self._graphs.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class StepGraph:
    """xdn_denoise_step of fixed slots captured into one HIP graph (xdn_denoise_graph_create)."""

    def __init__(self, slots, lr: float, capture_stream: int):
        arr = (XdnSlot * len(slots))(*slots)
        h = ctypes.c_void_p()
        check(lib.xdn_denoise_graph_create(ctypes.addressof(arr), len(slots), lr, capture_stream, ctypes.byref(h)),
              "graph capture")
        self.handle = h

    def launch(self, stream: int):
        check(lib.xdn_denoise_graph_launch(self.handle, stream), "graph launch")

    def destroy(self):
        if self.handle:
            lib.xdn_denoise_graph_destroy(self.handle)
            self.handle = ctypes.c_void_p()


def denoise_step_multi(denoisers, frames, stream=None):
    """One frame for each (denoiser, frame) pair in ONE batched step (all stages batched)."""
    if not denoisers:
        return
    if len(denoisers) > 16:
        raise ValueError("at most 16 streams per batched step")
    lr = denoisers[0].lr
    if any(d.lr != lr for d in denoisers):
        raise ValueError("streams of one batched step share the learning rate")
    arr = (XdnSlot * len(denoisers))(*[d.make_slot(f) for d, f in zip(denoisers, frames)])
    st = stream if stream is not None else torch.cuda.current_stream(denoisers[0].device).cuda_stream
    check(lib.xdn_denoise_step(ctypes.addressof(arr), len(denoisers), lr, st), "denoise_step")
    for d, f in zip(denoisers, frames):
        d.advance(f)


def rel_l2(pred: torch.Tensor, ref: torch.Tensor) -> float:
    """Mean relative squared error of the frame driver's report (RelL2, scripts/main.py:191-196):
    mean((ref - pred)^2 / (mean_c(ref)^2 + 1e-2)). An evaluation metric, not part of the step."""
    m = ref.mean(dim=-1, keepdim=True)
    return float(((ref - pred) ** 2 / (m * m + 1e-2)).mean())


STAGE_REGRESS, STAGE_REPROJECT, STAGE_FINISH = 1, 2, 4


def run_pipelined(den: Denoiser, frames, main=None, aux=None, on_frame=None):
    """Steps `den` through `frames` (eager launches) with frame t+1's cross-regression on a second
    HIP stream, overlapping frame t's screening, network, filter and backward.

    Frame t+1's cross-regression writes the ols buffers frame t reads as history, so it starts
    once frame t's reprojection is done; frame t's finishing stage waits for its own
    cross-regression. Results are identical to ``den.step`` frame by frame. ``on_frame(i, stream)``
    is called after frame i's last stage is enqueued (e.g. to record an event)."""
    frames = list(frames)
    if not frames:
        return None
    main = main if main is not None else torch.cuda.current_stream(den.device)
    aux = aux if aux is not None else torch.cuda.Stream(den.device)

    def stages(slot, mask, st):
        check(lib.xdn_denoise_stages(ctypes.addressof(slot), 1, den.lr, mask, st.cuda_stream), "denoise stages")

    stages(den.make_slot(frames[0]), STAGE_REGRESS, main)
    regressed = None  # event after the cross-regression of the next frame (on aux)
    for i, f in enumerate(frames):
        slot = den.make_slot(f)
        stages(slot, STAGE_REPROJECT, main)
        if regressed is not None:
            main.wait_event(regressed)
        if i + 1 < len(frames):
            rp = torch.cuda.Event()
            rp.record(main)
            aux.wait_event(rp)
            stages(den.make_slot(frames[i + 1], t=den.t + 1), STAGE_REGRESS, aux)
            regressed = torch.cuda.Event()
            regressed.record(aux)
        stages(slot, STAGE_FINISH, main)
        den.advance(f)
        if on_frame is not None:
            on_frame(i, main)
    return den.outputs()
