"""The custom ops of cross-denoiser behind their reference names (scripts/ops.py:18-40).

Same names, argument names, attributes and NHWC float32 layouts as the TensorFlow ops, on torch
CUDA (HIP) tensors:

    cross_regression(imga, imgb, albedo, normal, win_size, sparsity, recon_scale=1)
        -> (outa, outb, out_vara, out_varb)                          ops/regression.cc:25-57
    spatial_filter(imga, imgb, albedo, normal, band, win_size)
        -> (out_acca, out_accb, out_wgtsuma, out_wgtsumb)            ops/regression.cc:59-86
    spatial_filter_grad(grad_acca, grad_accb, grad_wgtsuma, grad_wgtsumb, imga, imgb,
                        img_albedo, img_normal, band, win_size) -> out_grad_band  :88-103
    temporal_screening(current1, current2, prev, prev1, prev2, reproj_success, win_curr, win_prev)
        -> (output, output1, output2, success)                       ops/temporal.cu.cc:263-285
    Reproject(mvec, input_list, linear_z, prev_linear_z, normal, prev_normal, pn_fwidth,
              opacity, prev_opacity) -> (success [bool], [outputs])  ops/reproject.cc:34-69

``spatial_filter`` is differentiable (torch.autograd.Function). Its gradient w.r.t. ``band`` is
``spatial_filter_grad`` exactly as the reference registers it (scripts/ops.py:23-32); unlike the
reference, gradients w.r.t. the four images are also available (the reference returns None).

Every op runs the hand-written gfx950 kernels of libxdn.so on the current HIP stream; there is no
CPU path. Argument errors raise ValueError with the reference's InvalidArgument text.
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib
from ._lib import check, lib, ptr, stream_of

__all__ = [
    "cross_regression", "spatial_filter", "spatial_filter_grad", "spatial_filter_grad_inputs",
    "temporal_screening", "erode_temporal_screening", "Reproject", "reproject",
]


def _img(t: torch.Tensor, name: str, channels: int | None = None) -> torch.Tensor:
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch.Tensor")
    if t.device.type != "cuda":
        raise ValueError(f"{name} must be a GPU tensor (got {t.device}); the ops have no CPU kernels")
    if t.dtype != torch.float32:
        raise TypeError(f"{name} must be float32 (got {t.dtype})")
    if t.dim() != 4:
        raise ValueError(f"{name} must be NHWC rank 4 (got shape {tuple(t.shape)})")
    if channels is not None and t.shape[3] != channels:
        raise ValueError(f"{name} must have {channels} channels (got {t.shape[3]})")
    return t.contiguous()


def _same_hw(ref: torch.Tensor, *ts) -> None:
    for name, t in ts:
        if t.shape[:3] != ref.shape[:3]:
            raise ValueError(f"{name} has shape {tuple(t.shape)}, expected [B,H,W] = {tuple(ref.shape[:3])}")


# ---------------------------------------------------------------------------------------------
def _cr_forward(imga, imgb, albedo, normal, win_size, sparsity, rs):
    B, H, W, C = imga.shape
    outa = torch.empty((B, H // rs, W // rs, C), device=imga.device, dtype=torch.float32)
    outb = torch.empty_like(outa)
    vara = torch.empty((B, H, W, 1), device=imga.device, dtype=torch.float32)
    varb = torch.empty_like(vara)
    ws_bytes = lib.xdn_cross_regression_workspace_size(B, H, W, int(sparsity))
    ws = torch.empty(max(ws_bytes, 4) // 4, device=imga.device, dtype=torch.float32)
    check(lib.xdn_cross_regression(ptr(imga), ptr(imgb), ptr(albedo), ptr(normal), ptr(outa), ptr(outb),
                                   ptr(vara), ptr(varb), ptr(ws), ws_bytes, B, H, W, int(win_size),
                                   int(sparsity), rs, stream_of(imga)), "cross_regression")
    return outa, outb, vara, varb


class _CrossRegressionFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, imga, imgb, albedo, normal, win_size, sparsity, rs):
        outs = _cr_forward(imga, imgb, albedo, normal, win_size, sparsity, rs)
        ctx.save_for_backward(imga, imgb, albedo, normal)
        ctx.args = (win_size, sparsity, rs)
        return outs

    @staticmethod
    def backward(ctx, g_outa, g_outb, g_vara, g_varb):
        imga, imgb, albedo, normal = ctx.saved_tensors
        win_size, sparsity, rs = ctx.args
        if rs != 1:
            raise NotImplementedError("cross_regression backward supports recon_scale 1 only")
        g = cross_regression_grad(imga, imgb, albedo, normal, g_outa, g_outb, g_vara, g_varb, win_size=win_size,
                                  sparsity=sparsity)
        return g[0], g[1], g[2], g[3], None, None, None


def cross_regression(imga, imgb, albedo, normal, win_size, sparsity, recon_scale=1, name=None):
    """Cross-regression of the two half-buffers (CrossRegressionOp, ops/regression.cc:125-196).

    Differentiable with respect to the four images through torch.autograd (recon_scale 1)."""
    imga, imgb = _img(imga, "imga", 3), _img(imgb, "imgb", 3)
    albedo, normal = _img(albedo, "albedo", 3), _img(normal, "normal", 3)
    _same_hw(imga, ("imgb", imgb), ("albedo", albedo), ("normal", normal))
    B, H, W, C = imga.shape
    if recon_scale > sparsity:
        raise ValueError("recon_scale must be less than or equal to sparsity")
    if H % recon_scale or W % recon_scale:
        raise ValueError("Image height and width must be divisible by recon_scale")
    return _CrossRegressionFn.apply(imga, imgb, albedo, normal, int(win_size), int(sparsity), int(recon_scale))


def cross_regression_grad(imga, imgb, albedo, normal, grad_outa, grad_outb, grad_vara=None, grad_varb=None, *,
                          win_size, sparsity):
    """(d imga, d imgb, d albedo, d normal) of cross_regression (recon_scale 1), window weights
    recomputed. ``grad_vara``/``grad_varb`` may be None (zero)."""
    imga, imgb = _img(imga, "imga", 3), _img(imgb, "imgb", 3)
    albedo, normal = _img(albedo, "albedo", 3), _img(normal, "normal", 3)
    B, H, W, _ = imga.shape

    def g(t, name, c):
        if t is None:
            return None if c == 1 else torch.zeros((B, H, W, c), device=imga.device, dtype=torch.float32)
        return _img(t, name, c)

    ga, gb = g(grad_outa, "grad_outa", 3), g(grad_outb, "grad_outb", 3)
    gva, gvb = g(grad_vara, "grad_vara", 1), g(grad_varb, "grad_varb", 1)
    _same_hw(imga, ("imgb", imgb), ("albedo", albedo), ("normal", normal), ("grad_outa", ga), ("grad_outb", gb))
    outs = [torch.empty_like(imga) for _ in range(4)]
    ws_bytes = lib.xdn_cross_regression_grad_workspace_size(B, H, W, int(sparsity))
    ws = torch.empty(max(ws_bytes, 4) // 4, device=imga.device, dtype=torch.float32)
    check(lib.xdn_cross_regression_grad(ptr(imga), ptr(imgb), ptr(albedo), ptr(normal), ptr(ga), ptr(gb), ptr(gva),
                                        ptr(gvb), *[ptr(o) for o in outs], ptr(ws), ws_bytes, B, H, W, int(win_size),
                                        int(sparsity), stream_of(imga)), "cross_regression_grad")
    return tuple(outs)


def cross_regression_from_color(color, color2, albedo, normal, win_size, sparsity):
    """prepare_input (scripts/main.py:207-212) fused with cross_regression.

    Returns (y1, y2, outa, outb, out_vara, out_varb) with y = log1p(colour).
    """
    color, color2 = _img(color, "color", 3), _img(color2, "color2", 3)
    albedo, normal = _img(albedo, "albedo", 3), _img(normal, "normal", 3)
    _same_hw(color, ("color2", color2), ("albedo", albedo), ("normal", normal))
    B, H, W, _ = color.shape
    y1, y2 = torch.empty_like(color), torch.empty_like(color)
    outa, outb = torch.empty_like(color), torch.empty_like(color)
    vara = torch.empty((B, H, W, 1), device=color.device, dtype=torch.float32)
    varb = torch.empty_like(vara)
    ws_bytes = lib.xdn_cross_regression_workspace_size(B, H, W, int(sparsity))
    ws = torch.empty(max(ws_bytes, 4) // 4, device=color.device, dtype=torch.float32)
    check(lib.xdn_cross_regression_from_color(ptr(color), ptr(color2), ptr(albedo), ptr(normal), ptr(y1), ptr(y2),
                                              ptr(outa), ptr(outb), ptr(vara), ptr(varb), ptr(ws), ws_bytes, B, H,
                                              W, int(win_size), int(sparsity), stream_of(color)),
          "cross_regression_from_color")
    return y1, y2, outa, outb, vara, varb


# ---------------------------------------------------------------------------------------------
def _sf_forward(imga, imgb, albedo, normal, band, win_size):
    B, H, W, _ = imga.shape
    acca, accb = torch.empty_like(imga), torch.empty_like(imgb)
    wa = torch.empty((B, H, W, 1), device=imga.device, dtype=torch.float32)
    wb = torch.empty_like(wa)
    check(lib.xdn_spatial_filter(ptr(imga), ptr(imgb), ptr(albedo), ptr(normal), ptr(band), ptr(acca), ptr(accb),
                                 ptr(wa), ptr(wb), B, H, W, int(win_size), stream_of(imga)), "spatial_filter")
    return acca, accb, wa, wb


def _sf_check(imga, imgb, albedo, normal, band):
    imga, imgb = _img(imga, "imga", 3), _img(imgb, "imgb", 3)
    albedo, normal = _img(albedo, "albedo", 3), _img(normal, "normal", 3)
    band = _img(band, "band", 5)
    _same_hw(imga, ("imgb", imgb), ("albedo", albedo), ("normal", normal), ("band", band))
    return imga, imgb, albedo, normal, band


class _SpatialFilterFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, imga, imgb, albedo, normal, band, win_size):
        outs = _sf_forward(imga, imgb, albedo, normal, band, win_size)
        ctx.save_for_backward(imga, imgb, albedo, normal, band)
        ctx.win_size = win_size
        return outs

    @staticmethod
    def backward(ctx, g_acca, g_accb, g_wa, g_wb):
        imga, imgb, albedo, normal, band = ctx.saved_tensors
        B, H, W, _ = imga.shape

        def z(g, c):
            return torch.zeros((B, H, W, c), device=imga.device, dtype=torch.float32) if g is None else g.contiguous()

        g_acca, g_accb, g_wa, g_wb = z(g_acca, 3), z(g_accb, 3), z(g_wa, 1), z(g_wb, 1)
        need = ctx.needs_input_grad
        g_band = None
        if need[4]:
            g_band = spatial_filter_grad(g_acca, g_accb, g_wa, g_wb, imga, imgb, albedo, normal, band,
                                         win_size=ctx.win_size)
        gi = [None] * 4
        if any(need[:4]):
            gi = list(spatial_filter_grad_inputs(g_acca, g_accb, g_wa, g_wb, imga, imgb, albedo, normal, band,
                                                 win_size=ctx.win_size, which=need[:4]))
        return gi[0], gi[1], gi[2], gi[3], g_band, None


def spatial_filter(imga, imgb, albedo, normal, band, win_size, name=None):
    """Bandwidth-driven bilateral filter (SpatialFilterOp, ops/regression.cc:235-282)."""
    imga, imgb, albedo, normal, band = _sf_check(imga, imgb, albedo, normal, band)
    return _SpatialFilterFn.apply(imga, imgb, albedo, normal, band, int(win_size))


def spatial_filter_grad(grad_acca, grad_accb, grad_wgtsuma, grad_wgtsumb, imga, imgb, img_albedo, img_normal, band,
                        win_size, name=None):
    """d L / d band (SpatialFilterGradOp, ops/regression.cc:283-323), weights recomputed."""
    imga, imgb, img_albedo, img_normal, band = _sf_check(imga, imgb, img_albedo, img_normal, band)
    ga, gb = _img(grad_acca, "grad_acca", 3), _img(grad_accb, "grad_accb", 3)
    gwa, gwb = _img(grad_wgtsuma, "grad_wgtsuma", 1), _img(grad_wgtsumb, "grad_wgtsumb", 1)
    _same_hw(imga, ("grad_acca", ga), ("grad_accb", gb), ("grad_wgtsuma", gwa), ("grad_wgtsumb", gwb))
    B, H, W, _ = imga.shape
    out = torch.empty_like(band)
    check(lib.xdn_spatial_filter_grad(ptr(ga), ptr(gb), ptr(gwa), ptr(gwb), ptr(imga), ptr(imgb), ptr(img_albedo),
                                      ptr(img_normal), ptr(band), ptr(out), B, H, W, int(win_size),
                                      stream_of(imga)), "spatial_filter_grad")
    return out


def spatial_filter_grad_inputs(grad_acca, grad_accb, grad_wgtsuma, grad_wgtsumb, imga, imgb, img_albedo, img_normal,
                               band, win_size, which=(True, True, True, True)):
    """(d imga, d imgb, d albedo, d normal); entries not requested in ``which`` are None."""
    imga, imgb, img_albedo, img_normal, band = _sf_check(imga, imgb, img_albedo, img_normal, band)
    ga, gb = _img(grad_acca, "grad_acca", 3), _img(grad_accb, "grad_accb", 3)
    gwa, gwb = _img(grad_wgtsuma, "grad_wgtsuma", 1), _img(grad_wgtsumb, "grad_wgtsumb", 1)
    B, H, W, _ = imga.shape
    outs = [torch.empty_like(imga) if w else None for w in which]
    check(lib.xdn_spatial_filter_grad_inputs(ptr(ga), ptr(gb), ptr(gwa), ptr(gwb), ptr(imga), ptr(imgb),
                                             ptr(img_albedo), ptr(img_normal), ptr(band), *[ptr(o) for o in outs],
                                             B, H, W, int(win_size), stream_of(imga)), "spatial_filter_grad_inputs")
    return tuple(outs)


# ---------------------------------------------------------------------------------------------
def temporal_screening(current1, current2, prev, prev1, prev2, reproj_success, win_curr, win_prev, name=None):
    """Temporal screening (TemporalScreeningOp, ops/temporal.cu.cc:204-261)."""
    c1, c2 = _img(current1, "current1", 3), _img(current2, "current2", 3)
    p, p1, p2 = _img(prev, "prev", 3), _img(prev1, "prev1", 3), _img(prev2, "prev2", 3)
    rs = _img(reproj_success, "reproj_success", 1)
    _same_hw(c1, ("current2", c2), ("prev", p), ("prev1", p1), ("prev2", p2), ("reproj_success", rs))
    B, H, W, _ = c1.shape
    o, o1, o2 = torch.empty_like(c1), torch.empty_like(c1), torch.empty_like(c1)
    s = torch.empty((B, H, W, 1), device=c1.device, dtype=torch.float32)
    check(lib.xdn_temporal_screening(ptr(c1), ptr(c2), ptr(p), ptr(p1), ptr(p2), ptr(rs), ptr(o), ptr(o1), ptr(o2),
                                     ptr(s), B, H, W, int(win_curr), int(win_prev), stream_of(c1)),
          "temporal_screening")
    return o, o1, o2, s


def erode_temporal_screening(current1, current2, prev, prev1, prev2, success_mask, win_curr, win_prev):
    """3x3 erosion of the boolean reprojection mask (scripts/main.py:270-276) fused with screening.

    Returns (output, output1, output2, success, eroded).
    """
    c1, c2 = _img(current1, "current1", 3), _img(current2, "current2", 3)
    p, p1, p2 = _img(prev, "prev", 3), _img(prev1, "prev1", 3), _img(prev2, "prev2", 3)
    m = success_mask.contiguous()
    if m.dtype not in (torch.bool, torch.uint8) or m.shape[:3] != c1.shape[:3]:
        raise ValueError("success_mask must be a [B,H,W,1] bool/uint8 tensor")
    B, H, W, _ = c1.shape
    o, o1, o2 = torch.empty_like(c1), torch.empty_like(c1), torch.empty_like(c1)
    s = torch.empty((B, H, W, 1), device=c1.device, dtype=torch.float32)
    e = torch.empty_like(s)
    check(lib.xdn_erode_temporal_screening(ptr(c1), ptr(c2), ptr(p), ptr(p1), ptr(p2), ptr(m), ptr(e), ptr(o),
                                           ptr(o1), ptr(o2), ptr(s), B, H, W, int(win_curr), int(win_prev),
                                           stream_of(c1)), "erode_temporal_screening")
    return o, o1, o2, s, e


# ---------------------------------------------------------------------------------------------
def Reproject(mvec, input_list, linear_z, prev_linear_z, normal, prev_normal, pn_fwidth, opacity, prev_opacity,
              name=None):
    """Motion-vector reprojection (ReprojectOp, ops/reproject.cc:87-178).

    Returns (success [B,H,W,1] bool, [outputs shaped like input_list]).
    """
    mvec = _img(mvec, "mvec", 3)
    lz, plz = _img(linear_z, "linear_z", 3), _img(prev_linear_z, "prev_linear_z", 3)
    n, pn = _img(normal, "normal", 3), _img(prev_normal, "prev_normal", 3)
    fw = _img(pn_fwidth, "pn_fwidth", 3)
    op, pop = _img(opacity, "opacity", 1), _img(prev_opacity, "prev_opacity", 1)
    ins = [_img(t, f"input_list[{i}]") for i, t in enumerate(input_list)]
    _same_hw(mvec, ("linear_z", lz), ("prev_linear_z", plz), ("normal", n), ("prev_normal", pn),
             ("pn_fwidth", fw), ("opacity", op), ("prev_opacity", pop), *[(f"input_list[{i}]", t) for i, t in enumerate(ins)])
    B, H, W, _ = mvec.shape
    success = torch.empty((B, H, W, 1), device=mvec.device, dtype=torch.bool)
    outs = [torch.empty_like(t) for t in ins]
    L = len(ins)
    in_arr = (ctypes.c_void_p * max(L, 1))(*[t.data_ptr() for t in ins])
    out_arr = (ctypes.c_void_p * max(L, 1))(*[t.data_ptr() for t in outs])
    dims = (ctypes.c_int * max(L, 1))(*[int(t.shape[3]) for t in ins])
    check(lib.xdn_reproject(ptr(mvec), L, in_arr, dims, ptr(lz), ptr(plz), ptr(n), ptr(pn), ptr(fw), ptr(op), ptr(pop),
                            ptr(success), out_arr, B, H, W, stream_of(mvec)), "Reproject")
    return success, outs


reproject = Reproject
