"""FP64 oracle of cross-denoiser's per-frame filtering path.

TEST INFRASTRUCTURE ONLY. Imported by tests/, by the smoke check of __graft_entry__ and by the
cpu_baseline leg of bench.py, and only to check results against or to time as the CPU baseline.
The product (the ``xdn`` package and libxdn.so) never imports it.

The stencil kernels are a C restatement (``xdn_oracle.c``, built into ``liboracle.so`` by the
Makefile next to this file); the pointwise parts of ``scripts/main.py`` are restated here in numpy
float64. All functions take NHWC float32 numpy arrays ``[B,H,W,C]`` (the same layout as the
reference ops) and return float64 arrays unless the output is a mask or a selection, which is
returned bit-exact in its reference dtype.

Each function is pinned by tests/test_oracle.py (the reference's golden vectors where they exist,
closed forms, invariants and brute force on small inputs).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_i = ctypes.c_int


def build(force: bool = False) -> str:
    """Compile liboracle.so with the Makefile beside this file."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.check_call(["make", "-s", "-C", _HERE] + (["-B"] if force else []))
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_variance.argtypes = [_f32p, _f64p, _i, _i]
        L.orc_regression.argtypes = [_f32p, _f32p, _f64p, _f32p, _f32p, _f64p, _i, _i, _i, _i]
        L.orc_recon.argtypes = [_f64p, _f32p, _f64p, _f32p, _f32p, _f64p, _i, _i, _i, _i, _i]
        L.orc_cross_regression.argtypes = [_f32p] * 4 + [_f64p] * 6 + [_i] * 5
        L.orc_cross_regression_grad.argtypes = [_f32p] * 4 + [_f32p, _f32p, ctypes.c_void_p, ctypes.c_void_p] + \
            [_f64p] * 4 + [_i] * 4
        L.orc_spatial_filter.argtypes = [_f32p] * 5 + [_f64p] * 4 + [_i] * 3
        L.orc_spatial_filter_grad.argtypes = [_f32p] * 9 + [_f64p] + [_i] * 3
        L.orc_spatial_filter_grad_inputs.argtypes = [_f32p] * 9 + [_f64p] * 4 + [_i] * 3
        L.orc_reproject.argtypes = [
            _f32p, _i, ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_int),
            _f32p, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p, _u8p,
            ctypes.POINTER(ctypes.c_void_p), _i, _i]
        L.orc_erosion.argtypes = [_u8p, ctypes.c_void_p, _i, _i]
        L.orc_screening.argtypes = [_f32p] * 6 + [ctypes.c_void_p] * 4 + [_i] * 4
        L.orc_set_threads.argtypes = [_i]
        L.orc_get_threads.restype = _i
        _lib = L
    return _lib


def set_threads(n: int) -> None:
    """Host threads of the C oracle's OpenMP loops."""
    lib().orc_set_threads(int(n))


def threads() -> int:
    return int(lib().orc_get_threads())


def _f32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32))


def _check4(a, c=None):
    if a.ndim != 4 or (c is not None and a.shape[3] != c):
        raise ValueError(f"expected NHWC with {c} channels, got {a.shape}")


# --------------------------------------------------------------------------------------------
# a1: prepare_input / output transform, scripts/main.py:207-212, 248
# --------------------------------------------------------------------------------------------
def prepare_input(color):
    return np.log1p(np.asarray(color, dtype=np.float64))


def output_transform(out_log):
    return np.maximum(0.0, np.expm1(np.asarray(out_log, dtype=np.float64)))


# --------------------------------------------------------------------------------------------
# a2-a5: cross regression (ops/regression.cu.cc:244-500)
# --------------------------------------------------------------------------------------------
def variance(y):
    y = _f32(y); _check4(y, 3)
    B, H, W, _ = y.shape
    out = np.zeros((B, H, W, 1), np.float64)
    for b in range(B):
        o = np.zeros((H, W), np.float64)
        lib().orc_variance(y[b].copy(), o, H, W)
        out[b, ..., 0] = o
    return out


def regression(target, feat, var, albedo, normal, win_size=17, sparsity=4):
    """beta [B, H/s, W/s, 6, 3] for each center (RegressionKernel regression.cu.cc:367-475)."""
    target, feat, albedo, normal = map(_f32, (target, feat, albedo, normal))
    var = np.ascontiguousarray(np.asarray(var, np.float64))
    B, H, W, _ = feat.shape
    Hs, Ws = H // sparsity, W // sparsity
    out = np.zeros((B, Hs, Ws, 6, 3), np.float64)
    for b in range(B):
        o = np.zeros((Hs, Ws, 6, 3), np.float64)
        lib().orc_regression(target[b].copy(), feat[b].copy(), var[b, ..., 0].copy(), albedo[b].copy(),
                             normal[b].copy(), o, H, W, win_size, sparsity)
        out[b] = o
    return out


def recon(beta, feat, var, albedo, normal, win_size=17, sparsity=4, recon_scale=1):
    feat, albedo, normal = map(_f32, (feat, albedo, normal))
    beta = np.ascontiguousarray(np.asarray(beta, np.float64))
    var = np.ascontiguousarray(np.asarray(var, np.float64))
    B, H, W, _ = feat.shape
    out = np.zeros((B, H // recon_scale, W // recon_scale, 3), np.float64)
    for b in range(B):
        o = np.zeros(out.shape[1:], np.float64)
        lib().orc_recon(beta[b].copy(), feat[b].copy(), var[b, ..., 0].copy(), albedo[b].copy(),
                        normal[b].copy(), o, H, W, win_size, sparsity, recon_scale)
        out[b] = o
    return out


def cross_regression(imga, imgb, albedo, normal, win_size=17, sparsity=4, recon_scale=1):
    """(outa, outb, out_vara, out_varb), CrossRegressionFunc regression.cu.cc:477-500."""
    imga, imgb, albedo, normal = map(_f32, (imga, imgb, albedo, normal))
    _check4(imga, 3)
    B, H, W, _ = imga.shape
    rs = recon_scale
    outa = np.zeros((B, H // rs, W // rs, 3)); outb = np.zeros_like(outa)
    vara = np.zeros((B, H, W, 1)); varb = np.zeros_like(vara)
    nb = (H // sparsity) * (W // sparsity) * 18
    for b in range(B):
        oa = np.zeros((H // rs, W // rs, 3)); ob = np.zeros_like(oa)
        va = np.zeros((H, W)); vb = np.zeros_like(va)
        ba = np.zeros(nb); bb = np.zeros(nb)
        lib().orc_cross_regression(imga[b].copy(), imgb[b].copy(), albedo[b].copy(), normal[b].copy(),
                                   oa, ob, va, vb, ba, bb, H, W, win_size, sparsity, rs)
        outa[b], outb[b], vara[b, ..., 0], varb[b, ..., 0] = oa, ob, va, vb
    return outa, outb, vara, varb


def cross_regression_grad(imga, imgb, albedo, normal, grad_outa, grad_outb, grad_vara=None, grad_varb=None,
                          win_size=17, sparsity=4):
    """FP64 adjoint of cross_regression (recon_scale 1): (grad_imga, grad_imgb, grad_albedo, grad_normal)."""
    imga, imgb, albedo, normal, grad_outa, grad_outb = map(_f32, (imga, imgb, albedo, normal, grad_outa, grad_outb))
    B, H, W, _ = imga.shape
    outs = [np.zeros((B, H, W, 3)) for _ in range(4)]
    for b in range(B):
        gv = [None if g is None else np.ascontiguousarray(_f32(g)[b].reshape(H, W)) for g in (grad_vara, grad_varb)]
        o = [np.zeros((H, W, 3)) for _ in range(4)]
        lib().orc_cross_regression_grad(imga[b].copy(), imgb[b].copy(), albedo[b].copy(), normal[b].copy(),
                                        grad_outa[b].copy(), grad_outb[b].copy(),
                                        None if gv[0] is None else gv[0].ctypes.data,
                                        None if gv[1] is None else gv[1].ctypes.data, *o, H, W, win_size, sparsity)
        for k in range(4):
            outs[k][b] = o[k]
    return tuple(outs)


# --------------------------------------------------------------------------------------------
# a6/a7: spatial filter and its gradients (regression.cu.cc:33-195)
# --------------------------------------------------------------------------------------------
def spatial_filter(imga, imgb, albedo, normal, band, win_size=11):
    imga, imgb, albedo, normal, band = map(_f32, (imga, imgb, albedo, normal, band))
    B, H, W, _ = imga.shape
    accA = np.zeros((B, H, W, 3)); accB = np.zeros_like(accA)
    wA = np.zeros((B, H, W, 1)); wB = np.zeros_like(wA)
    for b in range(B):
        a1 = np.zeros((H, W, 3)); a2 = np.zeros_like(a1); w1 = np.zeros((H, W)); w2 = np.zeros_like(w1)
        lib().orc_spatial_filter(imga[b].copy(), imgb[b].copy(), albedo[b].copy(), normal[b].copy(),
                                 band[b].copy(), a1, a2, w1, w2, H, W, win_size)
        accA[b], accB[b], wA[b, ..., 0], wB[b, ..., 0] = a1, a2, w1, w2
    return accA, accB, wA, wB


def spatial_filter_grad(grad_acca, grad_accb, grad_wgtsuma, grad_wgtsumb, imga, imgb, img_albedo,
                        img_normal, band, win_size=11):
    args = list(map(_f32, (grad_acca, grad_accb, grad_wgtsuma, grad_wgtsumb, imga, imgb, img_albedo,
                           img_normal, band)))
    B, H, W, _ = args[4].shape
    out = np.zeros((B, H, W, 5))
    for b in range(B):
        o = np.zeros((H, W, 5))
        lib().orc_spatial_filter_grad(*[a[b].copy() for a in args], o, H, W, win_size)
        out[b] = o
    return out


def spatial_filter_grad_inputs(grad_acca, grad_accb, grad_wgtsuma, grad_wgtsumb, imga, imgb,
                               img_albedo, img_normal, band, win_size=11):
    """(d imga, d imgb, d albedo, d normal): analytic adjoint of the filter w.r.t. its images."""
    args = list(map(_f32, (grad_acca, grad_accb, grad_wgtsuma, grad_wgtsumb, imga, imgb, img_albedo,
                           img_normal, band)))
    B, H, W, _ = args[4].shape
    outs = [np.zeros((B, H, W, 3)) for _ in range(4)]
    for b in range(B):
        o = [np.zeros((H, W, 3)) for _ in range(4)]
        lib().orc_spatial_filter_grad_inputs(*[a[b].copy() for a in args], *o, H, W, win_size)
        for k in range(4):
            outs[k][b] = o[k]
    return tuple(outs)


# --------------------------------------------------------------------------------------------
# a8: reprojection (ops/reproject.cu.cc:32-238)
# --------------------------------------------------------------------------------------------
def reproject(mvec, input_list, linear_z, prev_linear_z, normal, prev_normal, pn_fwidth, opacity,
              prev_opacity):
    """(success bool [B,H,W,1], [outputs float64]) with exact-FP32 masks."""
    mvec, linear_z, prev_linear_z, normal, prev_normal, pn_fwidth, opacity, prev_opacity = map(
        _f32, (mvec, linear_z, prev_linear_z, normal, prev_normal, pn_fwidth, opacity, prev_opacity))
    ins = [_f32(a) for a in input_list]
    B, H, W, _ = mvec.shape
    L = len(ins)
    dims = [int(a.shape[3]) for a in ins]
    success = np.zeros((B, H, W, 1), np.uint8)
    outs = [np.zeros((B, H, W, d)) for d in dims]
    for b in range(B):
        fin = [a[b].copy() for a in ins]
        fout = [np.zeros((H, W, d)) for d in dims]
        in_ptrs = (ctypes.c_void_p * max(L, 1))(*[a.ctypes.data for a in fin])
        out_ptrs = (ctypes.c_void_p * max(L, 1))(*[a.ctypes.data for a in fout])
        cdims = (ctypes.c_int * max(L, 1))(*dims)
        s = np.zeros((H, W), np.uint8)
        lib().orc_reproject(mvec[b].copy(), L, in_ptrs, cdims, linear_z[b].copy(), prev_linear_z[b].copy(),
                            normal[b].copy(), prev_normal[b].copy(), pn_fwidth[b].copy(),
                            opacity[b].copy(), prev_opacity[b].copy(), s, out_ptrs, H, W)
        success[b, ..., 0] = s
        for k in range(L):
            outs[k][b] = fout[k]
    return success.astype(bool), outs


# --------------------------------------------------------------------------------------------
# a9/a10: erosion and temporal screening (scripts/main.py:270-276, ops/temporal.cu.cc:57-198)
# --------------------------------------------------------------------------------------------
def erosion(success):
    s = np.ascontiguousarray(np.asarray(success).astype(np.uint8))
    B, H, W, _ = s.shape
    out = np.zeros((B, H, W, 1), np.float32)
    for b in range(B):
        o = np.zeros((H, W), np.float32)
        lib().orc_erosion(s[b, ..., 0].copy(), o.ctypes.data, H, W)
        out[b, ..., 0] = o
    return out


def temporal_screening(current1, current2, prev, prev1, prev2, reproj_success, win_curr=7, win_prev=1):
    """(output, output1, output2, success) float32, bit-exact selections."""
    args = list(map(_f32, (current1, current2, prev, prev1, prev2, reproj_success)))
    B, H, W, _ = args[0].shape
    outs = [np.zeros((B, H, W, 3), np.float32) for _ in range(3)] + [np.zeros((B, H, W, 1), np.float32)]
    for b in range(B):
        o = [np.zeros((H, W, 3), np.float32) for _ in range(3)] + [np.zeros((H, W), np.float32)]
        lib().orc_screening(*[a[b].copy() for a in args], *[x.ctypes.data for x in o], H, W, win_curr, win_prev)
        for k in range(3):
            outs[k][b] = o[k]
        outs[3][b, ..., 0] = o[3]
    return tuple(outs)


# --------------------------------------------------------------------------------------------
# a11: aggregation + loss (scripts/main.py:214-248) and its analytic backward
# --------------------------------------------------------------------------------------------
def aggregation_loss(sum1, sumw1, sum2, sumw2, alpha, success, success_prev, prev_out, opacity,
                     label1, label2, prev_label1, prev_label2):
    """(out_log, out, loss, avg_loss) in float64 (main.py:214-248)."""
    f = lambda a: np.asarray(a, np.float64)
    sum1, sumw1, sum2, sumw2, alpha, success, success_prev, prev_out, opacity = map(
        f, (sum1, sumw1, sum2, sumw2, alpha, success, success_prev, prev_out, opacity))
    label1, label2, prev_label1, prev_label2 = map(f, (label1, label2, prev_label1, prev_label2))
    out1 = sum1 / (sumw1 + 1e-4)
    out2 = sum2 / (sumw2 + 1e-4)
    out = (sum1 + sum2) / (sumw1 + sumw2 + 1e-4)
    op = np.where(opacity > 0, 1.0, 0.0)
    s = success * success_prev
    a = s * alpha + (1 - s)
    out = (a * out + (1 - a) * prev_out) * op
    o1 = a * out1 + (1 - a) * prev_out
    o2 = a * out2 + (1 - a) * prev_out
    loss1 = np.square(o1 - label1) / (np.square(label1) + 1e-2)
    loss2 = np.square(o2 - label2) / (np.square(label2) + 1e-2)
    t1 = np.square(o1 - prev_label1) / (np.square(prev_label1) + 1e-2)
    t2 = np.square(o2 - prev_label2) / (np.square(prev_label2) + 1e-2)
    loss = ((loss1 + loss2) * 0.5 + (t1 + t2) * 0.5) * 0.5 * op
    return out, output_transform(out), loss, float(loss.mean())


def aggregation_loss_grad(sum1, sumw1, sum2, sumw2, alpha, success, success_prev, prev_out, opacity,
                          label1, label2, prev_label1, prev_label2):
    """d avg_loss / d (sum1, sumw1, sum2, sumw2, alpha), analytic, float64."""
    f = lambda a: np.asarray(a, np.float64)
    sum1, sumw1, sum2, sumw2, alpha, success, success_prev, prev_out, opacity = map(
        f, (sum1, sumw1, sum2, sumw2, alpha, success, success_prev, prev_out, opacity))
    label1, label2, prev_label1, prev_label2 = map(f, (label1, label2, prev_label1, prev_label2))
    n = sum1.size  # mean over B*H*W*3 elements
    op = np.where(opacity > 0, 1.0, 0.0)
    s = success * success_prev
    a = s * alpha + (1 - s)
    d1 = sumw1 + 1e-4
    d2 = sumw2 + 1e-4
    out1, out2 = sum1 / d1, sum2 / d2
    o1 = a * out1 + (1 - a) * prev_out
    o2 = a * out2 + (1 - a) * prev_out
    c = 0.25 * op / n
    g1 = c * (2 * (o1 - label1) / (label1 ** 2 + 1e-2) + 2 * (o1 - prev_label1) / (prev_label1 ** 2 + 1e-2))
    g2 = c * (2 * (o2 - label2) / (label2 ** 2 + 1e-2) + 2 * (o2 - prev_label2) / (prev_label2 ** 2 + 1e-2))
    gout1, gout2 = a * g1, a * g2
    ga = np.sum(g1 * (out1 - prev_out) + g2 * (out2 - prev_out), axis=-1, keepdims=True)
    gsum1 = gout1 / d1
    gsum2 = gout2 / d2
    gsumw1 = -np.sum(gout1 * sum1, axis=-1, keepdims=True) / d1 ** 2
    gsumw2 = -np.sum(gout2 * sum2, axis=-1, keepdims=True) / d2 ** 2
    galpha = s * ga
    return gsum1, gsumw1, gsum2, gsumw2, galpha


# --------------------------------------------------------------------------------------------
# a12 head: bandwidth/alpha transform of the network output (scripts/main.py:110-113)
# --------------------------------------------------------------------------------------------
def band_alpha(param):
    p = np.asarray(param, np.float64)
    band = 1.0 / (np.square(p[..., :5]) + 1e-4)
    alpha = 1.0 / (1.0 + np.exp(-p[..., 5:6]))
    return band, alpha


def band_alpha_grad(param, gband, galpha):
    p = np.asarray(param, np.float64)
    x = p[..., :5]
    band = 1.0 / (np.square(x) + 1e-4)
    alpha = 1.0 / (1.0 + np.exp(-p[..., 5:6]))
    gx = np.asarray(gband, np.float64) * (-2.0 * x * band * band)
    ga = np.asarray(galpha, np.float64) * alpha * (1 - alpha)
    return np.concatenate([gx, ga], axis=-1)


def filter_train(ols1, ols2, albedo, normal, param, success, success_prev, prev_out, opacity, prev_ols1,
                 prev_ols2, win_size=11):
    """The spatiotemporal part of train_step (scripts/main.py:283-302) for a fixed network output
    `param` [B,H,W,6]: band/alpha head (main.py:110-113), spatial_filter(ols1, ols2, ...) and
    aggregation_loss with the labels crossed as main.py:283-286 sets them (label1 = ols2,
    label2 = ols1, prev_label1 = prev_ols2, prev_label2 = prev_ols1), then the analytic gradient of
    each image's mean loss back to `param`.

    Returns (out_log, out, loss, avg_loss [B], dparam) in float64."""
    band, alpha = band_alpha(param)
    band32 = band.astype(np.float32)
    acc1, acc2, wsum1, wsum2 = spatial_filter(ols1, ols2, albedo, normal, band32, win_size)
    agg = (acc1, wsum1, acc2, wsum2, alpha, success, success_prev, prev_out, opacity,
           ols2, ols1, prev_ols2, prev_ols1)
    out_log, out, loss, _ = aggregation_loss(*agg)
    B = np.asarray(ols1).shape[0]
    avg = np.array([aggregation_loss(*[np.asarray(a)[b:b + 1] for a in agg])[3] for b in range(B)])
    grads = [aggregation_loss_grad(*[np.asarray(a)[b:b + 1] for a in agg]) for b in range(B)]
    gsum1, gsumw1, gsum2, gsumw2, galpha = [np.concatenate([g[k] for g in grads], 0) for k in range(5)]
    gband = spatial_filter_grad(gsum1, gsum2, gsumw1, gsumw2, ols1, ols2, albedo, normal, band32, win_size)
    dparam = band_alpha_grad(param, gband, galpha)
    return out_log, out, loss, avg, dparam


def frame_report(out, ref, emissive=None, env_light=None, ref_emissive=None, ref_env_light=None):
    """Displayed frame and its RelL2 error of the driver loop (scripts/main.py:344-346), with RelL2
    as defined at main.py:191-196: per element (y_true - y_pred)^2 / (mean_c(y_true)^2 + 1e-2),
    averaged over the image. Returns (out + emissive + env_light [B,H,W,3], err [B]) in float64;
    absent terms are zero."""
    d = lambda a: 0.0 if a is None else np.asarray(a, np.float64)
    o = np.asarray(out, np.float64) + d(emissive) + d(env_light)
    r = np.asarray(ref, np.float64) + d(ref_emissive) + d(ref_env_light)
    m = r.mean(axis=-1, keepdims=True)
    e = np.square(r - o) / (np.square(m) + 1e-2)
    return o, e.reshape(e.shape[0], -1).mean(axis=1)
