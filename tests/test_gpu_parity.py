"""GPU parity of every op against the FP64 oracle, through the C ABI (libxdn.so).

Tolerances (stated per output, relative L2 = ||gpu - oracle|| / ||oracle||):
  filtered radiance and sums (spatial filter, fused step)       <= 1e-4 (the stated bar); observed ~1e-6
  cross-regression outputs (FP32 normal equations + Cholesky)   <= 1e-4
  gradients (FP32 accumulation of 121-tap sums)                 <= 1e-4
  reprojection success mask and screening selections            bit-exact
  reprojected values (FP32 bilinear vs FP64)                    <= 1e-6
"""
import numpy as np
import pytest
import torch

from helpers import log_images, rand_img, rel_l2, to_dev, to_np, unit_normals
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-4


@pytest.fixture(scope="module")
def ops(dev):
    from xdn import ops as xops
    return xops


def _report(name, err):
    print(f"  {name}: rel L2 vs FP64 oracle = {err:.3e}", flush=True)


# ------------------------------------------------------------------------------------------------
# 24x300 and 20x264 run the paired row tails of the regression launch (75 and 66 centers per row:
# a last block of at most 32 centers); 20x264 with an odd count of center rows (5) in a batch of 2.
@pytest.mark.parametrize("B,H,W", [(1, 64, 64), (1, 48, 40), (2, 36, 52), (1, 256, 256), (1, 24, 300),
                                   (2, 20, 264)])
def test_cross_regression(ops, dev, B, H, W):
    ya, yb = log_images(H, W, seed=H * W, B=B)
    alb, nrm = rand_img(H, W, seed=3, B=B), unit_normals(H, W, seed=4, B=B)
    ref = O.cross_regression(ya, yb, alb, nrm, 17, 4, 1)
    got = ops.cross_regression(to_dev(ya, dev), to_dev(yb, dev), to_dev(alb, dev), to_dev(nrm, dev), win_size=17,
                               sparsity=4)
    for name, g, r in zip(("outa", "outb", "out_vara", "out_varb"), got, ref):
        err = rel_l2(to_np(g), r)
        _report(f"cross_regression {name} {B}x{H}x{W}", err)
        assert err <= TOL, name
    assert got[0].shape == (B, H, W, 3) and got[2].shape == (B, H, W, 1)


@pytest.mark.parametrize("win,s,rs", [(13, 2, 1), (17, 4, 2), (9, 4, 4)])
def test_cross_regression_generic_attrs(ops, dev, win, s, rs):
    H, W = 32, 40
    ya, yb = log_images(H, W, seed=7)
    alb, nrm = rand_img(H, W, seed=5), unit_normals(H, W, seed=6)
    ref = O.cross_regression(ya, yb, alb, nrm, win, s, rs)
    got = ops.cross_regression(to_dev(ya, dev), to_dev(yb, dev), to_dev(alb, dev), to_dev(nrm, dev), win_size=win,
                               sparsity=s, recon_scale=rs)
    for name, g, r in zip(("outa", "outb", "out_vara", "out_varb"), got, ref):
        err = rel_l2(to_np(g), r)
        _report(f"cross_regression {name} win={win} s={s} rs={rs}", err)
        assert err <= TOL, name


def cr_grad_case(H, W, seed, B=1):
    ya, yb = log_images(H, W, seed=seed, B=B)
    return [ya, yb, rand_img(H, W, seed=seed + 1, B=B), unit_normals(H, W, seed=seed + 2, B=B)]


# Gradients pass through the per-center 6x6 solves (u = M^-1 g_beta) in FP32; stated tolerance.
CR_GRAD_TOL = 1e-4


@pytest.mark.parametrize("B,H,W,win,s", [(1, 48, 64, 17, 4), (2, 36, 20, 17, 4), (1, 30, 27, 9, 3), (1, 24, 32, 13, 2)])
def test_cross_regression_grad(ops, dev, B, H, W, win, s):
    args = cr_grad_case(H, W, seed=H + W, B=B)
    cots = _cots(B, H, W, 12)
    ref = O.cross_regression_grad(*args, *cots, win_size=win, sparsity=s)
    got = ops.cross_regression_grad(*[to_dev(a, dev) for a in args], *[to_dev(c, dev) for c in cots],
                                    win_size=win, sparsity=s)
    for name, g, r in zip(("imga", "imgb", "albedo", "normal"), got, ref):
        err = rel_l2(to_np(g), r)
        _report(f"cross_regression_grad {name} {B}x{H}x{W} win={win} s={s}", err)
        assert err <= CR_GRAD_TOL, name


def test_cross_regression_grad_512_sampled(ops, dev):
    """512x512 inputs, gradient of L = sum r . outputs on 1024 random pixels against the FP64
    adjoint (itself pinned by central differences of the FP64 forward)."""
    H = W = 512
    args = cr_grad_case(H, W, seed=512)
    cots = _cots(1, H, W, 13)
    ref = O.cross_regression_grad(*args, *cots)
    got = ops.cross_regression_grad(*[to_dev(a, dev) for a in args], *[to_dev(c, dev) for c in cots],
                                    win_size=17, sparsity=4)
    r = np.random.default_rng(14)
    ys, xs = r.integers(0, H, 1024), r.integers(0, W, 1024)
    for name, g, rr in zip(("imga", "imgb", "albedo", "normal"), got, ref):
        err = rel_l2(to_np(g)[0, ys, xs], rr[0, ys, xs])
        _report(f"cross_regression_grad {name} 512x512 (1024 px)", err)
        assert err <= CR_GRAD_TOL, name


def test_cross_regression_autograd(ops, dev):
    H, W = 32, 40
    args = [to_dev(a, dev).requires_grad_(True) for a in cr_grad_case(H, W, seed=5)]
    outs = ops.cross_regression(*args, win_size=17, sparsity=4)
    cots = [to_dev(c, dev) for c in _cots(1, H, W, 15)]
    L = sum((o * c).sum() for o, c in zip(outs, cots))
    L.backward()
    direct = ops.cross_regression_grad(*[a.detach() for a in args], *cots, win_size=17, sparsity=4)
    for a, g in zip(args, direct):
        assert torch.equal(a.grad, g)


def test_cross_regression_errors(ops, dev):
    x = torch.zeros(1, 32, 32, 3, device=dev)
    with pytest.raises(ValueError, match="recon_scale must be less than or equal to sparsity"):
        ops.cross_regression(x, x, x, x, win_size=17, sparsity=2, recon_scale=4)
    y = torch.zeros(1, 30, 32, 3, device=dev)
    with pytest.raises(ValueError, match="divisible by recon_scale"):
        ops.cross_regression(y, y, y, y, win_size=17, sparsity=4, recon_scale=4)


def test_cross_regression_from_color(ops, dev):
    H, W = 64, 72
    r = np.random.default_rng(0)
    c1 = r.lognormal(0, 1, (1, H, W, 3)).astype(np.float32)
    c2 = r.lognormal(0, 1, (1, H, W, 3)).astype(np.float32)
    alb, nrm = rand_img(H, W, seed=1), unit_normals(H, W, seed=2)
    y1, y2, oa, ob, va, vb = ops.cross_regression_from_color(to_dev(c1, dev), to_dev(c2, dev), to_dev(alb, dev),
                                                             to_dev(nrm, dev), 17, 4)
    assert rel_l2(to_np(y1), O.prepare_input(c1)) < 1e-6
    ref = O.cross_regression(np.log1p(c1.astype(np.float64)).astype(np.float32),
                             np.log1p(c2.astype(np.float64)).astype(np.float32), alb, nrm, 17, 4, 1)
    for name, g, rr in zip(("outa", "outb", "vara", "varb"), (oa, ob, va, vb), ref):
        err = rel_l2(to_np(g), rr)
        _report(f"cross_regression_from_color {name}", err)
        assert err <= TOL


# ------------------------------------------------------------------------------------------------
def sf_case(H, W, seed, B=1, band_lo=0.1, band_hi=10.0):
    ia, ib = log_images(H, W, seed=seed, B=B)
    alb, nrm = rand_img(H, W, seed=seed + 1, B=B), unit_normals(H, W, seed=seed + 2, B=B)
    r = np.random.default_rng(seed + 3)
    band = np.exp(r.uniform(np.log(band_lo), np.log(band_hi), (B, H, W, 5))).astype(np.float32)
    return ia, ib, alb, nrm, band


@pytest.mark.parametrize("B,H,W,win", [(1, 64, 64, 11), (1, 37, 29, 11), (2, 48, 80, 11), (1, 512, 512, 11),
                                       (1, 30, 33, 7), (1, 20, 24, 15)])
def test_spatial_filter(ops, dev, B, H, W, win):
    args = sf_case(H, W, seed=H + W, B=B)
    ref = O.spatial_filter(*args, win_size=win)
    got = ops.spatial_filter(*[to_dev(a, dev) for a in args], win_size=win)
    for name, g, r in zip(("acca", "accb", "wgtsuma", "wgtsumb"), got, ref):
        err = rel_l2(to_np(g), r)
        _report(f"spatial_filter {name} {B}x{H}x{W} win={win}", err)
        assert err <= TOL, name


def _cots(B, H, W, seed):
    r = np.random.default_rng(seed)
    return [r.normal(size=(B, H, W, c)).astype(np.float32) for c in (3, 3, 1, 1)]


@pytest.mark.parametrize("B,H,W,win", [(1, 64, 64, 11), (1, 37, 29, 11), (2, 24, 40, 11), (1, 33, 31, 9)])
def test_spatial_filter_grad(ops, dev, B, H, W, win):
    args = sf_case(H, W, seed=2 * H + W, B=B)
    cots = _cots(B, H, W, 9)
    ref = O.spatial_filter_grad(*cots, *args, win_size=win)
    got = ops.spatial_filter_grad(*[to_dev(c, dev) for c in cots], *[to_dev(a, dev) for a in args], win_size=win)
    err = rel_l2(to_np(got), ref)
    _report(f"spatial_filter_grad {B}x{H}x{W} win={win}", err)
    assert err <= TOL


@pytest.mark.parametrize("H,W,win", [(32, 32, 11), (23, 29, 11), (16, 40, 7)])
def test_spatial_filter_grad_inputs(ops, dev, H, W, win):
    args = sf_case(H, W, seed=H + 3 * W, band_lo=0.1, band_hi=3.0)
    cots = _cots(1, H, W, 10)
    ref = O.spatial_filter_grad_inputs(*cots, *args, win_size=win)
    got = ops.spatial_filter_grad_inputs(*[to_dev(c, dev) for c in cots], *[to_dev(a, dev) for a in args],
                                         win_size=win)
    for name, g, r in zip(("imga", "imgb", "albedo", "normal"), got, ref):
        err = rel_l2(to_np(g), r)
        _report(f"spatial_filter_grad_inputs {name} {H}x{W} win={win}", err)
        assert err <= TOL, name


def test_spatial_filter_autograd(ops, dev):
    H, W = 40, 44
    args = [to_dev(a, dev) for a in sf_case(H, W, seed=77)]
    for a in args:
        a.requires_grad_(True)
    outs = ops.spatial_filter(*args, win_size=11)
    cots = [to_dev(c, dev) for c in _cots(1, H, W, 11)]
    L = sum((o * c).sum() for o, c in zip(outs, cots))
    L.backward()
    gb = ops.spatial_filter_grad(*cots, *[a.detach() for a in args], win_size=11)
    assert torch.equal(args[4].grad, gb)
    ref = O.spatial_filter_grad_inputs(*[to_np(c) for c in cots], *[to_np(a.detach()) for a in args])
    for a, r in zip(args[:4], ref):
        assert rel_l2(to_np(a.grad), r) <= TOL


def _sf_local_forward(ia, ib, alb, nrm, band, ys, xs, half=5):
    """FP64 restatement of the spatial filter at the pixels (ys, xs) only (band per pixel, float64;
    reflect-101 boundary), for central differences in the band: [n, 8] = accA, accB, wsumA, wsumB."""
    H, W = ia.shape[1:3]
    dy, dx = np.meshgrid(np.arange(-half, half + 1), np.arange(-half, half + 1), indexing="ij")
    dy, dx = dy.ravel(), dx.ravel()
    ty, tx = ys[:, None] + dy[None], xs[:, None] + dx[None]
    ty = np.where(ty < 0, -ty, np.where(ty >= H, 2 * H - 2 - ty, ty))
    tx = np.where(tx < 0, -tx, np.where(tx >= W, 2 * W - 2 - tx, tx))
    g = lambda img: img[0].astype(np.float64)
    c = lambda img: g(img)[ys, xs][:, None, :]
    t = lambda img: g(img)[ty, tx]
    d2 = lambda img: ((c(img) - t(img)) ** 2).sum(-1)
    pd = ((dx / half) ** 2 + (dy / half) ** 2)[None]
    common = band[:, 2:3] * d2(alb) + band[:, 3:4] * d2(nrm) + band[:, 4:5] * pd
    wA = np.exp(-band[:, 0:1] * d2(ia) - common)
    wB = np.exp(-band[:, 1:2] * d2(ib) - common)
    accA = (wA[..., None] * t(ia)).sum(1)
    accB = (wB[..., None] * t(ib)).sum(1)
    return np.concatenate([accA, accB, wA.sum(1, keepdims=True), wB.sum(1, keepdims=True)], 1)


def test_spatial_filter_grad_512_sampled(ops, dev):
    """The gradient-check configuration for the filter: 512x512 inputs of the synthetic generator
    (log half-buffers, albedo, normal), bandwidths log-uniform in [0.1, 10] per pixel and channel,
    the autograd backward of spatial_filter (band and image gradients of L = sum r . outputs) on 1024
    random pixels against the FP64 oracle's gradient, and the band gradient also against FP64
    central differences (h = 1e-4) of the filter restated at those pixels. Bar: rel. L2 <= 1e-4."""
    from xdn import synth
    H = W = 512
    f = synth.make_frame(synth.StreamSpec(H, W, seed=2), 0)
    ia = np.log1p(f["color"].numpy().astype(np.float64)).astype(np.float32)
    ib = np.log1p(f["color2"].numpy().astype(np.float64)).astype(np.float32)
    alb, nrm = f["albedo"].numpy(), f["normal"].numpy()
    band = synth.log_uniform_band((1, H, W, 5), 0.1, 10.0, seed=22).numpy()
    cots = _cots(1, H, W, 23)
    args = [to_dev(a, dev).requires_grad_(True) for a in (ia, ib, alb, nrm, band)]
    outs = ops.spatial_filter(*args, win_size=11)
    L = sum((o * to_dev(c, dev)).sum() for o, c in zip(outs, cots))
    L.backward()
    got = [to_np(a.grad) for a in args]
    ref_band = O.spatial_filter_grad(*cots, ia, ib, alb, nrm, band)
    ref_in = O.spatial_filter_grad_inputs(*cots, ia, ib, alb, nrm, band)
    r = np.random.default_rng(24)
    ys, xs = r.integers(0, H, 1024), r.integers(0, W, 1024)
    for name, g, rr in zip(("imga", "imgb", "albedo", "normal", "band"), got, list(ref_in) + [ref_band]):
        err = rel_l2(g[0, ys, xs], rr[0, ys, xs])
        _report(f"spatial_filter autograd d{name} 512x512 (1024 px) vs FP64 oracle", err)
        assert err <= TOL, name
    # central differences of the band gradient: band_p only enters the outputs at p
    cw = np.concatenate([c[0, ys, xs].astype(np.float64) for c in cots], 1)
    b64 = band[0, ys, xs].astype(np.float64)
    h = 1e-4
    fd = np.zeros((1024, 5))
    for k in range(5):
        bp, bm = b64.copy(), b64.copy()
        bp[:, k] += h
        bm[:, k] -= h
        fp = (_sf_local_forward(ia, ib, alb, nrm, bp, ys, xs) * cw).sum(1)
        fm = (_sf_local_forward(ia, ib, alb, nrm, bm, ys, xs) * cw).sum(1)
        fd[:, k] = (fp - fm) / (2 * h)
    err = rel_l2(got[4][0, ys, xs], fd)
    _report("spatial_filter autograd dband 512x512 (1024 px) vs FP64 central differences", err)
    assert err <= TOL
    assert rel_l2(ref_band[0, ys, xs], fd) <= 1e-6  # the oracle's analytic gradient, same check


# ------------------------------------------------------------------------------------------------
def reproject_case(H, W, seed, pan=(1.3, -0.6), dis=0.1, B=1):
    from xdn import synth
    out = []
    for b in range(B):
        spec = synth.StreamSpec(H, W, seed=seed, stream=b, pan=pan, disocclusion=dis)
        f0, f1 = synth.make_frame(spec, 0), synth.make_frame(spec, 1)
        out.append((f0, f1))
    cat = lambda k, i: np.concatenate([to_np(o[i][k]) for o in out], 0)
    r = np.random.default_rng(seed)
    prev_imgs = [r.uniform(0, 2, (B, H, W, 3)).astype(np.float32), r.uniform(0, 2, (B, H, W, 3)).astype(np.float32),
                 r.uniform(0, 2, (B, H, W, 1)).astype(np.float32)]
    opac = np.ones((B, H, W, 1), np.float32)
    popac = np.ones((B, H, W, 1), np.float32)
    popac[:, H // 3, :] = 0.0  # a row of transparent history
    opac[:, :, W // 5] = 0.0
    return (cat("mvec", 1), prev_imgs, cat("linearZ", 1), cat("linearZ", 0), cat("normal", 1), cat("normal", 0),
            cat("pnFwidth", 1), opac, popac)


@pytest.mark.parametrize("B,H,W,pan", [(1, 64, 80, (1.3, -0.6)), (2, 45, 61, (-2.7, 2.2)), (1, 270, 480, (3.0, -3.0)),
                                       (1, 33, 35, (0.0, 0.0)), (1, 40, 40, (-0.4, 0.45))])
def test_reproject_bit_exact_mask(ops, dev, B, H, W, pan):
    case = reproject_case(H, W, seed=H * 7 + W, pan=pan, B=B)
    mvec, ins, *rest = case
    s_ref, o_ref = O.reproject(mvec, ins, *rest)
    s, outs = ops.Reproject(to_dev(mvec, dev), [to_dev(a, dev) for a in ins], *[to_dev(a, dev) for a in rest])
    assert s.dtype == torch.bool
    s = to_np(s)
    n_diff = int((s != s_ref).sum())
    print(f"  reproject mask {B}x{H}x{W} pan={pan}: {int(s_ref.sum())} valid, {n_diff} differ", flush=True)
    assert n_diff == 0
    for g, r in zip(outs, o_ref):
        err = rel_l2(to_np(g), r)
        _report(f"reproject values {B}x{H}x{W}", err)
        assert err <= 1e-6
        assert np.all(to_np(g)[~s[..., 0]] == 0)


def test_reproject_many_buffers(ops, dev):
    """More buffers than one launch carries: split launches, identical masks."""
    H, W = 32, 48
    mvec, ins, *rest = reproject_case(H, W, seed=5)
    r = np.random.default_rng(1)
    ins = [r.uniform(0, 1, (1, H, W, c)).astype(np.float32) for c in (3, 1, 2, 4, 3, 1)]
    s_ref, o_ref = O.reproject(mvec, ins, *rest)
    s, outs = ops.Reproject(to_dev(mvec, dev), [to_dev(a, dev) for a in ins], *[to_dev(a, dev) for a in rest])
    assert np.array_equal(to_np(s), s_ref)
    for g, rr in zip(outs, o_ref):
        assert rel_l2(to_np(g), rr) <= 1e-6


@pytest.mark.parametrize("H,W,pan,dis", [(64, 80, (1.3, -0.6), 0.1), (96, 130, (-2.7, 2.2), 0.3),
                                         (40, 40, (-0.4, 0.45), 0.6)])
def test_reproject_three_rgb_buffers(ops, dev, H, W, pan, dis):
    """Three 3-channel buffers (the training step's list, with its own kernel): mask bit-exact, the
    bilinear and 3x3-fallback averages against the oracle, with many disoccluded pixels."""
    mvec, ins, *rest = reproject_case(H, W, seed=H + W, pan=pan, dis=dis)
    r = np.random.default_rng(H)
    ins = [r.uniform(0, 2, (1, H, W, 3)).astype(np.float32) for _ in range(3)]
    s_ref, o_ref = O.reproject(mvec, ins, *rest)
    s, outs = ops.Reproject(to_dev(mvec, dev), [to_dev(a, dev) for a in ins], *[to_dev(a, dev) for a in rest])
    s = to_np(s)
    assert np.array_equal(s, s_ref) and 0 < s_ref.sum() < s_ref.size
    for g, rr in zip(outs, o_ref):
        assert rel_l2(to_np(g), rr) <= 1e-6
        assert np.all(to_np(g)[~s[..., 0]] == 0)


# ------------------------------------------------------------------------------------------------
def screening_case(H, W, seed, B=1):
    ia, ib = log_images(H, W, seed=seed, B=B)
    r = np.random.default_rng(seed)
    # history: mostly close to the current frame so both outcomes occur
    p = ((ia + ib) * 0.5 + r.normal(0, 0.05, ia.shape)).astype(np.float32)
    p1 = (ia + r.normal(0, 0.05, ia.shape)).astype(np.float32)
    p2 = (ib + r.normal(0, 0.3, ia.shape)).astype(np.float32)
    m = r.random((B, H, W, 1)) < 0.9
    return ia, ib, p, p1, p2, m


@pytest.mark.parametrize("B,H,W,wc,wp", [(1, 64, 64, 7, 1), (1, 37, 45, 7, 1), (2, 29, 70, 7, 1), (1, 40, 36, 5, 3),
                                         (1, 270, 480, 7, 1)])
def test_temporal_screening_bit_exact(ops, dev, B, H, W, wc, wp):
    ia, ib, p, p1, p2, m = screening_case(H, W, seed=H + W, B=B)
    rs = O.erosion(m)
    ref = O.temporal_screening(ia, ib, p, p1, p2, rs, wc, wp)
    got = ops.temporal_screening(*[to_dev(a, dev) for a in (ia, ib, p, p1, p2, rs)], win_curr=wc, win_prev=wp)
    for name, g, r in zip(("output", "output1", "output2", "success"), got, ref):
        nd = int((to_np(g) != r).sum())
        print(f"  temporal_screening {name} {B}x{H}x{W}: {nd} elements differ", flush=True)
        assert nd == 0, name
    # fused erosion variant
    got = ops.erode_temporal_screening(*[to_dev(a, dev) for a in (ia, ib, p, p1, p2)], to_dev(m, dev), wc, wp)
    for g, r in zip(got, list(ref) + [rs]):
        assert np.array_equal(to_np(g), r)


# ------------------------------------------------------------------------------------------------
def train_case(H, W, seed, B=1):
    ia, ib = log_images(H, W, seed=seed, B=B)
    r = np.random.default_rng(seed + 1)
    alb, nrm = rand_img(H, W, seed=seed + 2, B=B), unit_normals(H, W, seed=seed + 3, B=B)
    param = r.normal(0, 1.0, (B, H, W, 6)).astype(np.float32)
    succ = (r.random((B, H, W, 1)) < 0.8).astype(np.float32)
    sprev = ((r.random((B, H, W, 1)) < 0.8) * succ).astype(np.float32)
    prev_out = (ia + r.normal(0, 0.1, ia.shape)).astype(np.float32)
    opac = (r.random((B, H, W, 1)) < 0.95).astype(np.float32)
    pols1 = (ia + r.normal(0, 0.1, ia.shape)).astype(np.float32)
    pols2 = (ib + r.normal(0, 0.1, ia.shape)).astype(np.float32)
    return ia, ib, alb, nrm, param, succ, sprev, prev_out, opac, pols1, pols2


def train_oracle(ia, ib, alb, nrm, param, succ, sprev, prev_out, opac, pols1, pols2):
    return O.filter_train(ia, ib, alb, nrm, param, succ, sprev, prev_out, opac, pols1, pols2)


@pytest.mark.parametrize("B,H,W", [(1, 64, 64), (1, 37, 51), (2, 48, 40)])
def test_filter_train(dev, B, H, W):
    from xdn import pipeline
    case = train_case(H, W, seed=H + W, B=B)
    ref = train_oracle(*case)
    got = pipeline.filter_train(*[to_dev(a, dev) for a in case], want_loss=True)
    names = ("out_log", "out", "loss", "avg_loss", "dparam")
    for name, g, r in zip(names, got, ref):
        err = rel_l2(to_np(g), r)
        _report(f"filter_train {name} {B}x{H}x{W}", err)
        assert err <= TOL, name


def test_adam_keras(dev):
    from xdn._lib import check, lib, ptr
    r = np.random.default_rng(0)
    n = 1000
    p, g = r.normal(size=n).astype(np.float32), r.normal(size=n).astype(np.float32)
    m, v = np.zeros(n, np.float32), np.zeros(n, np.float32)
    P, G, M, V = (to_dev(a, dev) for a in (p, g, m, v))
    pr, mr, vr = p.astype(np.float64), m.astype(np.float64), v.astype(np.float64)
    for step in (1, 2, 3):
        check(lib.xdn_adam(ptr(P), ptr(G), ptr(M), ptr(V), n, step, 1e-3, 0.9, 0.999, 1e-7,
                           torch.cuda.current_stream().cuda_stream), "adam")
        alpha = 1e-3 * np.sqrt(1 - 0.999 ** step) / (1 - 0.9 ** step)
        mr += (g - mr) * 0.1
        vr += (g.astype(np.float64) ** 2 - vr) * 0.001
        pr -= alpha * mr / (np.sqrt(vr) + 1e-7)
    torch.cuda.synchronize()
    assert rel_l2(to_np(P), pr) < 1e-6


# ------------------------------------------------------------------------------------------------
def test_config0_forward_chain(ops, dev):
    """The plain method on one 256x256 frame with the previous frame as history (mvec zero,
    bandwidths fixed at 1): cross-regression, reprojection of the previous results, erosion +
    temporal screening and the spatial filter, each stage against the FP64 oracle on that stage's
    own inputs (masks and selections bit-exact, values <= 1e-4 relative L2)."""
    from xdn import synth
    H = W = 256
    spec = synth.StreamSpec(H, W, seed=0, pan=(0.0, 0.0))
    f0, f1 = synth.make_frame(spec, 0), synth.make_frame(spec, 1)
    n = lambda t: t.detach().cpu().numpy()
    g = lambda a: to_dev(np.asarray(a, np.float32), dev)
    ones5 = np.ones((1, H, W, 5), np.float32)

    def regress(fr):
        y1 = O.prepare_input(n(fr["color"])).astype(np.float32)
        y2 = O.prepare_input(n(fr["color2"])).astype(np.float32)
        got = ops.cross_regression(g(y1), g(y2), g(n(fr["albedo"])), g(n(fr["normal"])), win_size=17, sparsity=4)
        ref = O.cross_regression(y1, y2, n(fr["albedo"]), n(fr["normal"]))
        for a, r in zip(got, ref):
            assert rel_l2(n(a), r) <= TOL
        return n(got[0]), n(got[1])

    # frame 0: no history; its filtered output (bandwidths 1) becomes the history
    o1, o2 = regress(f0)
    acc = ops.spatial_filter(g(o1), g(o2), g(n(f0["albedo"])), g(n(f0["normal"])), g(ones5), win_size=11)
    acc_ref = O.spatial_filter(o1, o2, n(f0["albedo"]), n(f0["normal"]), ones5)
    for a, r in zip(acc, acc_ref):
        assert rel_l2(n(a), r) <= TOL
    out_log0 = ((n(acc[0]) + n(acc[1])) / (n(acc[2]) + n(acc[3]) + 1e-4)).astype(np.float32)
    # frame 1
    p1, p2 = regress(f1)
    hist = [out_log0, o1, o2]
    ones1 = np.ones((1, H, W, 1), np.float32)
    rp_args = (n(f1["linearZ"]), n(f0["linearZ"]), n(f1["normal"]), n(f0["normal"]), n(f1["pnFwidth"]), ones1, ones1)
    s, rp = ops.Reproject(g(n(f1["mvec"])), [g(h) for h in hist], *[g(a) for a in rp_args])
    s_ref, rp_ref = O.reproject(n(f1["mvec"]), hist, *rp_args)
    assert np.array_equal(n(s), s_ref)
    # mvec zero, identical depth/normal: everything but row 0 / column 0 reprojects
    assert s_ref[0, 1:, 1:].all() and not s_ref[0, 0, :].any() and not s_ref[0, :, 0].any()
    for a, r in zip(rp, rp_ref):
        assert rel_l2(n(a), r) <= 1e-6
    se = O.erosion(n(s).astype(bool))
    scr = ops.temporal_screening(g(p1), g(p2), *rp, g(se), win_curr=7, win_prev=1)
    scr_ref = O.temporal_screening(p1, p2, *[n(a) for a in rp], se, 7, 1)
    for a, r in zip(scr, scr_ref):
        assert np.array_equal(n(a), r)
    acc1 = ops.spatial_filter(g(p1), g(p2), g(n(f1["albedo"])), g(n(f1["normal"])), g(ones5), win_size=11)
    acc1_ref = O.spatial_filter(p1, p2, n(f1["albedo"]), n(f1["normal"]), ones5)
    for a, r in zip(acc1, acc1_ref):
        err = rel_l2(n(a), r)
        assert err <= TOL
    _report("configuration 0 chain, last spatial filter acca", rel_l2(n(acc1[0]), acc1_ref[0]))


def test_empty_batch(ops, dev):
    """Zero-image batches through the drop-in ops give zero-image outputs; the network backwards,
    whose weight gradient has no batch dimension, write zeros."""
    import ctypes
    from xdn._lib import lib
    H, W = 32, 48
    e3 = torch.empty((0, H, W, 3), device=dev)
    e1 = torch.empty((0, H, W, 1), device=dev)
    outa, outb, va, vb = ops.cross_regression(e3, e3, e3, e3, win_size=17, sparsity=4)
    assert outa.shape == (0, H, W, 3) and va.shape == (0, H, W, 1)
    acc = ops.spatial_filter(e3, e3, e3, e3, torch.empty((0, H, W, 5), device=dev), win_size=11)
    assert all(t.shape[0] == 0 for t in acc)
    o = ops.temporal_screening(e3, e3, e3, e3, e3, e1, 7, 1)
    assert all(t.shape[0] == 0 for t in o)
    s, outs = ops.Reproject(e3, [e3, e3, e1], e3, e3, e3, e3, e3, e1, e1)
    assert s.shape == (0, H, W, 1) and [t.shape[3] for t in outs] == [3, 3, 1]
    for count, fn in ((lib.xdn_cnn5_param_count(), lib.xdn_cnn5_backward),
                      (lib.xdn_unet_param_count(), lib.xdn_unet_backward)):
        grad = torch.full((int(count),), float("nan"), device=dev)
        rc = fn(*([None] * 7), ctypes.c_void_p(grad.data_ptr()), None, 0, 0, H, W, None)
        assert rc == 0, lib.xdn_last_error().decode()
        torch.cuda.synchronize()
        assert torch.count_nonzero(grad).item() == 0
