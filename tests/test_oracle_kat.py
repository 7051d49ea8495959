"""Known-answer tests of the FP64 oracle, written from the reference's definitions by hand.

The reference ships no golden vectors, so these pin the terms that a consistent slip in both the
oracle's forward and its adjoint would leave invisible to the finite-difference and restatement
checks of test_oracle.py:
  - the sign of the reconstruction delta (pixel minus center, ops/regression.cu.cc:341-348);
  - the regression weight, delta_1 denominator, ridge and target operand of one center's normal
    equations on a two-valued image (regression.cu.cc:278-287, 395-445, 458-461);
  - which distance each of the five filter bandwidths weights (regression.cu.cc:133-146);
  - the aggregation loss's 1/2 factors and temporal blend (scripts/main.py:214-248);
  - the label crossing of train_step (scripts/main.py:283-286).

Each known answer is a ``kat_*`` function of the oracle module, so tests/test_oracle_mutations.py
can run it against copies of the oracle with one planted mistake and show that it fails there.
"""
import numpy as np
import pytest

from oracle import oracle as O

F32 = np.float32


def _close(got, expect, rtol, atol=0.0, what=""):
    got = np.asarray(got, np.float64)
    expect = np.asarray(expect, np.float64)
    assert got.shape == expect.shape, (what, got.shape, expect.shape)
    bad = np.abs(got - expect) > atol + rtol * np.abs(expect)
    assert not bad.any(), f"{what}: max |diff| {np.max(np.abs(got - expect)):.3e}"


# ------------------------------------------------------------------------------------------------
# reconstruction: a pixel that sees exactly one center
# ------------------------------------------------------------------------------------------------
def kat_recon_single_center(O):
    """win 3, sparsity 4: pixels x in {c-1, c, c+1} of the centers c in {4, 8} see only that center
    (the clamped range of regression.cu.cc:312-315), so the weight cancels and
    out = max(0, beta . delta) with delta = [1, (f_p - f_c)/(sqrt v_p + sqrt v_c + 1e-2),
    a_p - a_c, n_p - n_c] (regression.cu.cc:341-360)."""
    H = W = 12
    r = np.random.default_rng(901)
    feat = r.uniform(0.2, 1.5, (1, H, W, 3)).astype(F32)
    alb = r.uniform(0, 1, (1, H, W, 3)).astype(F32)
    n = r.normal(size=(1, H, W, 3))
    nrm = (n / np.linalg.norm(n, axis=-1, keepdims=True)).astype(F32)
    var = O.variance(feat)
    beta = np.zeros((1, H // 4, W // 4, 6, 3))
    beta[..., 1:, :] = r.uniform(-1, 1, beta[..., 1:, :].shape)
    beta[..., 0, :] = 5.0
    out = O.recon(beta, feat, var, alb, nrm, win_size=3, sparsity=4)
    f = feat[0].astype(np.float64)
    a = alb[0].astype(np.float64)
    nn = nrm[0].astype(np.float64)
    v = var[0, ..., 0]
    checked = 0
    for cy in (4, 8):
        for cx in (4, 8):
            b = beta[0, cy // 4, cx // 4]
            for py in (cy - 1, cy, cy + 1):
                for px in (cx - 1, cx, cx + 1):
                    den = np.sqrt(v[py, px]) + np.sqrt(v[cy, cx]) + 1e-2
                    for ch in range(3):
                        delta = [1.0, (f[py, px, ch] - f[cy, cx, ch]) / den, a[py, px, ch] - a[cy, cx, ch],
                                 nn[py, px, 0] - nn[cy, cx, 0], nn[py, px, 1] - nn[cy, cx, 1],
                                 nn[py, px, 2] - nn[cy, cx, 2]]
                        expect = max(0.0, sum(b[k, ch] * delta[k] for k in range(6)))
                        assert expect > 0  # the clamp must not hide the sign
                        _close(out[0, py, px, ch], expect, 1e-12, 1e-13, f"recon ({py},{px},{ch})")
                        checked += 1
    assert checked == 4 * 9 * 3


# ------------------------------------------------------------------------------------------------
# regression: one center's normal equations on a checkerboard
# ------------------------------------------------------------------------------------------------
def kat_regression_checkerboard(O):
    """Own half f and target t are two-valued checkerboards (f0/t0 where y+x is even, f1/t1 where it
    is odd); albedo and normal are constant. For the center (12,12) of a 24x24 image the 17x17
    window stays inside rows/columns 1..22, where every pixel has 4 equal and 4 opposite
    neighbours, so var = |(f0+f1)/2 - f_self|^2 = D/4 with D = |f1-f0|^2 (regression.cu.cc:244-276).
    The 145 same-colour taps have w = 1 + 1e-7 and delta_1 = 0; the 144 opposite ones
    w = exp(-D / (D/4 + D/4 + 1e-4)) + 1e-7 and delta_1 = (f1-f0)_c / (2 sqrt(D/4) + 1e-2).
    Albedo and normal deltas vanish, so per channel the system is 2x2 in (beta_0, beta_1) plus the
    1e-3 ridge (regression.cu.cc:458-461), and beta_2..5 = 0; solved here by Cramer's rule."""
    H = W = 24
    f0, f1 = np.array([0.2, 0.5, 0.9]), np.array([0.6, 0.1, 0.4])
    t0, t1 = np.array([1.0, 2.0, 0.5]), np.array([0.3, 1.1, 2.5])
    yy, xx = np.mgrid[0:H, 0:W]
    odd = ((yy + xx) % 2 == 1)[..., None]
    feat = np.where(odd, f1, f0)[None].astype(F32)
    target = np.where(odd, t1, t0)[None].astype(F32)
    alb = np.full((1, H, W, 3), 0.4, F32)
    nrm = np.broadcast_to(np.array([0, 0, 1], F32), (1, H, W, 3)).copy()
    f0, f1 = f0.astype(F32).astype(np.float64), f1.astype(F32).astype(np.float64)
    t0, t1 = t0.astype(F32).astype(np.float64), t1.astype(F32).astype(np.float64)
    D = float(np.sum((f1 - f0) ** 2))
    V = D / 4
    var = O.variance(feat)
    _close(var[0, 1:-1, 1:-1, 0], np.full((H - 2, W - 2), V), 1e-12, 0, "checkerboard variance")
    beta = O.regression(target, feat, var, alb, nrm, win_size=17, sparsity=4)[0, 3, 3]
    ws = 1.0 + 1e-7
    wo = np.exp(-D / (2 * V + 1e-4)) + 1e-7
    e = (f1 - f0) / (2 * np.sqrt(V) + 1e-2)  # delta_1 of an opposite tap, per channel
    for ch in range(3):
        a00 = 145 * ws + 144 * wo + 1e-3
        a01 = 144 * wo * e[ch]
        a11 = 144 * wo * e[ch] ** 2 + 1e-3
        b0 = 145 * ws * t0[ch] + 144 * wo * t1[ch]
        b1 = 144 * wo * e[ch] * t1[ch]
        det = a00 * a11 - a01 * a01
        beta0 = (b0 * a11 - a01 * b1) / det
        beta1 = (a00 * b1 - a01 * b0) / det
        _close(beta[0, ch], beta0, 1e-10, 0, f"beta_0 ch{ch}")
        _close(beta[1, ch], beta1, 1e-10, 0, f"beta_1 ch{ch}")
        _close(beta[2:, ch], np.zeros(4), 0, 1e-12, f"beta_2..5 ch{ch}")


# ------------------------------------------------------------------------------------------------
# spatial filter: one bandwidth at a time
# ------------------------------------------------------------------------------------------------
SF_H, SF_W, SF_HALF = 14, 16, 5


def _same_other(region, half=SF_HALF):
    """Per position: how many of the 11 taps (reflect-101 boundary, numpy.pad 'reflect') lie in the
    position's own region and how many in the other one."""
    pad = np.pad(region, half, mode="reflect")
    win = np.stack([pad[i:i + region.size] for i in range(2 * half + 1)])
    same = (win == region[None]).sum(0)
    return same.astype(np.float64), (2 * half + 1 - same).astype(np.float64)


def _sf_base(r):
    H, W = SF_H, SF_W
    ia = np.broadcast_to(np.array([0.7, 1.3, 0.2], F32), (1, H, W, 3)).copy()
    ib = np.broadcast_to(np.array([0.4, 0.9, 1.6], F32), (1, H, W, 3)).copy()
    alb = np.full((1, H, W, 3), 0.5, F32)
    nrm = np.broadcast_to(np.array([0.0, 0.6, 0.8], F32), (1, H, W, 3)).copy()
    return ia, ib, alb, nrm


def _band(k, value=0.8):
    b = np.zeros((1, SF_H, SF_W, 5), F32)
    b[..., k] = value
    return b


def kat_sf_albedo_band(O):
    """Only b2 set; albedo two-valued along x: w = exp(-b2 |a1-a0|^2) across the edge, 1 elsewhere,
    the same for both halves (regression.cu.cc:142-143)."""
    ia, ib, alb, nrm = _sf_base(None)
    a0, a1 = np.array([0.1, 0.3, 0.2], F32), np.array([0.8, 0.5, 0.9], F32)
    reg = np.arange(SF_W) >= 7
    alb[0] = np.where(reg[None, :, None], a1, a0)
    e = np.exp(-float(np.float32(0.8)) * float(np.sum((a1.astype(np.float64) - a0.astype(np.float64)) ** 2)))
    same, other = _same_other(reg)
    ws = np.broadcast_to(11 * (same + other * e), (SF_H, SF_W))
    accA, accB, wA, wB = O.spatial_filter(ia, ib, alb, nrm, _band(2))
    _close(wA[0, ..., 0], ws, 1e-12, 0, "wsumA (albedo band)")
    _close(wB[0, ..., 0], ws, 1e-12, 0, "wsumB (albedo band)")
    _close(accA[0], ws[..., None] * ia[0].astype(np.float64), 1e-12, 0, "accA (albedo band)")


def kat_sf_normal_band(O):
    """Only b3 set; normal two-valued along y: w = exp(-b3 |n1-n0|^2) across the edge."""
    ia, ib, alb, nrm = _sf_base(None)
    n0, n1 = np.array([0.0, 0.6, 0.8], F32), np.array([0.6, 0.0, 0.8], F32)
    reg = np.arange(SF_H) >= 6
    nrm[0] = np.where(reg[:, None, None], n1, n0)
    e = np.exp(-float(np.float32(0.8)) * float(np.sum((n1.astype(np.float64) - n0.astype(np.float64)) ** 2)))
    same, other = _same_other(reg)
    ws = np.broadcast_to((11 * (same + other * e))[:, None], (SF_H, SF_W))
    accA, accB, wA, wB = O.spatial_filter(ia, ib, alb, nrm, _band(3))
    _close(wA[0, ..., 0], ws, 1e-12, 0, "wsumA (normal band)")
    _close(wB[0, ..., 0], ws, 1e-12, 0, "wsumB (normal band)")
    _close(accB[0], ws[..., None] * ib[0].astype(np.float64), 1e-12, 0, "accB (normal band)")


def kat_sf_pixel_band(O):
    """Only b4 set: w = exp(-b4 ((dx/5)^2 + (dy/5)^2)) (the integer half width, :133-135), a
    separable Gaussian whose sum is (sum_d exp(-b4 d^2/25))^2 at every pixel."""
    ia, ib, alb, nrm = _sf_base(None)
    b4 = float(np.float32(0.8))
    g = sum(np.exp(-b4 * d * d / 25.0) for d in range(-5, 6))
    accA, accB, wA, wB = O.spatial_filter(ia, ib, alb, nrm, _band(4))
    _close(wA, np.full(wA.shape, g * g), 1e-12, 0, "wsumA (pixel band)")
    _close(wB, np.full(wB.shape, g * g), 1e-12, 0, "wsumB (pixel band)")


def kat_sf_color_bands(O):
    """b0 weights half A by its own colour distance only, b1 half B only (:140-141): imga two-valued
    along x with only b0 set leaves half B a plain box filter (wsum 121), and vice versa."""
    base = _sf_base(None)
    c0, c1 = np.array([0.2, 0.4, 0.1], F32), np.array([0.9, 0.3, 1.2], F32)
    e = np.exp(-float(np.float32(0.8)) * float(np.sum((c1.astype(np.float64) - c0.astype(np.float64)) ** 2)))
    reg = np.arange(SF_W) >= 9
    same, other = _same_other(reg)
    own = np.where(reg[:, None], c1, c0).astype(np.float64)
    opp = np.where(reg[:, None], c0, c1).astype(np.float64)
    ws = 11 * (same + other * e)
    acc = 11 * (same[:, None] * own + (other * e)[:, None] * opp)
    for k in (0, 1):
        ia, ib, alb, nrm = [x.copy() for x in base]
        two = np.broadcast_to(np.where(reg[None, :, None], c1, c0), (SF_H, SF_W, 3))
        if k == 0:
            ia[0] = two
        else:
            ib[0] = two
        accA, accB, wA, wB = O.spatial_filter(ia, ib, alb, nrm, _band(k))
        w_two, w_box = (wA, wB) if k == 0 else (wB, wA)
        acc_two = accA if k == 0 else accB
        _close(w_two[0, ..., 0], np.broadcast_to(ws, (SF_H, SF_W)), 1e-12, 0, f"two-valued half, b{k}")
        _close(acc_two[0], np.broadcast_to(acc, (SF_H, SF_W, 3)), 1e-12, 0, f"acc two-valued half, b{k}")
        _close(w_box, np.full(w_box.shape, 121.0), 1e-12, 0, f"other half, b{k}")


# ------------------------------------------------------------------------------------------------
# aggregation loss and the label crossing
# ------------------------------------------------------------------------------------------------
def kat_aggregation_loss(O):
    """Two pixels of one channel, every term of scripts/main.py:214-248 written out by hand.
    Pixel 0: success = success_prev = 1, alpha = 0.25 (temporal blend); pixel 1: success_prev = 0,
    so alpha' = 1 (spatial only)."""
    def col(v0, v1):
        return np.array([v0, v1], np.float64).reshape(1, 1, 2, 1)

    sum1, sumw1 = col(2.0, 1.0), col(0.9999, 1.9999)  # out1 = 2, 0.5
    sum2, sumw2 = col(4.0, 3.0), col(0.9999, 0.9999)  # out2 = 4, 3
    alpha = col(0.25, 0.6)
    success, success_prev = col(1.0, 1.0), col(1.0, 0.0)
    prev_out, opacity = col(4.0, 9.0), col(1.0, 1.0)
    label1, label2 = col(1.0, 0.5), col(3.0, 2.0)
    prev_label1, prev_label2 = col(0.0, 1.0), col(5.0, 4.0)
    out_log, out_lin, loss, avg = O.aggregation_loss(sum1, sumw1, sum2, sumw2, alpha, success, success_prev,
                                                     prev_out, opacity, label1, label2, prev_label1, prev_label2)
    # pixel 0: out1 = 2, out2 = 4, out = 6/2 = 3; blended with 0.25 / 0.75 against prev_out = 4
    o = 0.25 * (6.0 / (2.0 - 2e-4 + 1e-4)) + 0.75 * 4.0
    o1 = 0.25 * 2.0 + 0.75 * 4.0  # 3.5
    o2 = 0.25 * 4.0 + 0.75 * 4.0  # 4.0
    sp = 0.5 * ((o1 - 1.0) ** 2 / (1.0 + 1e-2) + (o2 - 3.0) ** 2 / (9.0 + 1e-2))
    tp = 0.5 * ((o1 - 0.0) ** 2 / (0.0 + 1e-2) + (o2 - 5.0) ** 2 / (25.0 + 1e-2))
    l0 = 0.5 * (sp + tp)
    # pixel 1: no blend; out1 = 1/2, out2 = 3, out = 4/(1.9999+0.9999+1e-4)
    q = 4.0 / (1.9999 + 0.9999 + 1e-4)
    q1, q2 = 0.5, 3.0
    sp = 0.5 * ((q1 - 0.5) ** 2 / (0.25 + 1e-2) + (q2 - 2.0) ** 2 / (4.0 + 1e-2))
    tp = 0.5 * ((q1 - 1.0) ** 2 / (1.0 + 1e-2) + (q2 - 4.0) ** 2 / (16.0 + 1e-2))
    l1 = 0.5 * (sp + tp)
    _close(out_log.ravel(), [o, q], 1e-12, 0, "aggregation out")
    _close(loss.ravel(), [l0, l1], 1e-9, 0, "aggregation loss")
    _close(avg, (l0 + l1) / 2, 1e-9, 0, "aggregation mean loss")
    # a transparent pixel contributes neither output nor loss
    _, _, loss_t, _ = O.aggregation_loss(sum1, sumw1, sum2, sumw2, alpha, success, success_prev, prev_out,
                                         col(0.0, 1.0), label1, label2, prev_label1, prev_label2)
    assert loss_t.ravel()[0] == 0.0 and loss_t.ravel()[1] == pytest.approx(l1, rel=1e-9)


def kat_filter_train_labels(O):
    """train_step's labels cross over (scripts/main.py:283-286): out1 (filtered ols1) is scored against
    ols2 and prev_ols2, out2 against ols1 and prev_ols1. A network output of 0 for the five bands
    (band = 1/(0 + 1e-4) = 1e4) turns the filter into the identity (every off-center weight is at most
    exp(-1e4/25)), and success = 0 makes alpha' = 1, so out1 = ols1 / (1 + 1e-4)."""
    H, W = 5, 6
    r = np.random.default_rng(907)
    ols1, ols2, pols1, pols2, prev_out = [r.uniform(0.2, 2.0, (1, H, W, 3)).astype(F32) for _ in range(5)]
    alb = r.uniform(0, 1, (1, H, W, 3)).astype(F32)
    nrm = np.broadcast_to(np.array([0, 0, 1], F32), (1, H, W, 3)).copy()
    param = np.zeros((1, H, W, 6), F32)
    param[..., 5] = 0.3
    zero = np.zeros((1, H, W, 1), F32)
    one = np.ones((1, H, W, 1), F32)
    out_log, out, loss, avg, _ = O.filter_train(ols1, ols2, alb, nrm, param, zero, one, prev_out, one, pols1, pols2)
    a, b, p1, p2 = (x.astype(np.float64) for x in (ols1, ols2, pols1, pols2))
    o1, o2 = a / (1 + 1e-4), b / (1 + 1e-4)
    spatial = 0.5 * ((o1 - b) ** 2 / (b ** 2 + 1e-2) + (o2 - a) ** 2 / (a ** 2 + 1e-2))
    temporal = 0.5 * ((o1 - p2) ** 2 / (p2 ** 2 + 1e-2) + (o2 - p1) ** 2 / (p1 ** 2 + 1e-2))
    expect = 0.5 * (spatial + temporal)
    _close(out_log, (a + b) / (2 + 1e-4), 1e-12, 0, "filter_train out")
    _close(loss, expect, 1e-12, 0, "filter_train loss")
    _close(avg, [expect.mean()], 1e-12, 0, "filter_train mean loss")


KATS = [kat_recon_single_center, kat_regression_checkerboard, kat_sf_albedo_band, kat_sf_normal_band,
        kat_sf_pixel_band, kat_sf_color_bands, kat_aggregation_loss, kat_filter_train_labels]


@pytest.mark.parametrize("kat", KATS, ids=[k.__name__ for k in KATS])
def test_oracle_known_answer(kat):
    kat(O)
