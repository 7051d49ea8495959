"""Parity of the U-Net kernels (the reference's bandwidth network, scripts/main.py:83-118)
against a float64 torch-CPU reference.

The reference rounds to bf16 where the kernels store bf16 (network input, 3x3 weights, every
activation, every backward gradient buffer), keeps the 1x1 layers' weights in FP32 (as the
kernels do), and computes everything else in float64 with library operators: conv2d and
torch.nn.grad for the 3x3 layers, conv_transpose2d (stride 2, output_padding 1) for the
transposed 1x1 layers, max_pool2d / max_unpool2d for the pools (torch's CPU max pool sends the
gradient to the first maximum of a window in row-major order, as the kernel does; the
reference's own TF max-pool gradient on ties is not pinned here).

Two bars:
  * every kernel on its own operands (the bf16 tensors it read, taken from the workspace):
    weight and bias gradients and the FP32 head output <= 1e-5 relative L2; bf16 outputs equal
    to the rounded float64 result except for rare one-ulp rounding-boundary cases; max pool
    bit-exact;
  * end to end: network output <= 5e-3 against the float64 forward, weight and bias gradients
    <= 1e-2 against the float64 backward chain run on the kernels' own activations.
"""
import ctypes

import numpy as np
import pytest
import torch
import torch.nn.functional as F

from helpers import rel_l2

# workspace buffers (level, channels, fp32) in the order of xdn_unet_workspace_layout
BUFS = [(0, 8, 0), (0, 8, 0), (1, 8, 0), (1, 16, 0), (1, 16, 0), (2, 16, 0), (2, 32, 0), (2, 32, 0),
        (1, 32, 0), (1, 16, 0), (1, 16, 0), (0, 16, 0), (0, 8, 0),
        (0, 8, 0), (0, 24, 0), (1, 16, 0), (1, 16, 0), (1, 48, 0), (2, 32, 0), (2, 32, 0), (2, 16, 0),
        (1, 16, 0), (1, 16, 0), (1, 8, 0), (0, 8, 0), (0, 8, 0)]
(A0, A1, P1, A2, A3, P2, A4, A5, A6, A7, A8, A9, A10, DY10, DC5, DY8, DY7, DC4, DY5, DY4, DP2, DY3, DY2, DP1,
 DY1, DY0) = range(26)


def bf(x):
    return x.to(torch.bfloat16).to(torch.float64)


def conv3(x, w, b=None):
    y = F.conv2d(x, bf(w.double()).permute(3, 2, 0, 1), padding=1)
    return y if b is None else y + b.double().view(1, -1, 1, 1)


def convT(x, w, b):
    co, ci = w.shape
    return F.conv_transpose2d(x, w.double().t().reshape(ci, co, 1, 1), stride=2, output_padding=1) + \
        b.double().view(1, -1, 1, 1)


def head(x, w, b):
    ci, co = w.shape
    return F.conv2d(x, w.double().t().reshape(co, ci, 1, 1)) + b.double().view(1, -1, 1, 1)


def unet_ref_forward(imgs, L):
    """float64 forward with the kernels' bf16 rounding points; returns (out NHWC, {buffer: act})."""
    x0 = bf(torch.cat(imgs, -1).double()).permute(0, 3, 1, 2)
    a = {}
    a[A0] = bf(torch.relu(conv3(x0, *L[0])))
    a[A1] = bf(torch.relu(conv3(a[A0], *L[1])))
    a[P1] = F.max_pool2d(a[A1], 2)
    a[A2] = bf(torch.relu(conv3(a[P1], *L[2])))
    a[A3] = bf(torch.relu(conv3(a[A2], *L[3])))
    a[P2] = F.max_pool2d(a[A3], 2)
    a[A4] = bf(torch.relu(conv3(a[P2], *L[4])))
    a[A5] = bf(torch.relu(conv3(a[A4], *L[5])))
    a[A6] = bf(convT(a[A5], *L[6]))
    a[A7] = bf(torch.relu(conv3(torch.cat([a[A6], a[A3]], 1), *L[7])))
    a[A8] = bf(torch.relu(conv3(a[A7], *L[8])))
    a[A9] = bf(convT(a[A8], *L[9]))
    a[A10] = bf(torch.relu(conv3(torch.cat([a[A9], a[A1]], 1), *L[10])))
    out = head(a[A10], *L[11])
    return out.permute(0, 2, 3, 1), a, x0


def unpool(dp, act):
    _, idx = F.max_pool2d(act, 2, return_indices=True)
    return F.max_unpool2d(dp, idx, 2, output_size=act.shape[-2:])


def wgrad3(x, w, dy):
    dw = torch.nn.grad.conv2d_weight(x, bf(w.double()).permute(3, 2, 0, 1).shape, dy, padding=1)
    return dw.permute(2, 3, 1, 0), dy.sum((0, 2, 3))


def dgrad3(x_shape, w, dy):
    return torch.nn.grad.conv2d_input(x_shape, bf(w.double()).permute(3, 2, 0, 1), dy, padding=1)


def convT_bwd(x, w, b, dout):
    """(dW [co][ci], db, dX unmasked) of convT given the gradient of its output."""
    g = dout[:, :, ::2, ::2]
    dw = torch.einsum("nihw,nohw->oi", x, g)
    dx = torch.einsum("oi,nohw->nihw", w.double(), g)
    return dw, dout.sum((0, 2, 3)), dx


def head_bwd(x, w, dparam):
    dw = torch.einsum("nihw,nohw->io", x, dparam)
    dx = torch.einsum("io,nohw->nihw", w.double(), dparam)
    return dw, dparam.sum((0, 2, 3)), dx


def unet_ref_backward(L, a, x0, dparam, take=None):
    """float64 backward chain. Activations `a` (dict buffer -> NCHW); with `take`, each gradient
    the kernels store is replaced by the kernels' own (take(buffer) -> NCHW) before it is used, so
    every step runs on exactly the operands its kernel read. Returns ({layer: (dW, db)},
    {buffer: reference gradient before replacement})."""
    g, r = {}, {}

    def keep(buf, val):
        r[buf] = val
        return take(buf) if take else (bf(val) if not BUFS[buf][2] else val)

    dp = dparam.permute(0, 3, 1, 2)
    dw, db, dx = head_bwd(a[A10], L[11][0], dp)
    g[11] = (dw, db)
    dy10 = keep(DY10, dx * (a[A10] > 0))
    x10 = torch.cat([a[A9], a[A1]], 1)
    g[10] = wgrad3(x10, L[10][0], dy10)
    dc5 = keep(DC5, dgrad3(x10.shape, L[10][0], dy10))
    dw, db, dx = convT_bwd(a[A8], *L[9], dc5[:, :16])
    g[9] = (dw, db)
    dy8 = keep(DY8, dx * (a[A8] > 0))
    g[8] = wgrad3(a[A7], L[8][0], dy8)
    dy7 = keep(DY7, dgrad3(a[A7].shape, L[8][0], dy8) * (a[A7] > 0))
    x7 = torch.cat([a[A6], a[A3]], 1)
    g[7] = wgrad3(x7, L[7][0], dy7)
    dc4 = keep(DC4, dgrad3(x7.shape, L[7][0], dy7))
    dw, db, dx = convT_bwd(a[A5], *L[6], dc4[:, :32])
    g[6] = (dw, db)
    dy5 = keep(DY5, dx * (a[A5] > 0))
    g[5] = wgrad3(a[A4], L[5][0], dy5)
    dy4 = keep(DY4, dgrad3(a[A4].shape, L[5][0], dy5) * (a[A4] > 0))
    g[4] = wgrad3(a[P2], L[4][0], dy4)
    dp2 = keep(DP2, dgrad3(a[P2].shape, L[4][0], dy4))
    dy3 = keep(DY3, (unpool(dp2, a[A3]) + dc4[:, 32:48]) * (a[A3] > 0))
    g[3] = wgrad3(a[A2], L[3][0], dy3)
    dy2 = keep(DY2, dgrad3(a[A2].shape, L[3][0], dy3) * (a[A2] > 0))
    g[2] = wgrad3(a[P1], L[2][0], dy2)
    dp1 = keep(DP1, dgrad3(a[P1].shape, L[2][0], dy2))
    dy1 = keep(DY1, (unpool(dp1, a[A1]) + dc5[:, 16:24]) * (a[A1] > 0))
    g[1] = wgrad3(a[A0], L[1][0], dy1)
    dy0 = keep(DY0, dgrad3(a[A0].shape, L[1][0], dy1) * (a[A0] > 0))
    g[0] = wgrad3(x0, L[0][0], dy0)
    return g, r


def bf16_mismatch(got, ref_exact):
    """(elements differing from the bf16-rounded reference, elements off by more than one ulp)."""
    ref = bf(ref_exact)
    diff = ref != got
    ulp = (ref.abs() * 2.0 ** -7).clamp_min(1e-30)
    far = ((ref - got).abs() > 1.01 * ulp) & diff
    return int(diff.sum()), int(far.sum())


def ws_tensor(net, i):
    B, H, W = net.shape
    lv, ch, f32 = BUFS[i]
    raw = net.workspace_buffer(i).view(torch.float32 if f32 else torch.bfloat16)
    return raw.view(B, H >> lv, W >> lv, ch).cpu().to(torch.float64).permute(0, 3, 1, 2)


def test_unet_param_layout_cpu():
    from xdn import cnn
    from xdn._lib import lib
    flat = cnn.flatten_params(cnn.unet_init_params(seed=1))
    assert flat.numel() == int(lib.xdn_unet_param_count()) == 31414
    back = cnn.flatten_params(cnn.unet_unflatten(flat))
    assert torch.equal(back, flat)
    # workspace buffers are disjoint, 256-byte aligned and sized [B][h][w][ch]
    B, H, W = 2, 36, 44
    spans = []
    for i, (lv, ch, f32) in enumerate(BUFS):
        off, nb = ctypes.c_size_t(), ctypes.c_size_t()
        assert lib.xdn_unet_workspace_layout(B, H, W, i, ctypes.byref(off), ctypes.byref(nb)) == 0
        assert off.value % 256 == 0
        assert nb.value == B * (H >> lv) * (W >> lv) * ch * (4 if f32 else 2)
        spans.append((off.value, off.value + nb.value))
    spans.sort()
    assert all(a[1] <= b[0] for a, b in zip(spans, spans[1:]))
    assert spans[-1][1] <= int(lib.xdn_unet_workspace_size(B, H, W))
    off, nb = ctypes.c_size_t(), ctypes.c_size_t()
    assert lib.xdn_unet_workspace_layout(B, H, W, 26, ctypes.byref(off), ctypes.byref(nb)) != 0


def test_unet_rejects_bad_shapes_cpu():
    """Argument checks run before anything touches the device."""
    from xdn._lib import lib
    ws = int(lib.xdn_unet_workspace_size(1, 32, 32))
    for B, H, W in ((1, 30, 32), (1, 32, 34), (-1, 32, 32), (17, 32, 32), (0, 30, 32)):
        rc = lib.xdn_unet_forward(None, None, None, None, None, None, None, None, ws, B, H, W, None)
        assert rc == -1, (B, H, W)
        msg = lib.xdn_last_error().decode()
        assert "unet" in msg
    rc = lib.xdn_unet_backward(None, None, None, None, None, None, None, None, None, 16, 1, 32, 32, None)
    assert rc == -1 and "workspace" in lib.xdn_last_error().decode()


def test_reference_backward_matches_autograd_cpu(monkeypatch):
    """The hand-written float64 backward chain above equals torch autograd through the same
    forward (with the bf16 roundings switched off), so the GPU tests compare against a pinned
    reference."""
    from xdn import cnn
    monkeypatch.setitem(globals(), "bf", lambda x: x.double())
    B, H, W = 2, 12, 16
    g = torch.Generator().manual_seed(3)
    imgs = [torch.rand(B, H, W, 3, generator=g, dtype=torch.float64) for _ in range(5)]
    L = [(w.double().requires_grad_(), (torch.randn(b.shape, generator=g, dtype=torch.float64) * 0.1).requires_grad_())
         for w, b in cnn.unet_init_params(seed=4)]
    out, a, x0 = unet_ref_forward(imgs, L)
    dparam = torch.randn(B, H, W, 6, generator=g, dtype=torch.float64)
    (out * dparam).sum().backward()
    Ld = [(w.detach(), b.detach()) for w, b in L]
    gr, _ = unet_ref_backward(Ld, {k: v.detach() for k, v in a.items()}, x0, dparam)
    for l in range(12):
        assert torch.allclose(gr[l][0], L[l][0].grad, rtol=1e-9, atol=1e-12), l
        assert torch.allclose(gr[l][1], L[l][1].grad, rtol=1e-9, atol=1e-12), l


@pytest.mark.gpu
@pytest.mark.parametrize("B,H,W", [(1, 48, 80), (2, 32, 64), (1, 36, 44), (1, 256, 320), (1, 1080, 1920)])
def test_unet_forward_backward(dev, B, H, W):
    from xdn import cnn
    torch.manual_seed(H * W + B)
    imgs = [torch.rand(B, H, W, 3) * 2 for _ in range(5)]
    net = cnn.HipUNet(dev, seed=11, shape=(B, H, W))
    for _, b in net.layers():  # non-zero biases exercise the bias paths
        b.copy_(torch.randn(b.shape, device=dev) * 0.05)
    net.pack()
    out = net.forward(*[t.to(dev).contiguous() for t in imgs])
    torch.cuda.synchronize()
    L = [(w.cpu(), b.cpu()) for w, b in net.layers()]
    ref_out, ref_a, x0 = unet_ref_forward(imgs, L)
    e = rel_l2(out.cpu().numpy(), ref_out.numpy())
    print(f"  unet forward {B}x{H}x{W}: output rel L2 {e:.2e}", flush=True)
    assert e <= 5e-3

    # every forward kernel on its own operands
    own = {i: ws_tensor(net, i) for i in range(26)}
    x_of = {A0: None, A1: own[A0], A2: own[P1], A3: own[A2], A4: own[P2], A5: own[A4],
            A7: torch.cat([own[A6], own[A3]], 1), A8: own[A7], A10: torch.cat([own[A9], own[A1]], 1)}
    layer_of = {A0: 0, A1: 1, A2: 2, A3: 3, A4: 4, A5: 5, A7: 7, A8: 8, A10: 10}
    n_el = 0
    for buf, l in layer_of.items():
        x = x0 if x_of[buf] is None else x_of[buf]
        d, far = bf16_mismatch(own[buf], torch.relu(conv3(x, *L[l])))
        n_el = own[buf].numel()
        print(f"  conv layer {l} forward on own operands: {d} of {n_el} differ, {far} by more than 1 ulp", flush=True)
        assert far <= 2 + 1e-5 * n_el and d <= 2e-3 * n_el + 2, l
    assert torch.equal(own[P1], F.max_pool2d(own[A1], 2)) and torch.equal(own[P2], F.max_pool2d(own[A3], 2))
    for buf, src, l in ((A6, A5, 6), (A9, A8, 9)):
        d, far = bf16_mismatch(own[buf], convT(own[src], *L[l]))
        print(f"  convT layer {l} forward on own operands: {d} differ, {far} by more than 1 ulp", flush=True)
        # the same size-scaled bar as the 3x3 layers (1080p: 3 of 3.3e7 elements beyond one ulp)
        assert far <= 2 + 1e-5 * own[buf].numel() and d <= 2e-3 * own[buf].numel() + 2
    e = rel_l2(out.cpu().numpy(), head(own[A10], *L[11]).permute(0, 2, 3, 1).numpy())
    print(f"  head forward on own operands: rel L2 {e:.2e}", flush=True)
    assert e <= 1e-5

    dparam = torch.randn(B, H, W, 6, dtype=torch.float64) * 1e-2
    net.backward(dparam.float().to(dev).contiguous())
    torch.cuda.synchronize()
    dparam = dparam.float().double()
    got = [(w.cpu().double(), b.cpu().double()) for w, b in cnn.unet_unflatten(net.grad)]
    own.update({i: ws_tensor(net, i) for i in range(13, 26)})
    # backward kernels on their own operands
    g_own, r_own = unet_ref_backward(L, own, x0, dparam, take=lambda i: own[i])
    for i in range(13, 26):
        if BUFS[i][2]:
            e = rel_l2(own[i].numpy(), r_own[i].numpy())
            print(f"  gradient buffer {i} (FP32) on own operands: rel L2 {e:.2e}", flush=True)
            assert e <= 1e-5, i
        else:
            d, far = bf16_mismatch(own[i], r_own[i])
            print(f"  gradient buffer {i} (bf16) on own operands: {d} of {own[i].numel()} differ, {far} by more",
                  flush=True)
            assert far <= 2 + 1e-5 * own[i].numel() and d <= 2e-3 * own[i].numel() + 2, i
    for l in range(12):
        ew = rel_l2(got[l][0].numpy(), g_own[l][0].numpy())
        eb = rel_l2(got[l][1].numpy(), g_own[l][1].numpy())
        print(f"  layer {l} weight gradient on own operands: dW {ew:.2e} db {eb:.2e}", flush=True)
        assert ew <= 1e-5 and eb <= 1e-5, l
    # end to end: the reference chain makes its own gradients from the kernels' activations
    acts = {i: own[i] for i in range(13)}
    g_ref, _ = unet_ref_backward(L, acts, x0, dparam)
    for l in range(12):
        ew = rel_l2(got[l][0].numpy(), g_ref[l][0].numpy())
        eb = rel_l2(got[l][1].numpy(), g_ref[l][1].numpy())
        print(f"  layer {l} weight gradient end to end: dW {ew:.2e} db {eb:.2e}", flush=True)
        assert ew <= 1e-2 and eb <= 1e-2, l


@pytest.mark.gpu
def test_unet_adam_step_and_repack(dev):
    """A Keras Adam step on the U-Net parameters, then the packed fragments follow the update."""
    from xdn import cnn
    B, H, W = 1, 32, 48
    net = cnn.HipUNet(dev, seed=5, shape=(B, H, W))
    imgs = [torch.rand(B, H, W, 3, device=dev) for _ in range(5)]
    net.forward(*imgs)
    net.backward(torch.randn(B, H, W, 6, device=dev) * 1e-2)
    p0 = net.params.clone()
    net.adam_step(1e-3)
    torch.cuda.synchronize()
    moved = (net.params - p0).abs()
    nz = net.grad.abs() > 1e-3
    assert int(nz.sum()) > 1000
    assert torch.allclose(moved[nz], torch.full_like(moved[nz], 1e-3), rtol=1e-2, atol=1e-6)
    out = net.forward(*imgs)
    torch.cuda.synchronize()
    L = [(w.cpu(), b.cpu()) for w, b in net.layers()]
    ref_out, _, _ = unet_ref_forward([t.cpu() for t in imgs], L)
    assert rel_l2(out.cpu().numpy(), ref_out.numpy()) <= 5e-3


@pytest.mark.gpu
@pytest.mark.parametrize("H,W", [(128, 160), (256, 320)])
def test_unet_error_against_unrounded_float64(dev, H, W, monkeypatch):
    """The U-Net's error in the reference's own arithmetic (FP32 Keras, scripts/main.py:83-118,
    stood in for by float64 with no bf16 rounding anywhere): the network output, every layer's
    weight gradient, and what the difference does downstream, the filtered radiance and the
    network-output gradient of the fused training filter (FP32 on the GPU for both)."""
    from test_gpu_parity import train_case
    from xdn import cnn, pipeline
    case = train_case(H, W, seed=H + 3)
    ia, ib, alb, nrm, _, succ, sprev, prev_out, opac, pols1, pols2 = case
    imgs = [ia, ib, prev_out, alb, nrm]
    net = cnn.HipUNet(dev, seed=1234, shape=(1, H, W))  # the step's Glorot-uniform initialisation
    out = net.forward(*[torch.from_numpy(x).to(dev) for x in imgs])
    torch.cuda.synchronize()
    L = [(w.cpu(), b.cpu()) for w, b in net.layers()]
    monkeypatch.setitem(globals(), "bf", lambda x: x.double())
    ref_out, ref_a, x0 = unet_ref_forward([torch.from_numpy(x) for x in imgs], L)
    e_out = rel_l2(out.cpu().numpy(), ref_out.numpy())

    # downstream: the same filter step fed the kernels' output and the unrounded one
    rest = [to_dev_np(x, dev) for x in (succ, sprev, prev_out, opac, pols1, pols2)]
    head4 = [to_dev_np(x, dev) for x in (ia, ib, alb, nrm)]
    got = pipeline.filter_train(*head4, out, *rest)
    ref = pipeline.filter_train(*head4, ref_out.float().to(dev), *rest)
    e_rad = rel_l2(got[1].cpu().numpy(), ref[1].cpu().numpy())
    e_dp = rel_l2(got[4].cpu().numpy(), ref[4].cpu().numpy())

    # weight gradients for the same upstream gradient (the unrounded chain's)
    dparam = ref[4].double().cpu()
    net.backward(ref[4].contiguous())
    torch.cuda.synchronize()
    gw = [(w.cpu().double(), b.cpu().double()) for w, b in cnn.unet_unflatten(net.grad)]
    g_ref, _ = unet_ref_backward(L, ref_a, x0, dparam)
    e_w = max(rel_l2(gw[l][0].numpy(), g_ref[l][0].numpy()) for l in range(12))
    e_b = max(rel_l2(gw[l][1].numpy(), g_ref[l][1].numpy()) for l in range(12))
    print(f"  unet {H}x{W} vs unrounded float64: output {e_out:.2e}, weight gradients <= {e_w:.2e}, "
          f"bias gradients <= {e_b:.2e}; downstream radiance {e_rad:.2e}, network-output gradient {e_dp:.2e}",
          flush=True)
    assert e_out <= 3e-2 and e_w <= 1e-1 and e_b <= 1e-1 and e_rad <= 1e-2 and e_dp <= 1e-1


def to_dev_np(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)
