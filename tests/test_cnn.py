"""Parity of the MFMA bandwidth CNN (cnn5) against a float64 torch-CPU reference.

The reference applies the bf16 roundings at the same points as the kernels (network input,
weights, each hidden activation, each backward activation gradient) and computes everything else
in float64 with library convolutions (torch.nn.functional.conv2d and torch.nn.grad); the kernels
accumulate in FP32.

Two bars:
  * per kernel, fed with the kernels' own bf16 operands read back from the workspace (so no
    rounding difference propagates): weight and bias gradients <= 1e-5 relative L2, data gradient
    equal to the bf16-rounded float64 result except for rare rounding-boundary cases;
  * end to end: network output <= 2e-3 against the float64 forward; the five weight and bias
    gradients <= 5e-3 against the float64 backward chain run on the kernels' own activations.
"""
import numpy as np
import pytest
import torch

from helpers import rel_l2

pytestmark = pytest.mark.gpu


def bf(x):
    return x.to(torch.bfloat16).to(torch.float64)


def ref_forward(imgs, layers):
    x = bf(torch.cat(imgs, -1).double()).permute(0, 3, 1, 2)
    acts = [x]
    out = None
    for l, (w, b) in enumerate(layers):
        wt = bf(w.double()).permute(3, 2, 0, 1)
        y = torch.nn.functional.conv2d(x, wt, padding=1) + b.double().view(1, -1, 1, 1)
        if l < 4:
            x = bf(torch.relu(y))
            acts.append(x)
        else:
            out = y
    return out.permute(0, 2, 3, 1), acts


def ref_backward(layers, acts, dparam):
    g = bf(dparam.double()).permute(0, 3, 1, 2)
    grads = [None] * 5
    for l in range(4, -1, -1):
        w, b = layers[l]
        wt = bf(w.double()).permute(3, 2, 0, 1)
        x = acts[l]
        dw = torch.nn.grad.conv2d_weight(x, wt.shape, g, padding=1)
        grads[l] = (dw.permute(2, 3, 1, 0), g.sum((0, 2, 3)))
        if l > 0:
            dx = torch.nn.grad.conv2d_input(x.shape, wt, g, padding=1)
            g = bf(dx * (acts[l] > 0))
    return grads


@pytest.mark.parametrize("B,H,W", [(1, 48, 80), (1, 37, 45), (2, 32, 64), (1, 128, 96), (1, 1080, 1920)])
def test_cnn5_forward_backward(dev, B, H, W):
    from xdn import cnn
    torch.manual_seed(H * W)
    imgs = [torch.rand(B, H, W, 3) * 2 for _ in range(5)]
    net = cnn.HipCNN5(dev, seed=7, shape=(B, H, W))
    # non-zero biases so the bias paths are exercised
    layers = net.layers()
    for w, b in layers:
        b.copy_(torch.randn(b.shape, device=dev) * 0.05)
    net.pack()
    g_imgs = [t.to(dev).contiguous() for t in imgs]
    out = net.forward(*g_imgs)
    torch.cuda.synchronize()
    layers_cpu = [(w.cpu(), b.cpu()) for w, b in net.layers()]
    ref_out, acts = ref_forward(imgs, layers_cpu)
    e = rel_l2(out.cpu().numpy(), ref_out.numpy())
    print(f"  cnn5 forward {B}x{H}x{W}: rel L2 {e:.2e}", flush=True)
    assert e <= 2e-3
    dparam = torch.randn(B, H, W, 6) * 1e-3
    net.backward(dparam.to(dev).contiguous())
    torch.cuda.synchronize()
    # the reference backward runs on the kernels' own bf16 activations: an activation that rounds
    # to zero on one side only would flip a ReLU mask, and a bias gradient (a signed sum over all
    # pixels) moves by ~1/sqrt(pixels) per flipped element
    acts_gpu = [acts[0]] + [_ws_act(net, i, B, H, W) for i in range(4)]
    ref = ref_backward(layers_cpu, acts_gpu, dparam)
    from xdn.cnn import unflatten_params
    got = unflatten_params(net.grad.cpu())
    errs = []
    for l in range(5):
        ew = rel_l2(got[l][0].numpy(), ref[l][0].numpy())
        eb = rel_l2(got[l][1].numpy(), ref[l][1].numpy())
        print(f"  cnn5 layer {l + 1} dW {ew:.2e} db {eb:.2e}", flush=True)
        errs.append((ew, eb))
    for l, (ew, eb) in enumerate(errs):
        assert ew <= 5e-3 and eb <= 5e-3, l
    check_kernels_on_own_operands(net, imgs, layers_cpu, B, H, W)


def _ws_act(net, i, B, H, W):
    """bf16 tensor i of the cnn5 workspace (0-3 hidden activations, 4-5 gradient ping-pong) as float64 NCHW."""
    act = (B * H * W * 32 * 2 + 255) // 256 * 256
    raw = net.ws[i * act:i * act + B * H * W * 32 * 2].view(torch.bfloat16).view(B, H, W, 32)
    return raw.cpu().to(torch.float64).permute(0, 3, 1, 2)


def check_kernels_on_own_operands(net, imgs, layers_cpu, B, H, W):
    """After backward: ping-pong slot 1 holds dY of layer 2 (its pre-activation gradient) and X[0]
    the layer-1 activation, the operands of the kernel that runs layer 2's backward together with
    layer 1's weight gradient. Layer 2's weight gradient is checked on exactly those operands. Layer
    1's pre-activation gradient (layer 2's data gradient where the activation is nonzero) stays in
    that kernel's LDS, so layer 1's weight gradient is checked against the one computed from its
    exact value rounded to bf16 once: the kernel's FP32 sums round a few elements to the
    neighbouring bf16 value (<= 1e-3 of them, one ulp = 2^-8 relative, random signs), which moves
    the weight and bias gradient by about sqrt(1e-3) * 2^-8 ~ 1e-4 relative at most; a wrong tap,
    channel, row or mask moves them by O(1)."""
    from xdn.cnn import unflatten_params
    got = unflatten_params(net.grad.cpu())
    x0 = bf(torch.cat(imgs, -1).double()).permute(0, 3, 1, 2)
    a1, dy2 = _ws_act(net, 0, B, H, W), _ws_act(net, 5, B, H, W)
    w1 = bf(layers_cpu[1][0].double()).permute(3, 2, 0, 1)
    dy1 = (torch.nn.grad.conv2d_input(a1.shape, w1, dy2, padding=1) * (a1 > 0)).to(torch.bfloat16).to(torch.float64)
    for l, (x, dy, tol) in enumerate(((x0, dy1, 1e-3), (a1, dy2, 1e-5))):
        w = layers_cpu[l][0]
        dw = torch.nn.grad.conv2d_weight(x, bf(w.double()).permute(3, 2, 0, 1).shape, dy, padding=1)
        ew = rel_l2(got[l][0].numpy(), dw.permute(2, 3, 1, 0).numpy())
        eb = rel_l2(got[l][1].numpy(), dy.sum((0, 2, 3)).numpy())
        print(f"  wgrad layer {l + 1} on its operands: dW {ew:.2e} db {eb:.2e}", flush=True)
        assert ew <= tol and eb <= tol, l


def test_cnn5_adam_and_pack_roundtrip(dev):
    """After an Adam step the packed fragments are rebuilt from the updated master weights."""
    from xdn import cnn
    B, H, W = 1, 32, 32
    net = cnn.HipCNN5(dev, seed=3, shape=(B, H, W))
    imgs = [torch.rand(B, H, W, 3, device=dev) for _ in range(5)]
    net.forward(*imgs)
    net.backward(torch.randn(B, H, W, 6, device=dev) * 1e-2)
    p0 = net.params.clone()
    net.adam_step(1e-3)
    torch.cuda.synchronize()
    assert not torch.equal(p0, net.params)
    # first Keras Adam step moves every parameter with a non-zero gradient by ~lr
    moved = (net.params - p0).abs()
    nz = net.grad.abs() > 1e-3  # eps 1e-7 against sqrt(v)=0.0316|g| stays below 0.3%
    assert torch.allclose(moved[nz], torch.full_like(moved[nz], 1e-3), rtol=1e-2, atol=1e-6)
    out = net.forward(*imgs)
    ref_out, _ = ref_forward([t.cpu() for t in imgs], [(w.cpu(), b.cpu()) for w, b in net.layers()])
    assert rel_l2(out.cpu().numpy(), ref_out.numpy()) <= 2e-3
