"""Parity of the MFMA bandwidth CNN (cnn5) against a float64 torch-CPU reference.

The reference applies the bf16 roundings at the same points as the kernels (network input,
weights, each hidden activation, each backward activation gradient) and computes everything else
in float64 with library convolutions (torch.nn.functional.conv2d and torch.nn.grad); the kernels
accumulate in FP32. Tolerances (relative L2): network output <= 2e-3, weight/bias gradients
<= 5e-3 (FP32 accumulation over all pixels and rare bf16 rounding flips of near-tie values).
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


@pytest.mark.parametrize("B,H,W", [(1, 48, 80), (1, 37, 45), (2, 32, 64), (1, 128, 96)])
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
    ref = ref_backward(layers_cpu, acts, dparam)
    from xdn.cnn import unflatten_params
    got = unflatten_params(net.grad.cpu())
    for l in range(5):
        ew = rel_l2(got[l][0].numpy(), ref[l][0].numpy())
        eb = rel_l2(got[l][1].numpy(), ref[l][1].numpy())
        print(f"  cnn5 layer {l + 1} dW {ew:.2e} db {eb:.2e}", flush=True)
        assert ew <= 5e-3 and eb <= 5e-3, l


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
