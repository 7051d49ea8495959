"""Shared helpers of the test suite."""
import numpy as np
import torch


def rel_l2(got, ref) -> float:
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    den = np.linalg.norm(ref.ravel())
    num = np.linalg.norm((got - ref).ravel())
    return float(num / den) if den > 0 else float(num)


def to_np(t: torch.Tensor) -> np.ndarray:
    return t.detach().cpu().numpy()


def to_dev(a, dev) -> torch.Tensor:
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def rand_img(H, W, C=3, lo=0.0, hi=1.0, seed=0, B=1):
    r = np.random.default_rng(seed)
    return r.uniform(lo, hi, (B, H, W, C)).astype(np.float32)


def unit_normals(H, W, seed=0, B=1):
    r = np.random.default_rng(seed)
    n = r.normal(size=(B, H, W, 3))
    n[..., 2] = np.abs(n[..., 2]) + 1.0
    return (n / np.linalg.norm(n, axis=-1, keepdims=True)).astype(np.float32)


def log_images(H, W, seed=0, B=1):
    """Two noisy log-radiance half buffers with shared structure (like prepare_input output)."""
    r = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:H, 0:W]
    base = 1.0 + 0.8 * np.sin(xx / 7.0)[..., None] * np.cos(yy / 5.0)[..., None] * np.array([1.0, 0.7, 0.4])
    base = np.broadcast_to(base, (B, H, W, 3))
    c1 = base * np.exp(r.normal(0, 0.5, (B, H, W, 3)))
    c2 = base * np.exp(r.normal(0, 0.5, (B, H, W, 3)))
    return np.log1p(c1).astype(np.float32), np.log1p(c2).astype(np.float32)
