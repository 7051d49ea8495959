"""Writes the golden vectors of tests/golden/*.npz from the FP64 oracle.

The reference holds no fixtures or known-answer files for this path (it has no tests), so these
vectors are produced by the oracle, whose functions are pinned by closed forms, brute force and
central differences in tests/test_oracle.py. They freeze the oracle's outputs at small, awkward
shapes (sizes not multiples of the tile sizes, borders, mirrored windows) so that both the
oracle and the device kernels are checked against fixed data. Inputs are float32, outputs are
stored as float64 (masks as uint8). Run from the repository root:

    python tests/golden/make_fixtures.py
"""
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "cross-denoiser_amd"))
sys.path.insert(0, os.path.dirname(HERE))

from oracle import oracle as O  # noqa: E402
from helpers import log_images, rand_img, unit_normals  # noqa: E402


def _save(name, seed, **arrays):
    try:
        rev = subprocess.run(["git", "-C", ROOT, "rev-parse", "--short", "HEAD"], capture_output=True,
                             text=True).stdout.strip()
    except OSError:
        rev = "unknown"
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, seed=np.int64(seed), oracle_rev=np.array(rev), **arrays)
    print(f"{path}: {os.path.getsize(path)} bytes")


def cross_regression(seed=101):
    H, W = 16, 20
    ya, yb = log_images(H, W, seed=seed)
    alb, nrm = rand_img(H, W, seed=seed + 1), unit_normals(H, W, seed=seed + 2)
    oa, ob, va, vb = O.cross_regression(ya, yb, alb, nrm, 17, 4, 1)
    r = np.random.default_rng(seed + 3)
    cots = [r.normal(size=(1, H, W, c)).astype(np.float32) for c in (3, 3, 1, 1)]
    g = O.cross_regression_grad(ya, yb, alb, nrm, *cots, win_size=17, sparsity=4)
    _save("cross_regression_16x20", seed, imga=ya, imgb=yb, albedo=alb, normal=nrm, outa=oa, outb=ob, vara=va,
          varb=vb, grad_outa=cots[0], grad_outb=cots[1], grad_vara=cots[2], grad_varb=cots[3], grad_imga=g[0],
          grad_imgb=g[1], grad_albedo=g[2], grad_normal=g[3])


def spatial_filter(seed=202):
    H, W = 13, 17
    ia, ib = log_images(H, W, seed=seed)
    alb, nrm = rand_img(H, W, seed=seed + 1), unit_normals(H, W, seed=seed + 2)
    r = np.random.default_rng(seed + 3)
    band = np.exp(r.uniform(np.log(0.1), np.log(10.0), (1, H, W, 5))).astype(np.float32)
    acc = O.spatial_filter(ia, ib, alb, nrm, band, win_size=11)
    cots = [r.normal(size=(1, H, W, c)).astype(np.float32) for c in (3, 3, 1, 1)]
    gband = O.spatial_filter_grad(*cots, ia, ib, alb, nrm, band, win_size=11)
    gin = O.spatial_filter_grad_inputs(*cots, ia, ib, alb, nrm, band, win_size=11)
    _save("spatial_filter_13x17", seed, imga=ia, imgb=ib, albedo=alb, normal=nrm, band=band, acca=acc[0],
          accb=acc[1], wgtsuma=acc[2], wgtsumb=acc[3], grad_acca=cots[0], grad_accb=cots[1], grad_wgtsuma=cots[2],
          grad_wgtsumb=cots[3], grad_band=gband, grad_imga=gin[0], grad_imgb=gin[1], grad_albedo=gin[2],
          grad_normal=gin[3])


def reproject(seed=303):
    from xdn import synth
    H, W = 18, 22
    spec = synth.StreamSpec(H, W, seed=seed, pan=(1.7, -1.2), disocclusion=0.15)
    f0, f1 = synth.make_frame(spec, 0), synth.make_frame(spec, 1)
    n = lambda t: t.numpy()
    r = np.random.default_rng(seed)
    ins = [r.uniform(0, 2, (1, H, W, 3)).astype(np.float32), r.uniform(0, 2, (1, H, W, 1)).astype(np.float32)]
    opac = np.ones((1, H, W, 1), np.float32)
    popac = np.ones((1, H, W, 1), np.float32)
    popac[:, 5, :] = 0.0
    args = (n(f1["mvec"]), ins, n(f1["linearZ"]), n(f0["linearZ"]), n(f1["normal"]), n(f0["normal"]),
            n(f1["pnFwidth"]), opac, popac)
    s, outs = O.reproject(*args)
    _save("reproject_18x22", seed, mvec=args[0], in0=ins[0], in1=ins[1], linear_z=args[2], prev_linear_z=args[3],
          normal=args[4], prev_normal=args[5], pn_fwidth=args[6], opacity=opac, prev_opacity=popac,
          success=s.astype(np.uint8), out0=outs[0], out1=outs[1])


def temporal_screening(seed=404):
    H, W = 15, 19
    c1, c2 = log_images(H, W, seed=seed)
    r = np.random.default_rng(seed)
    p = (0.5 * (c1 + c2) + r.normal(0, 0.05, c1.shape)).astype(np.float32)
    p1 = (c1 + r.normal(0, 0.3, c1.shape)).astype(np.float32)
    p2 = (c2 + r.normal(0, 0.3, c1.shape)).astype(np.float32)
    succ = (r.random((1, H, W, 1)) > 0.1)
    se = O.erosion(succ)
    outs = O.temporal_screening(c1, c2, p, p1, p2, se, 7, 1)
    _save("temporal_screening_15x19", seed, current1=c1, current2=c2, prev=p, prev1=p1, prev2=p2,
          success=succ.astype(np.uint8), success_eroded=se, output=outs[0], output1=outs[1], output2=outs[2],
          success_prev=outs[3])


if __name__ == "__main__":
    cross_regression()
    spatial_filter()
    reproject()
    temporal_screening()
