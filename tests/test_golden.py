"""Golden vectors (tests/golden/*.npz, written by tests/golden/make_fixtures.py).

CPU: the FP64 oracle still reproduces every stored vector. GPU: the device ops match them, masks
bit-exact and values within the tolerances of the parity tests.
"""
import os

import numpy as np
import pytest

from helpers import rel_l2
from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TOL = 1e-4


def load(name):
    with np.load(os.path.join(GOLD, name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def close(a, b):
    np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-13)


def test_oracle_cross_regression_golden():
    g = load("cross_regression_16x20")
    out = O.cross_regression(g["imga"], g["imgb"], g["albedo"], g["normal"], 17, 4, 1)
    for k, o in zip(("outa", "outb", "vara", "varb"), out):
        close(o, g[k])
    gr = O.cross_regression_grad(g["imga"], g["imgb"], g["albedo"], g["normal"], g["grad_outa"], g["grad_outb"],
                                 g["grad_vara"], g["grad_varb"])
    for k, o in zip(("grad_imga", "grad_imgb", "grad_albedo", "grad_normal"), gr):
        close(o, g[k])


def test_oracle_spatial_filter_golden():
    g = load("spatial_filter_13x17")
    ins = [g[k] for k in ("imga", "imgb", "albedo", "normal", "band")]
    cots = [g[k] for k in ("grad_acca", "grad_accb", "grad_wgtsuma", "grad_wgtsumb")]
    for k, o in zip(("acca", "accb", "wgtsuma", "wgtsumb"), O.spatial_filter(*ins, win_size=11)):
        close(o, g[k])
    close(O.spatial_filter_grad(*cots, *ins, win_size=11), g["grad_band"])
    for k, o in zip(("grad_imga", "grad_imgb", "grad_albedo", "grad_normal"),
                    O.spatial_filter_grad_inputs(*cots, *ins, win_size=11)):
        close(o, g[k])


def _rp_args(g):
    return (g["mvec"], [g["in0"], g["in1"]], g["linear_z"], g["prev_linear_z"], g["normal"], g["prev_normal"],
            g["pn_fwidth"], g["opacity"], g["prev_opacity"])


def test_oracle_reproject_golden():
    g = load("reproject_18x22")
    s, outs = O.reproject(*_rp_args(g))
    assert np.array_equal(s.astype(np.uint8), g["success"])
    assert 0 < g["success"].sum() < g["success"].size
    close(outs[0], g["out0"])
    close(outs[1], g["out1"])


def test_oracle_screening_golden():
    g = load("temporal_screening_15x19")
    se = O.erosion(g["success"].astype(bool))
    assert np.array_equal(se, g["success_eroded"])
    outs = O.temporal_screening(g["current1"], g["current2"], g["prev"], g["prev1"], g["prev2"], se, 7, 1)
    for k, o in zip(("output", "output1", "output2", "success_prev"), outs):
        assert np.array_equal(o, g[k])


# ------------------------------------------------------------------------------------------------
@pytest.mark.gpu
def test_device_ops_match_golden(dev):
    import torch
    from xdn import ops
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    n = lambda x: x.detach().cpu().numpy()
    g = load("cross_regression_16x20")
    out = ops.cross_regression(t(g["imga"]), t(g["imgb"]), t(g["albedo"]), t(g["normal"]), win_size=17, sparsity=4)
    for k, o in zip(("outa", "outb", "vara", "varb"), out):
        assert rel_l2(n(o), g[k]) <= TOL, k
    gr = ops.cross_regression_grad(*[t(g[k]) for k in ("imga", "imgb", "albedo", "normal", "grad_outa", "grad_outb",
                                                       "grad_vara", "grad_varb")], win_size=17, sparsity=4)
    for k, o in zip(("grad_imga", "grad_imgb", "grad_albedo", "grad_normal"), gr):
        assert rel_l2(n(o), g[k]) <= TOL, k

    g = load("spatial_filter_13x17")
    ins = [t(g[k]) for k in ("imga", "imgb", "albedo", "normal", "band")]
    cots = [t(g[k]) for k in ("grad_acca", "grad_accb", "grad_wgtsuma", "grad_wgtsumb")]
    for k, o in zip(("acca", "accb", "wgtsuma", "wgtsumb"), ops.spatial_filter(*ins, win_size=11)):
        assert rel_l2(n(o), g[k]) <= TOL, k
    assert rel_l2(n(ops.spatial_filter_grad(*cots, *ins, win_size=11)), g["grad_band"]) <= TOL
    for k, o in zip(("grad_imga", "grad_imgb", "grad_albedo", "grad_normal"),
                    ops.spatial_filter_grad_inputs(*cots, *ins, win_size=11)):
        assert rel_l2(n(o), g[k]) <= TOL, k

    g = load("reproject_18x22")
    a = _rp_args(g)
    s, outs = ops.Reproject(t(a[0]), [t(x) for x in a[1]], *[t(x) for x in a[2:]])
    assert np.array_equal(n(s).astype(np.uint8), g["success"])
    assert rel_l2(n(outs[0]), g["out0"]) <= 1e-6 and rel_l2(n(outs[1]), g["out1"]) <= 1e-6

    g = load("temporal_screening_15x19")
    outs = ops.temporal_screening(*[t(g[k]) for k in ("current1", "current2", "prev", "prev1", "prev2",
                                                      "success_eroded")], win_curr=7, win_prev=1)
    for k, o in zip(("output", "output1", "output2", "success_prev"), outs):
        assert np.array_equal(n(o), g[k]), k
