"""Each known answer of test_oracle_kat.py catches the mistake it is there for.

A scratch copy of oracle/ gets one planted mistake (a plausible slip: a flipped sign, swapped
operands, a dropped factor), is rebuilt with its own Makefile into a temporary directory and loaded
as a separate module; the known answers named for that mistake must fail on it. The mistakes are
planted consistently wherever the oracle repeats the term (forward and adjoint alike), which is the
case the finite-difference checks cannot see. CPU only (gcc).
"""
import importlib.util
import os
import shutil
import subprocess

import pytest

import test_oracle_kat as K

ORACLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")

# (name, file, [(old, new, expected count or None for all)], known answers that must fail)
MUTATIONS = [
    ("recon delta sign flipped (center minus pixel), forward and adjoint", "xdn_oracle.c", [
        ("{1.0, (cw[ch] - iw[ch]) / den, ca[ch] - ia[ch],\n"
         "                                              cn[0] - in[0], cn[1] - in[1], cn[2] - in[2]}",
         "{1.0, (iw[ch] - cw[ch]) / den, ia[ch] - ca[ch],\n"
         "                                              in[0] - cn[0], in[1] - cn[1], in[2] - cn[2]}", 1),
        ("d[ch][1] = (cw[ch] - iw[ch]) / den; d[ch][2] = ca[ch] - ia[ch];\n"
         "                            d[ch][3] = cn[0] - in[0]; d[ch][4] = cn[1] - in[1]; d[ch][5] = cn[2] - in[2];",
         "d[ch][1] = (iw[ch] - cw[ch]) / den; d[ch][2] = ia[ch] - ca[ch];\n"
         "                            d[ch][3] = in[0] - cn[0]; d[ch][4] = in[1] - cn[1]; d[ch][5] = in[2] - cn[2];", 1),
        ("crg_pair(&h, p, c, gw,", "crg_pair(&h, c, p, gw,", 1)],
     ["kat_recon_single_center"]),
    ("albedo and normal bandwidths swapped (b2 <-> b3)", "xdn_oracle.c", [
        ("- b[2] * dal - b[3] * dn", "- b[3] * dal - b[2] * dn", None)],
     ["kat_sf_albedo_band", "kat_sf_normal_band"]),
    ("colour bandwidths swapped (b0 <-> b1)", "xdn_oracle.c", [
        ("exp(-b[0] * dA", "exp(-B1_ * dA", None), ("exp(-b[1] * dB", "exp(-b[0] * dB", None),
        ("exp(-B1_ * dA", "exp(-b[1] * dA", None)],
     ["kat_sf_color_bands"]),
    ("pixel distance divided by the window instead of its half width", "xdn_oracle.c", [
        ("(double)(cy - iy) / half, dx = (double)(cx - ix) / half", "(double)(cy - iy) / win, dx = (double)(cx - ix) / win",
         None)],
     ["kat_sf_pixel_band"]),
    ("regression weight denominator epsilon 1e-3 instead of 1e-4", "xdn_oracle.c", [
        ("exp(-d / ((cvar + ivar) + 1e-4)) + 1e-7", "exp(-d / ((cvar + ivar) + 1e-3)) + 1e-7", 1)],
     ["kat_regression_checkerboard"]),
    ("regression weight without the +1e-7 offset", "xdn_oracle.c", [
        ("exp(-d / ((cvar + ivar) + 1e-4)) + 1e-7", "exp(-d / ((cvar + ivar) + 1e-4))", 1)],
     ["kat_regression_checkerboard"]),
    ("delta_1 normalised by the variances instead of the standard deviations", "xdn_oracle.c", [
        ("const double den = (cstd + istd) + 1e-2;\n                    for (int ch = 0; ch < 3; ++ch)\n"
         "                    {\n                        const double d[P6] = {1.0, (iw[ch] - cw[ch]) / den",
         "const double den = (cvar + ivar) + 1e-2;\n                    for (int ch = 0; ch < 3; ++ch)\n"
         "                    {\n                        const double d[P6] = {1.0, (iw[ch] - cw[ch]) / den", 1)],
     ["kat_regression_checkerboard"]),
    ("regression target read from the own half instead of the other half", "xdn_oracle.c", [
        ("load3(nrm, i, in); load3(target, i, it);", "load3(nrm, i, in); load3(feat, i, it);", 1)],
     ["kat_regression_checkerboard"]),
    ("ridge 1e-2 instead of 1e-3", "xdn_oracle.c", [
        ("M[i][i] += 1e-3;", "M[i][i] += 1e-2;", 1)],
     ["kat_regression_checkerboard"]),
    ("final 1/2 of the aggregation loss dropped", "oracle.py", [
        ("loss = ((loss1 + loss2) * 0.5 + (t1 + t2) * 0.5) * 0.5 * op",
         "loss = ((loss1 + loss2) * 0.5 + (t1 + t2) * 0.5) * op", 1)],
     ["kat_aggregation_loss"]),
    ("temporal blend without the success mask", "oracle.py", [
        ("a = s * alpha + (1 - s)", "a = alpha", None)],
     ["kat_aggregation_loss"]),
    ("labels not crossed (out1 scored against ols1)", "oracle.py", [
        ("ols2, ols1, prev_ols2, prev_ols1)", "ols1, ols2, prev_ols1, prev_ols2)", 1)],
     ["kat_filter_train_labels"]),
]


def _mutated_oracle(tmp_path, file, edits):
    d = tmp_path / "oracle"
    shutil.copytree(ORACLE, d, ignore=shutil.ignore_patterns("liboracle.so", "__pycache__", "_ref"))
    path = d / file
    src = path.read_text()
    for old, new, count in edits:
        n = src.count(old)
        assert n > 0 and (count is None or n == count), f"edit site {old[:40]!r} found {n} times"
        src = src.replace(old, new)
    path.write_text(src)
    subprocess.check_call(["make", "-s", "-C", str(d)])
    spec = importlib.util.spec_from_file_location(f"oracle_mutant_{tmp_path.name}", str(d / "oracle.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    assert mod._LIB_PATH == str(d / "liboracle.so")
    return mod


@pytest.mark.skipif(shutil.which("gcc") is None and shutil.which("cc") is None, reason="needs a C compiler")
@pytest.mark.parametrize("name,file,edits,kats", MUTATIONS, ids=[m[0] for m in MUTATIONS])
def test_known_answers_catch_planted_mistake(tmp_path, name, file, edits, kats):
    mutant = _mutated_oracle(tmp_path, file, edits)
    for k in kats:
        with pytest.raises(AssertionError):
            getattr(K, k)(mutant)
