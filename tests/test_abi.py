"""The C ABI library loads without a GPU and exports every function include/xdn.h declares;
argument errors are reported before anything is launched (no GPU needed for those paths)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "xdn.h")
LIB = os.path.join(ROOT, "cross-denoiser_amd", "xdn", "libxdn.so")


def declared():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(xdn_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_entry_points():
    names = declared()
    for must in ("xdn_cross_regression", "xdn_spatial_filter", "xdn_spatial_filter_grad", "xdn_reproject",
                 "xdn_temporal_screening", "xdn_filter_train", "xdn_adam"):
        assert must in names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(LIB):
        pytest.fail("libxdn.so is not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_prototypes_cover_header():
    from xdn import _lib
    assert sorted(_lib.PROTOTYPES) == declared()


def test_argument_errors_without_gpu():
    from xdn._lib import lib
    lib.xdn_cross_regression.restype = ctypes.c_int
    rc = lib.xdn_cross_regression(None, None, None, None, None, None, None, None, None, 0, 1, 32, 32, 17, 2, 4, None)
    assert rc == -1
    assert lib.xdn_last_error().decode() == "recon_scale must be less than or equal to sparsity"
    rc = lib.xdn_cross_regression(None, None, None, None, None, None, None, None, None, 0, 1, 30, 32, 17, 4, 4, None)
    assert lib.xdn_last_error().decode() == "Image height and width must be divisible by recon_scale"
    assert lib.xdn_cross_regression_workspace_size(1, 1080, 1920, 4) == 270 * 480 * 36 * 4
    rc = lib.xdn_reproject(None, 1, None, (ctypes.c_int * 1)(5), *([None] * 8), None, 1, 8, 8, None)
    assert rc == -1 and "channels" in lib.xdn_last_error().decode()
