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


def test_empty_batch_is_a_no_op_without_gpu():
    """Batch 0 returns 0 with every buffer NULL (zero-element tensors have null data pointers) and
    touches no device; attributes are still checked."""
    from xdn._lib import lib
    H, W = 32, 48
    N = None
    calls = {
        "xdn_cross_regression": [N] * 8 + [N, 0, 0, H, W, 17, 4, 1, N],
        "xdn_cross_regression_from_color": [N] * 10 + [N, 0, 0, H, W, 17, 4, N],
        "xdn_cross_regression_grad": [N] * 12 + [N, 0, 0, H, W, 17, 4, N],
        "xdn_spatial_filter": [N] * 9 + [0, H, W, 11, N],
        "xdn_spatial_filter_grad": [N] * 10 + [0, H, W, 11, N],
        "xdn_spatial_filter_grad_inputs": [N] * 13 + [0, H, W, 11, N],
        "xdn_reproject": [N, 3, N, (ctypes.c_int * 3)(3, 3, 1)] + [N] * 8 + [N, 0, H, W, N],
        "xdn_temporal_screening": [N] * 10 + [0, H, W, 7, 1, N],
        "xdn_erode_temporal_screening": [N] * 11 + [0, H, W, 7, 1, N],
        "xdn_filter_train": [N] * 16 + [N, 0, 0, H, W, 11, N],
        "xdn_filter_train_noalloc": [N] * 16 + [0, H, W, 11, N],
        "xdn_frame_error": [N] * 8 + [N, 0, 0, H, W, N],
        "xdn_cnn5_forward": [N] * 7 + [N, 0, 0, H, W, N],
        "xdn_unet_forward": [N] * 7 + [N, 0, 0, H, W, N],
    }
    for name, args in calls.items():
        rc = getattr(lib, name)(*args)
        assert rc == 0, (name, lib.xdn_last_error().decode())
    # attributes are checked for an empty batch too
    assert lib.xdn_cross_regression(*([N] * 8 + [N, 0, 0, H, W, 17, 2, 4, N])) == -1
    assert lib.xdn_filter_train(*([N] * 16 + [N, 0, 0, H, W, 7, N])) == -1
    assert lib.xdn_unet_forward(*([N] * 7 + [N, 0, 0, 30, W, N])) == -1
    assert lib.xdn_cross_regression(*([N] * 8 + [N, 0, -1, H, W, 17, 4, 1, N])) == -1
    # the weight gradients have no batch dimension: an empty batch still needs grad (zeroed)
    assert lib.xdn_cnn5_backward(*([N] * 8 + [N, 0, 0, H, W, N])) == -1
    assert lib.xdn_unet_backward(*([N] * 8 + [N, 0, 0, H, W, N])) == -1


def test_oversize_frames_are_rejected_without_gpu():
    """Frames past the kernels' 32-bit offsets are refused before anything is launched: cnn5 below
    2^25 pixels (its 64 B/px activations), the filter below 2^32 pixels."""
    from xdn import pipeline
    from xdn._lib import lib
    N = None
    assert lib.xdn_cnn5_forward(*([N] * 7 + [N, 1 << 40, 1, 8192, 4096, N])) == -1
    assert "2^25" in lib.xdn_last_error().decode()
    assert lib.xdn_spatial_filter(*([N] * 9 + [1, 65536, 65536, 11, N])) == -1
    assert "2^32" in lib.xdn_last_error().decode()
    s = pipeline.XdnSlot()
    s.height, s.width = 8192, 4096
    assert lib.xdn_denoise_step(ctypes.addressof(s), 1, 1e-3, None) == -1
    assert "2^25" in lib.xdn_last_error().decode()
    s.height, s.width = 4320, 7680  # 8K UHD fits; the next failure is the missing buffers
    assert lib.xdn_denoise_step(ctypes.addressof(s), 1, 1e-3, None) == -1
    assert "null" in lib.xdn_last_error().decode()


def test_make_slot_parity_selects_history_buffers():
    """A slot for step t takes its outputs from history buffer t % 2 and its history from the other
    (host logic only: CPU tensors stand in for device buffers)."""
    import torch
    from xdn import pipeline

    class Fake(pipeline.Denoiser):
        def __init__(self):  # no device allocation: only the fields make_slot reads
            self.H, self.W, self.t = 8, 12, 3
            self.device = torch.device("cpu")
            z = lambda: torch.zeros(1, 8, 12, 3)
            self.hist = [[z(), z(), z()], [z(), z(), z()]]
            self.out, self.zeros3, self.scratch = z(), z(), torch.zeros(16, dtype=torch.uint8)
            self.params = self.grad = self.adam_m = self.adam_v = torch.zeros(4)
            self.packed, self.avg_loss = torch.zeros(4, dtype=torch.uint8), torch.zeros(1)
            self.adam_t, self.prev_frame = torch.zeros(1, dtype=torch.int32), None
            self.network = "unet"

    d = Fake()
    frame = {k: torch.zeros(1, 8, 12, 3 if k != "opacity" else 1) for k in
             ("color", "color2", "albedo", "normal", "mvec", "linearZ", "pnFwidth", "opacity")}
    s3, s4 = d.make_slot(frame), d.make_slot(frame, t=4)
    assert s3.ols1 == d.hist[1][1].data_ptr() and s3.prev_ols1 == d.hist[0][1].data_ptr()
    assert s4.ols1 == d.hist[0][1].data_ptr() and s4.prev_ols1 == d.hist[1][1].data_ptr()
    assert s3.adam_step == 4 and s4.adam_step == 5 and s3.color == frame["color"].data_ptr()
    assert s3.network == pipeline.NETWORKS["unet"] == 1


def test_slot_network_checks_without_gpu():
    """Slot argument checks (network kind, one network per call) run before any launch."""
    from xdn import pipeline
    from xdn._lib import lib
    a, b = pipeline.XdnSlot(), pipeline.XdnSlot()
    for s in (a, b):
        s.height, s.width = 64, 64
    a.network = 5
    assert lib.xdn_denoise_step(ctypes.addressof(a), 1, 1e-3, None) == -1
    assert "unknown network" in lib.xdn_last_error().decode()
    pair = (pipeline.XdnSlot * 2)()
    for i, net in enumerate((0, 1)):
        pair[i].height, pair[i].width, pair[i].network = 64, 64, net
    assert lib.xdn_denoise_step(ctypes.addressof(pair), 2, 1e-3, None) == -1
    assert "one network" in lib.xdn_last_error().decode()
    assert pipeline.slot_bytes(1080, 1920, "unet") > 0 and pipeline.slot_bytes(1080, 1920, "cnn5") > 0


def test_null_buffers_are_rejected_without_gpu():
    """Entry points reject null buffers before anything is launched."""
    from xdn import pipeline
    from xdn._lib import lib
    ws = int(lib.xdn_unet_workspace_size(1, 32, 32))
    assert lib.xdn_unet_forward(None, None, None, None, None, None, None, None, ws, 1, 32, 32, None) == -1
    assert "null" in lib.xdn_last_error().decode()
    ws5 = int(lib.xdn_cnn5_workspace_size(1, 32, 32))
    assert lib.xdn_cnn5_backward(*([None] * 9), ws5, 1, 32, 32, None) == -1
    assert "null" in lib.xdn_last_error().decode()
    s = pipeline.XdnSlot()
    s.height, s.width = 64, 64
    assert lib.xdn_denoise_step(ctypes.addressof(s), 1, 1e-3, None) == -1
    assert "is null" in lib.xdn_last_error().decode()
    assert lib.xdn_slot_init(ctypes.addressof(s), None) == -1


def test_single_op_null_buffers_are_rejected_without_gpu():
    """The single-op entry points reject a missing required tensor before anything is launched; the
    documented optional ones (variance gradients, unrequested input gradients, the per-pixel loss) stay
    optional."""
    from xdn._lib import lib

    def rejected(rc):
        return rc == -1 and "null" in lib.xdn_last_error().decode()

    ws = 1 << 20
    assert rejected(lib.xdn_cross_regression(*([None] * 8), ws, ws, 1, 32, 32, 7, 4, 1, None))
    assert rejected(lib.xdn_spatial_filter(*([None] * 9), 1, 32, 32, 11, None))
    assert rejected(lib.xdn_spatial_filter_grad(*([None] * 10), 1, 32, 32, 11, None))
    assert rejected(lib.xdn_spatial_filter_grad_inputs(*([None] * 13), 1, 32, 32, 11, None))
    assert rejected(lib.xdn_reproject(None, 0, None, None, *([None] * 8), None, 1, 32, 32, None))
    assert rejected(lib.xdn_temporal_screening(*([None] * 5), 1, *([None] * 4), 1, 32, 32, 7, 7, None))
    assert rejected(lib.xdn_filter_train(*([None] * 16), ws, ws, 1, 32, 32, 11, None))
    assert rejected(lib.xdn_adam(None, None, None, None, 16, 1, 1e-3, 0.9, 0.999, 1e-7, None))
    assert rejected(lib.xdn_cnn5_pack(None, None, None))
    # nothing to update: no buffers needed
    assert lib.xdn_adam(None, None, None, None, 0, 1, 1e-3, 0.9, 0.999, 1e-7, None) in (0, -1)
