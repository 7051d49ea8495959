"""xdn_filter_train_noalloc, the fused training filter without a workspace argument, against
xdn_filter_train with a caller-owned workspace: identical outputs (the same kernels run), on two
calls in a row with the second on a smaller input (the library's fixed loss-partial array keeps the
first call's contents), and inputs beyond the fixed array's capacity rejected before any launch."""
import numpy as np
import pytest
import torch

from helpers import to_dev
from test_gpu_parity import train_case
from xdn._lib import lib


def _ptr(t):
    return None if t is None else t.data_ptr()


def _outputs(B, H, W, dev):
    e = lambda c: torch.full((B, H, W, c), float("nan"), device=dev)
    return e(3), e(3), e(6), e(3), torch.full((B,), float("nan"), device=dev)


def _run(noalloc, ins, B, H, W, dev):
    out_log, out, dparam, loss, avg = _outputs(B, H, W, dev)
    st = torch.cuda.current_stream(dev).cuda_stream
    args = [_ptr(t) for t in ins] + [_ptr(out_log), _ptr(out), _ptr(dparam), _ptr(loss), _ptr(avg)]
    if noalloc:
        rc = lib.xdn_filter_train_noalloc(*args, B, H, W, 11, st)
    else:
        ws_bytes = lib.xdn_filter_train_workspace_size(B, H, W)
        ws = torch.empty(max(ws_bytes, 4) // 4, device=dev)
        rc = lib.xdn_filter_train(*args, _ptr(ws), ws_bytes, B, H, W, 11, st)
    assert rc == 0, lib.xdn_last_error().decode()
    torch.cuda.synchronize()
    return out_log, out, dparam, loss, avg


def test_noalloc_rejects_inputs_beyond_its_scratch_without_gpu():
    cap = int(lib.xdn_filter_train_workspace_size(1, 16, 16)) * 262144
    assert cap == 1 << 20
    # 16 images of 2048x2048 need 16 x 128 x 128 = 262,144 tiles: the most that fits; one row more does not
    assert lib.xdn_filter_train_workspace_size(16, 2048, 2048) == cap
    rc = lib.xdn_filter_train_noalloc(*([None] * 16), 16, 2049, 2048, 11, None)
    assert rc == -1 and "fixed" in lib.xdn_last_error().decode()


@pytest.mark.gpu
def test_noalloc_matches_workspace_entry_point(dev):
    for (B, H, W, seed) in ((2, 96, 128, 5), (1, 40, 56, 9)):
        print(f"  filter_train_noalloc {B}x{H}x{W}", flush=True)
        ins = [to_dev(a, dev) for a in train_case(H, W, seed=seed, B=B)]
        want = _run(False, ins, B, H, W, dev)
        got = _run(True, ins, B, H, W, dev)
        for name, g, w in zip(("out_log", "out", "dparam", "loss", "avg_loss"), got, want):
            assert torch.isfinite(w).all(), name
            assert torch.equal(g, w), (name, B, H, W)
