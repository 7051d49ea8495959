"""ctypes binding of libxdn.so (the C ABI declared in include/xdn.h).

The library is built in-tree by ``cross-denoiser_amd/csrc/Makefile`` (``__graft_entry__.build()``).
There is no fallback: if the library is missing or fails to load, importing this module raises.
"""
from __future__ import annotations

import ctypes
import os

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libxdn.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"libxdn.so not found at {LIB_PATH}; build it with `make -C cross-denoiser_amd/csrc` "
        "(or __graft_entry__.build())")

lib = ctypes.CDLL(LIB_PATH)

_P = ctypes.c_void_p
_I = ctypes.c_int
_F = ctypes.c_float
_Z = ctypes.c_size_t
_PP = ctypes.POINTER(ctypes.c_void_p)
_IP = ctypes.POINTER(ctypes.c_int)

# name -> (restype, argtypes); mirrors include/xdn.h
PROTOTYPES = {
    "xdn_version": (ctypes.c_char_p, []),
    "xdn_last_error": (ctypes.c_char_p, []),
    "xdn_cross_regression_workspace_size": (_Z, [_I, _I, _I, _I]),
    "xdn_cross_regression": (_I, [_P] * 8 + [_P, _Z] + [_I] * 6 + [_P]),
    "xdn_cross_regression_from_color": (_I, [_P] * 10 + [_P, _Z] + [_I] * 5 + [_P]),
    "xdn_cross_regression_grad_workspace_size": (_Z, [_I, _I, _I, _I]),
    "xdn_cross_regression_grad": (_I, [_P] * 12 + [_P, _Z] + [_I] * 5 + [_P]),
    "xdn_spatial_filter": (_I, [_P] * 9 + [_I] * 4 + [_P]),
    "xdn_spatial_filter_grad": (_I, [_P] * 10 + [_I] * 4 + [_P]),
    "xdn_spatial_filter_grad_inputs": (_I, [_P] * 13 + [_I] * 4 + [_P]),
    "xdn_reproject": (_I, [_P, _I, _PP, _IP] + [_P] * 8 + [_PP] + [_I] * 3 + [_P]),
    "xdn_temporal_screening": (_I, [_P] * 10 + [_I] * 5 + [_P]),
    "xdn_erode_temporal_screening": (_I, [_P] * 11 + [_I] * 5 + [_P]),
    "xdn_filter_train_workspace_size": (_Z, [_I, _I, _I]),
    "xdn_filter_train": (_I, [_P] * 16 + [_P, _Z] + [_I] * 4 + [_P]),
    "xdn_filter_train_noalloc": (_I, [_P] * 16 + [_I] * 4 + [_P]),
    "xdn_adam": (_I, [_P] * 4 + [_I, _I, _F, _F, _F, _F, _P]),
    "xdn_frame_error_workspace_size": (_Z, [_I, _I, _I]),
    "xdn_frame_error": (_I, [_P] * 8 + [_P, _Z] + [_I] * 3 + [_P]),
    "xdn_cnn5_param_count": (_Z, []),
    "xdn_cnn5_packed_bytes": (_Z, []),
    "xdn_cnn5_workspace_size": (_Z, [_I, _I, _I]),
    "xdn_cnn5_pack": (_I, [_P, _P, _P]),
    "xdn_cnn5_forward": (_I, [_P] * 7 + [_P, _Z] + [_I] * 3 + [_P]),
    "xdn_cnn5_backward": (_I, [_P] * 8 + [_P, _Z] + [_I] * 3 + [_P]),
    "xdn_unet_param_count": (_Z, []),
    "xdn_unet_packed_bytes": (_Z, []),
    "xdn_unet_workspace_size": (_Z, [_I, _I, _I]),
    "xdn_unet_workspace_layout": (_I, [_I, _I, _I, _I, _P, _P]),
    "xdn_unet_pack": (_I, [_P, _P, _P]),
    "xdn_unet_forward": (_I, [_P] * 7 + [_P, _Z] + [_I] * 3 + [_P]),
    "xdn_unet_backward": (_I, [_P] * 8 + [_P, _Z] + [_I] * 3 + [_P]),
    "xdn_slot_scratch_bytes": (_Z, [_I, _I]),
    "xdn_slot_scratch_offsets": (_I, [_I, _I, _P]),
    "xdn_slot_init": (_I, [_P, _P]),
    "xdn_denoise_step": (_I, [_P, _I, _F, _P]),
    "xdn_denoise_stages": (_I, [_P, _I, _F, _I, _P]),
    "xdn_denoise_graph_create": (_I, [_P, _I, _F, _P, _PP]),
    "xdn_denoise_graph_launch": (_I, [_P, _P]),
    "xdn_denoise_graph_destroy": (_I, [_P]),
}

for _name, (_res, _args) in PROTOTYPES.items():
    _fn = getattr(lib, _name)
    _fn.restype = _res
    _fn.argtypes = _args


class XdnError(RuntimeError):
    pass


def check(rc: int, what: str) -> None:
    """Raise for a non-zero return code: ValueError for bad arguments, RuntimeError for HIP."""
    if rc == 0:
        return
    msg = lib.xdn_last_error().decode()
    if rc == -1:
        raise ValueError(f"{what}: {msg}")
    raise XdnError(f"{what}: {msg}")


def ptr(t) -> int | None:
    return None if t is None else t.data_ptr()


def stream_of(t: torch.Tensor) -> int:
    return torch.cuda.current_stream(t.device).cuda_stream


def version() -> str:
    return lib.xdn_version().decode()
