"""xdn: MI355X-native per-frame filtering path of cross-denoiser.

The hand-written gfx950 kernels live in libxdn.so (C ABI: include/xdn.h); ``xdn.ops`` exposes them
behind the reference's op names (scripts/ops.py), ``xdn.pipeline`` runs the per-frame online
training step (scripts/main.py:250-313), ``xdn.driver`` the reference's frame loop around it
(main.py:318-359: displayed frame and RelL2 report) and ``xdn.serving`` the multi-stream serving
loop.
Importing ``xdn.ops`` loads the library and fails if it is missing.
"""
__all__ = ["ops", "synth", "pipeline", "cnn", "driver", "serving"]
