"""MI355X-native backend for the UMANS per-step agent loop.

The compute path is libumans_gpu.so (hand-written HIP for gfx950 behind the C ABI in
include/umans_gpu.h); `umans_amd.gpu` binds it. Importing the package itself loads
nothing, so the registry and scenario generators work on machines without a GPU.
"""
__all__ = ["registry", "scenarios"]
