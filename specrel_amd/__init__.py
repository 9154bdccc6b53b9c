"""SpecRel per-pixel hot path on AMD Instinct MI355X (gfx950).

The renderer itself is libspecrel_hip.so (hand-written HIP kernels behind the C ABI of
include/specrel.h); libspecrel_scene.so produces and packs scenes on the host. This package
holds the Python mirrors of the reference's renderer interface (src/gpu.rs `GpuRenderer`) and
scene types.
"""
from . import _ffi  # noqa: F401
from .renderer import Context, GpuRenderer, SpecrelError  # noqa: F401
from .scene import (Builder, Camera4D, FlashEvent, FrameDynamics, RenderConfig, Scene, Tetrad,  # noqa: F401
                    build_frame, default_camera, gen_params)
