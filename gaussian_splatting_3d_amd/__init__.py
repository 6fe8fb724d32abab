"""MI355X (gfx950) 3D Gaussian-splatting rasteriser.

Drop-in for the rasteriser path of heheyas/gaussian_splatting_3d:
``renderer`` (render_sh, render_sh_bg, render_start_end, render,
project_gaussians), ``backend._backend`` (the reference extension's function
names), ``culling.tile_culling_aabb_count``, ``camera.CameraInfo`` and
``sh_renderer.SHRenderer``; ``rasterize`` is the fused gfx950 pipeline behind
SHRenderer.  The kernels live in libgsmi.so (C ABI: include/gsmi.h).
"""
from .camera import CameraInfo, get_c2w_from_up_and_look_at  # noqa: F401
from .scenes import RasterConfig  # noqa: F401

__all__ = ["CameraInfo", "get_c2w_from_up_and_look_at", "RasterConfig"]
