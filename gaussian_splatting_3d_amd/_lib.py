"""Loader for the gfx950 C-ABI library ``libgsmi.so`` (declared in include/gsmi.h).

torch is imported first so that the HIP runtime torch ships is the one the
library binds to (both carry the soname ``libamdhip64.so.7``); device
pointers and streams are then shared between torch and the kernels.

There is no fallback: if the library is missing the import of any GPU entry
point raises, so a run can never silently use something other than the HIP
kernels.
"""
from __future__ import annotations

import ctypes
import os

import torch  # noqa: F401  (must precede the library load, see module docstring)

_HERE = os.path.dirname(os.path.abspath(__file__))
# GSMI_LIB: an alternative build of the same library (e.g. a tuning variant).
LIB_PATH = os.environ.get("GSMI_LIB") or os.path.join(_HERE, "libgsmi.so")

_P = ctypes.c_void_p
_I32 = ctypes.c_int
_I64 = ctypes.c_int64
_F = ctypes.c_float
_D = ctypes.c_double
_SZ = ctypes.c_size_t

# name -> (restype, argtypes); mirrors include/gsmi.h
SIGNATURES = {
    "gsmi_status_string": (ctypes.c_char_p, [_I32]),
    "gsmi_version": (_I32, []),
    "gsmi_culling_gaussian_bsphere": (_I32, [_I64, _P, _P, _P, _P, _P, _P, _F, _P]),
    "gsmi_aabb_count_workspace_size": (_SZ, [_I64]),
    "gsmi_tile_culling_aabb_count": (_I32, [_I64, _P, _P, _F, _F, _F, _F, _I32, _I32, _F, _I32, _P, _P, _P, _P,
                                            _SZ, _P]),
    "gsmi_binning_workspace_size": (_SZ, [_I64, _I64, _I32]),
    "gsmi_tile_culling_aabb_start_end": (_I32, [_I64, _I64, _P, _P, _P, _I32, _I32, _P, _P, _P, _P, _P, _SZ, _I32,
                                                _P]),
    "gsmi_sort_workspace_size": (_SZ, [_I64]),
    "gsmi_sort_pairs": (_I32, [_I64, _P, _P, _I32, _P, _SZ, _P]),
    "gsmi_preprocess_workspace_size": (_SZ, [_I64]),
    "gsmi_preprocess_count": (_I32, [_I64, _P, _P, _P, _P, _P, _P, _F, _F, _F, _F, _I32, _I32, _F, _F, _I32, _I32,
                                     _P, _P, _P, _P, _P, _SZ, _P]),
    "gsmi_preprocess_emit": (_I32, [_I64, _P, _P, _P, _P, _P, _F, _F, _F, _F, _I32, _I32, _F, _I32, _P, _P, _P,
                                    _P, _P, _P, _P, _P, _P, _P]),
    "gsmi_tile_ranges": (_I32, [_I64, _P, _I32, _P, _P, _P]),
    # the reference's own compositing signatures (gs/src/render.h:70-133)
    "gsmi_tile_based_vol_rendering_sh": (_I32, [_I64, _I64, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _I32, _I32, _I32,
                                                _F, _F, _I32, _I32, _I32, _F, _P]),
    "gsmi_tile_based_vol_rendering_backward_sh": (_I32, [_I64, _I64, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P,
                                                         _P, _P, _P, _I32, _I32, _I32, _F, _F, _I32, _I32, _I32, _F,
                                                         _P]),
    "gsmi_tile_based_vol_rendering_sh_with_bg": (_I32, [_I64, _I64, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _I32,
                                                        _I32, _I32, _F, _F, _I32, _I32, _I32, _F, _P, _P]),
    "gsmi_tile_based_vol_rendering_backward_sh_with_bg": (_I32, [_I64, _I64, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P,
                                                                 _P, _P, _P, _P, _P, _I32, _I32, _I32, _F, _F, _I32,
                                                                 _I32, _I32, _F, _P, _P]),
    "gsmi_tile_based_vol_rendering_start_end": (_I32, [_I64, _I64, _P, _P, _P, _P, _P, _P, _P, _P, _P, _I32, _I32,
                                                       _I32, _F, _F, _I32, _I32, _F, _P]),
    "gsmi_tile_based_vol_rendering_backward_start_end": (_I32, [_I64, _I64, _P, _P, _P, _P, _P, _P, _P, _P, _P,
                                                                _P, _P, _P, _P, _P, _I32, _I32, _I32, _F, _F, _I32,
                                                                _I32, _F, _P]),
    # extended compositing entries
    "gsmi_render_sh_forward": (_I32, [_I64, _I64, _P, _P, _P, _I64, _I64, _P, _P, _P, _P, _P, _P, _P, _P, _I32,
                                      _I32, _I32, _F, _F, _I32, _I32, _I32, _F, _P, _P, _P, _P, _I32, _I32, _P]),
    "gsmi_backward_order_workspace_size": (_SZ, [_I32]),
    "gsmi_render_sh_backward": (_I32, [_I64, _I64, _P, _P, _P, _I64, _I64, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P,
                                       _P, _P, _P, _P, _P, _P, _I32, _I32, _I32, _F, _F, _I32, _I32, _I32, _F, _P,
                                       _I32, _P, _SZ, _P]),
    "gsmi_render_sh_backward_noalloc": (_I32, [_I64, _I64, _P, _P, _P, _I64, _I64, _P, _P, _P, _P, _P, _P, _P, _P,
                                               _P, _P, _P, _P, _P, _P, _P, _P, _I32, _I32, _I32, _F, _F, _I32, _I32,
                                               _I32, _F, _P, _I32, _P]),
    "gsmi_render_rgb_forward": (_I32, [_I64, _I64, _P, _P, _P, _P, _P, _P, _P, _P, _P, _I32, _I32, _I32, _F, _F,
                                       _I32, _I32, _F, _P, _P, _I32, _P]),
    "gsmi_render_rgb_backward": (_I32, [_I64, _I64, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P,
                                        _P, _I32, _I32, _I32, _F, _F, _I32, _I32, _F, _P]),
    "gsmi_shift_rects": (_I32, [_I64, _P, _P, _I32, _P, _P, _P]),
    "gsmi_two_level_offsets_workspace_size": (_SZ, [_I32]),
    "gsmi_two_level_offsets": (_I32, [_I32, _I32, _I32, _P, _P, _P, _P, _SZ, _P]),
    "gsmi_render_sh_two_level": (_I32, [_I64, _P, _P, _P, _I64, _I64, _P, _P, _P, _P, _P, _P, _P, _I32, _P, _P, _P,
                                        _P, _P, _P, _I32, _I32, _F, _F, _I32, _I32, _I32, _F, _P, _P, _I32, _P]),
    "gsmi_project_backward": (_I32, [_I64, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "gsmi_preprocess_views_workspace_size": (_SZ, [_I64, _I32]),
    "gsmi_preprocess_views_count": (_I32, [_I64, _I32, _P, _P, _P, _P, _F, _F, _I32, _P, _P, _P, _P, _SZ, _P]),
    "gsmi_preprocess_views_count_blocks": (_I32, [_I64, _I32, _P, _P, _P, _P, _F, _F, _I32, _I32, _P, _P, _P, _P,
                                                  _SZ, _P]),
    "gsmi_render_views_sh_two_level": (_I32, [_I64, _I32, _P, _P, _I32, _P, _P, _P, _I64, _I64, _P, _P, _P, _P, _P,
                                              _P, _P, _I32, _P, _I32, _F, _P, _P, _I32, _P]),
    "gsmi_preprocess_views_emit": (_I32, [_I64, _I32, _P, _P, _P, _P, _P, _F, _P, _P, _P, _P, _P, _P, _P, _P, _P,
                                          _P]),
    "gsmi_bin_views": (_I32, [_I64, _I64, _I32, _P, _P, _I32, _P, _P, _P, _P, _P, _P, _P, _SZ, _P]),
    "gsmi_render_views_sh": (_I32, [_I64, _I64, _I32, _P, _I32, _P, _P, _P, _I64, _I64, _P, _P, _P, _P, _P, _P,
                                    _I32, _F, _P, _P, _P, _P, _I32, _I32, _P]),
    "gsmi_render_views_backward_sh": (_I32, [_I64, _I64, _I32, _P, _I32, _P, _P, _P, _I64, _I64, _P, _P, _P, _P,
                                             _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _I32, _F, _P, _I32, _P, _SZ,
                                             _P]),
    "gsmi_two_level_offsets_views": (_I32, [_I32, _P, _P, _P, _I32, _P, _P, _P, _P, _SZ, _P]),
    "gsmi_render_views_sh_two_level_record": (_I32, [_I64, _I32, _P, _P, _I32, _P, _P, _P, _I64, _I64, _P, _P, _P,
                                                     _P, _P, _P, _P, _I32, _P, _P, _P, _P, _I32, _F, _P, _P, _I32,
                                                     _P]),
    "gsmi_render_views_backward_sh_lists": (_I32, [_I64, _I64, _I32, _P, _I32, _P, _P, _P, _I64, _I64, _P, _P, _P,
                                                   _P, _P, _P, _P, _P, _P, _P, _P, _P, _I32, _F, _I32, _P, _SZ,
                                                   _P]),
    "gsmi_project_backward_views": (_I32, [_I64, _I32, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P,
                                           _P]),
    "gsmi_loss_workspace_size": (_SZ, [_I64]),
    "gsmi_photometric_loss": (_I32, [_I64, _P, _P, _I32, _P, _P, _P, _SZ, _P]),
    "gsmi_ssim_loss_workspace_size": (_SZ, [_I32, _I32, _I32]),
    "gsmi_photometric_ssim_loss": (_I32, [_I32, _I32, _I32, _P, _P, _I32, _F, _I32, _P, _P, _P, _SZ, _P]),
    "gsmi_adam_step": (_I32, [_I32, _P, _P, _P, _P, _P, _P, _I32, _D, _D, _D, _I32, _P]),
    "gsmi_grad_stats": (_I32, [_I64, _P, _I32, _P, _I32, _P, _P]),
    "gsmi_count_visible": (_I32, [_I64, _P, _P, _P]),
    "gsmi_densify_workspace_size": (_SZ, [_I64]),
    "gsmi_split_clone_count": (_I32, [_I64, _P, _P, _I32, _F, _P, _F, _P, _P, _SZ, _P]),
    "gsmi_split_clone_write": (_I32, [_I64, _P, _P, _P, _P, _P, _I32, _P, _F, _P, _P, _P, _P, _P, _P, _SZ, _P]),
    "gsmi_prune_alpha_count": (_I32, [_I64, _P, _F, _P, _P, _SZ, _P]),
    "gsmi_prune_alpha_write": (_I32, [_I64, _P, _P, _P, _P, _P, _I32, _P, _P, _P, _P, _P, _P, _SZ, _P]),
    "gsmi_make_view": (_I32, [_P, _D, _D, _D, _D, _I32, _I32, _D, _D, _I32, _P]),
    "gsmi_server_create": (_I32, [_I32, _I64, _I32, _I32, _F, _F, _F, _I32, _P]),
    "gsmi_server_configure": (_I32, [_P, _I64, _I32]),
    "gsmi_server_add_scene": (_I32, [_P, _I64, _I32, _I64, _P, _P, _P, _P, _P, _P]),
    "gsmi_server_submit": (_I32, [_P, _I32, _P, _D, _D, _D, _D, _I32, _I32, _D, _D, _P, _P]),
    "gsmi_server_poll": (_I32, [_P, _P, _P, _P, _P, _P, _I32, _P]),
    "gsmi_server_drain": (_I32, [_P, _D]),
    "gsmi_server_stats": (_I32, [_P, _P]),
    "gsmi_server_stats_ex": (_I32, [_P, _P, _I32]),
    "gsmi_server_destroy": (_I32, [_P]),
    "gsmi_exp_spec": (_I32, [_I64, _P, _P, _P]),
    "gsmi_exp_spec_nonpos": (_I32, [_I64, _P, _P, _P]),
    "gsmi_div_staged": (_I32, [_I64, _P, _P, _P, _P, _P]),
    "gsmi_tile_order_from_costs": (_I32, [_I32, _P, _P, _P, _SZ, _P]),
    "gsmi_sched_arena_run": (_I32, [_I64, _I32, _P, _P]),
    "gsmi_sched_select_batch": (_I32, [_I32, _P, _P, _I32, _P, _I32, _I64, _P, _P, _P]),
    "gsmi_sched_admit": (_I32, [_I32, _P, _P, _P, ctypes.c_uint64, _P, _P, _P]),
}

_lib = None


def load():
    """Load libgsmi.so and bind every entry point; raises if it is missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"gaussian_splatting_3d_amd: HIP library {LIB_PATH} is not built; "
                "run `make -C gaussian_splatting_3d_amd/csrc` (or __graft_entry__.build())"
            )
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


class GsmiError(RuntimeError):
    pass


def check(rc: int, what: str):
    if rc != 0:
        msg = load().gsmi_status_string(rc).decode()
        raise GsmiError(f"{what} failed: {msg} (status {rc})")


def ptr(t):
    """Device/host address of a tensor (None -> NULL)."""
    if t is None:
        return None
    return ctypes.c_void_p(t.data_ptr())


def stream_of(device=None):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def host_f32(arr):
    """A ctypes float array holding `arr` (flattened) for host-array parameters."""
    import numpy as np

    a = np.ascontiguousarray(np.asarray(arr, dtype=np.float32).reshape(-1))
    return (ctypes.c_float * a.size)(*a.tolist())
