"""The C-ABI library loads and exports exactly what include/gsmi.h declares.

No compute call is made (no GPU here): only load, symbol lookup, parameter
counts against the header, and the pure-host helpers (status strings,
version, workspace-size queries).
"""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "gsmi.h")


def _declarations():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    decls = {}
    for m in re.finditer(r"\b(gsmi_\w+)\s*\(([^;{]*?)\)\s*;", text, flags=re.S):
        args = m.group(2).strip()
        n = 0 if args in ("", "void") else args.count(",") + 1
        decls[m.group(1)] = n
    return decls


@pytest.fixture(scope="module")
def lib():
    from gaussian_splatting_3d_amd import _lib

    if not os.path.exists(_lib.LIB_PATH):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "gaussian_splatting_3d_amd", "csrc")], check=True)
    return _lib.load()


def test_header_declares_entry_points():
    d = _declarations()
    for name in ("gsmi_tile_based_vol_rendering_sh", "gsmi_tile_based_vol_rendering_backward_sh",
                 "gsmi_tile_culling_aabb_start_end", "gsmi_culling_gaussian_bsphere", "gsmi_sort_pairs",
                 "gsmi_preprocess_count", "gsmi_preprocess_emit", "gsmi_project_backward"):
        assert name in d


def test_library_exports_every_declared_symbol(lib):
    from gaussian_splatting_3d_amd import _lib

    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    missing = set(_declarations()) - exported
    assert not missing, missing
    for name in _declarations():
        assert getattr(lib, name) is not None


def test_ctypes_signatures_match_header(lib):
    from gaussian_splatting_3d_amd import _lib

    decls = _declarations()
    assert set(_lib.SIGNATURES) == set(decls)
    for name, n in decls.items():
        assert len(_lib.SIGNATURES[name][1]) == n, name


def test_host_helpers(lib):
    assert lib.gsmi_version() >= 100
    assert lib.gsmi_status_string(0) == b"ok"
    assert b"workspace" in lib.gsmi_status_string(4)
    assert lib.gsmi_sort_workspace_size(0) > 0
    a, b = lib.gsmi_sort_workspace_size(10_000), lib.gsmi_sort_workspace_size(1_000_000)
    assert b > a > 10_000 * 12
    assert lib.gsmi_binning_workspace_size(1000, 5000, 256) >= lib.gsmi_sort_workspace_size(5000)
    assert lib.gsmi_preprocess_workspace_size(1_000_000) >= 8_000_000


def test_argument_validation_without_gpu(lib):
    """Invalid sizes are rejected before any launch."""
    assert lib.gsmi_sort_pairs(-1, None, None, 40, None, 0, None) == 1
    assert lib.gsmi_sort_pairs(10, None, None, 40, None, 0, None) == 1
    assert lib.gsmi_tile_ranges(0, None, 0, None, None, None) == 1
    P = ctypes.c_void_p(16)
    # C = 5 is outside the supported SH orders
    rc = lib.gsmi_render_sh_forward(1, 1, P, P, P, 75, 25, None, P, P, P, P, P, P, P, 16, 1, 1, 1.0, 1.0,
                                    16, 16, 5, 1e-4, None, None, None, None, 0, 0, None)
    assert rc == 2
    rc = lib.gsmi_render_sh_forward(1, 1, P, P, P, 3, 1, None, P, P, P, P, P, P, P, 8, 1, 1, 1.0, 1.0,
                                    16, 16, 1, 1e-4, None, None, None, None, 0, 0, None)
    assert rc == 2  # tile size 8
    # the reference-shaped entries: C = 5 and tile size 8 as above
    assert lib.gsmi_tile_based_vol_rendering_sh(1, 1, P, P, P, P, P, P, P, P, P, P, 16, 1, 1, 1.0, 1.0, 16, 16, 5,
                                                1e-4, None) == 2
    assert lib.gsmi_tile_based_vol_rendering_backward_sh(1, 1, P, P, P, P, P, P, P, P, P, P, P, P, P, P, P, 8, 1, 1,
                                                         1.0, 1.0, 16, 16, 1, 1e-4, None) == 2
    assert lib.gsmi_tile_based_vol_rendering_backward_start_end(1, 1, P, P, P, P, P, P, P, P, P, P, P, P, P, P, 16,
                                                                0, 1, 1.0, 1.0, 16, 16, 1e-4, None) == 1

    def bwd(T, nd, pre, bg, flags, ws, nws):
        return lib.gsmi_render_sh_backward(1, 1, P, P, P, 48, 16, None, P, P, P, P, P, T, nd, pre, P, P, P, P, P, P,
                                           P, 16, 1, 1, 1.0, 1.0, 16, 16, 4, 1e-4, bg, flags, ws, nws, None)

    assert bwd(P, None, None, None, 0, None, 0) == 1  # T_final without n_done
    assert bwd(None, P, None, None, 0, None, 0) == 1
    assert bwd(None, None, None, None, 1, None, 0) == 2  # the bf16 split needs saved state
    assert bwd(None, None, None, None, 0, P, 64) == 1  # the launch order needs saved n_done
    assert bwd(P, P, None, None, 4, None, 0) == 1  # undefined flag bit
    assert bwd(P, P, None, P, 0, None, 0) == 1  # background with saved state needs pre_bg


def test_make_view_matches_python_records(lib):
    """The server builds camera records natively; they must equal the Python
    ones bit for bit (frustum planes, image constants), or served images
    would differ from the per-call pipeline."""
    import numpy as np

    from gaussian_splatting_3d_amd import multiview, scenes, serving
    from gaussian_splatting_3d_amd.camera import CameraInfo

    cams = [scenes.orbit_camera(6.0, 0.1 + 0.7 * k, 1100.0 * w / 1297, 1100.0 * w / 1297, w, h, near=0.01,
                                far=1000.0, height=0.3 * k) for k, (w, h) in
            enumerate([(640, 480), (1280, 720), (1920, 1080), (1297, 840), (100, 70)])]
    cams.append((scenes.plain_forward_10k(n=10).cameras[0]))
    cams.append((cams[0][0], CameraInfo.from_fov_camera(0.9, 1.5, 333, 0.05, 50.0)))
    for c2w, cam in cams:
        want, _, _, _ = multiview.make_views([(c2w, cam)])
        got = serving.make_view_record(c2w, cam)
        assert np.array_equal(got.view(np.uint32), want[0].view(np.uint32)), (cam.w, cam.h)
