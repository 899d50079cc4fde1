"""The C-ABI library loads and exports every function include/gsr.h declares;
host-side error paths (no compute on the CPU)."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from gsr import lib


def test_library_exports_header_symbols():
    L = lib.load()
    names = lib.header_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_header_camera_block_is_176_bytes():
    assert ctypes.sizeof(lib.Camera) == 176
    assert ctypes.sizeof(lib.Options) == 40      # u32 u32 u64 u32 u32 u32 u32 u32 (+4 padding)


def _has_gpu():
    try:
        out = subprocess.run(["rocminfo"], capture_output=True, text=True, timeout=60).stdout
        return "gfx9" in out
    except Exception:
        return False


@pytest.mark.skipif(os.path.exists("/dev/kfd") and _has_gpu(), reason="a GPU is present")
def test_create_without_device_fails_cleanly():
    with pytest.raises(lib.GsrError) as e:
        lib.Context(0)
    assert e.value.code == lib.GSR_ERR_DEVICE
    assert "device" in str(e.value).lower()


def test_null_context_is_an_argument_error():
    L = lib.load()
    assert L.gsr_synchronize(None) == lib.GSR_ERR_ARGUMENT
    assert L.gsr_scene_select(None, b"x") == lib.GSR_ERR_ARGUMENT
    out = ctypes.c_void_p()
    assert L.gsr_create(None, 0) == lib.GSR_ERR_ARGUMENT
    assert "null" in lib.last_error(None)
    n = ctypes.c_int()
    assert L.gsr_get_timings(None, None, 0, ctypes.byref(n)) == lib.GSR_ERR_ARGUMENT
    del out


def test_uniform_block_layout_matches_reference():
    """GlobalUniforms slots (renderer.js:773-797): view@0, viewProj@16, eye@32,
    count@35 (u32), focal@36, tanFov@38, textureSize@40."""
    from gsr import scenes
    cam = scenes.Camera((0.0, 0.0, -4.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 60.0, 512, 256)
    u = cam.uniform(1234)
    c = lib.Camera.from_uniform(u)
    assert c.count == 1234 and c.texture_size[0] == 512 and c.texture_size[1] == 256
    assert np.isclose(c.tan_fov[1], np.tan(np.radians(30.0)))
    assert np.isclose(c.tan_fov[0], c.tan_fov[1] * 2.0)
    assert np.isclose(c.focal[0], 512 / (2 * c.tan_fov[0]))
    # view looks down -z: the origin is 4 units in front of the eye
    v = np.array(c.view[:]).reshape(4, 4).T
    assert np.allclose(v @ [0, 0, 0, 1], [0, 0, -4, 1], atol=1e-6)
