"""The Node host end to end on the GPU: PLY file -> JS loader -> N-API addon ->
C ABI -> HIP frame, image checked against the oracle."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

from gsr import scenes
from plyfile import write_ply

HERE = os.path.dirname(os.path.abspath(__file__))
NODE = shutil.which("node")
pytestmark = [pytest.mark.gpu, pytest.mark.skipif(NODE is None, reason="node is not installed")]


def test_node_render_matches_oracle(tmp_path, orc):
    sc, cam = scenes.workload_small()
    write_ply(tmp_path / "s.ply", sc)
    args = dict(ply=str(tmp_path / "s.ply"), width=cam.width, height=cam.height, eye=list(cam.eye),
                target=list(cam.target), up=list(cam.up), fov=cam.fovy_deg, out=str(tmp_path / "img.bin"),
                out_rgba=str(tmp_path / "img.rgba"))
    print("node render 16k splats", flush=True)
    r = subprocess.run([NODE, os.path.join(HERE, "js_driver.js"), "render", json.dumps(args)], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout)
    assert out["ok"], out
    res = out["result"]
    img = np.fromfile(tmp_path / "img.bin", dtype=np.float32).reshape(3, cam.height, cam.width)
    ref, _, _, I, _ = orc.render(sc.packed(), sc.count, cam.uniform(sc.count))
    assert np.abs(img - ref).max() <= 1e-4
    assert np.array_equal(img, ref)
    assert res["instances"] == I
    rgba = np.fromfile(tmp_path / "img.rgba", dtype=np.uint8).reshape(cam.height, cam.width, 4)
    expect = np.rint(np.clip(ref, 0, 1) * 255).astype(np.uint8).transpose(1, 2, 0)[::-1]
    assert np.array_equal(rgba[..., :3], expect) and (rgba[..., 3] == 255).all()
    assert res["profiles"] == 2
    assert set(res["lastProfile"]) >= {"preprocess", "offset scan", "emit", "sort", "tile ranges", "raster", "TOTAL"}
    assert "Invalid Gs buffer" in res["badScene"]
