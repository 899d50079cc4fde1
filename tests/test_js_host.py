"""Node host (web-gs-renderer_amd/js) on the CPU: PLY loader, camera uniform
block, profiler rows, and a clean failure when no GPU is present."""
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
pytestmark = pytest.mark.skipif(NODE is None, reason="node is not installed")


def run_node(cmd, **args):
    r = subprocess.run([NODE, os.path.join(HERE, "js_driver.js"), cmd, json.dumps(args)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return json.loads(r.stdout)


def _scene(n=37, seed=3):
    return scenes.uniform_scene(n, seed=seed)


def test_parse_matches_packed_layout(tmp_path):
    sc = _scene()
    cols = list(write_ply(tmp_path / "a.ply", sc).keys())
    rng = np.random.default_rng(0)
    order = list(rng.permutation(cols))               # any property order
    types = {"x": "double", "opacity": "float64", "f_rest_3": "double"}
    write_ply(tmp_path / "b.ply", sc, order=order, types=types,
              extra=[("nx", "float"), ("red", "uchar"), ("flags", "int")])
    for name in ("a.ply", "b.ply"):
        out = tmp_path / (name + ".bin")
        r = run_node("parse", ply=str(tmp_path / name), out=str(out))
        assert r["ok"], r
        assert r["result"]["count"] == sc.count
        got = np.fromfile(out, dtype=np.float32)
        assert np.array_equal(got, sc.packed())


def test_parse_integer_and_crlf_types(tmp_path):
    sc = _scene(5)
    sc.opacity[:] = [1, -2, 3, 120, -100]
    expect = write_ply(tmp_path / "c.ply", sc, types={"opacity": "char"}, crlf=True)
    out = tmp_path / "c.bin"
    r = run_node("parse", ply=str(tmp_path / "c.ply"), out=str(out))
    assert r["ok"], r
    got = np.fromfile(out, dtype=np.float32)
    assert np.array_equal(got[13 * 5:14 * 5], expect["opacity"])


@pytest.mark.parametrize("case,needle", [
    ("missing", 'missing required property "f_rest_44"'),
    ("big_endian", "unsupported format"),
    ("list", "list properties are not supported"),
    ("zero", "vertex count missing or zero"),
    ("unknown_type", 'unknown property type "half"'),
    ("no_end", '"end_header" not found'),
    ("truncated", "vertices need"),
])
def test_parse_errors(tmp_path, case, needle):
    sc = _scene(4)
    p = tmp_path / f"{case}.ply"
    if case == "missing":
        write_ply(p, sc, drop=("f_rest_44",))
    elif case == "big_endian":
        write_ply(p, sc, fmt="binary_big_endian")
    elif case == "list":
        write_ply(p, sc, list_prop=True)
    elif case == "zero":
        write_ply(p, sc, count=0)
    elif case == "unknown_type":
        p.write_bytes(b"ply\nformat binary_little_endian 1.0\nelement vertex 1\nproperty half x\nend_header\n")
    elif case == "no_end":
        write_ply(p, sc, no_end=True)
    elif case == "truncated":
        write_ply(p, sc, truncate=5)
    r = run_node("parse", ply=str(p), out=str(tmp_path / "x.bin"))
    assert not r["ok"]
    assert needle in r["error"] and f"{case}.ply" in r["error"]


def test_camera_uniform_matches_harness(tmp_path):
    cam = scenes.Camera((0.5, -1.0, -6.0), (0.1, 0.2, 0.0), (0.0, 1.0, 0.0), 47.0, 640, 360)
    out = tmp_path / "u.bin"
    r = run_node("uniform", width=640, height=360, eye=list(cam.eye), target=list(cam.target), up=list(cam.up),
                 fov=47.0, count=1234, out=str(out))
    assert r["ok"], r
    got = np.fromfile(out, dtype=np.float32)
    ref = cam.uniform(1234)
    assert got.view(np.uint32)[35] == 1234
    assert np.array_equal(got, ref)


def test_profiler_rows_average_and_reset():
    frames = [[["preprocess", 1.0], ["sort", 2.0], ["TOTAL", 3.0]],
              [["preprocess", 3.0], ["sort", 4.0], ["TOTAL", 7.0]],
              [["preprocess", 5.0], ["sort", 6.0], ["TOTAL", 11.0]]]
    r = run_node("profiler", printEvery=2, frames=frames)["result"]["rows"]
    assert r[0] == {"preprocess": {"ms": 1.0}, "sort": {"ms": 2.0}, "TOTAL": {"ms": 3.0}}
    assert r[1] == {"preprocess": {"ms": 2.0}, "sort": {"ms": 3.0}, "TOTAL": {"ms": 5.0}}
    assert r[2] == {"preprocess": {"ms": 5.0}, "sort": {"ms": 6.0}, "TOTAL": {"ms": 11.0}}   # reset after 2


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU may be present")
def test_renderer_without_gpu_throws():
    r = run_node("create")
    assert not r["ok"] and "[gsr 2]" in r["error"] and "device" in r["error"]
