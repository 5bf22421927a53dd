"""The JavaScript shim (home-measure-privacy_amd/js/gpu.js) over the N-API addon:
same results as the C ABI called from Python, errors thrown as JS Errors."""
import json
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

import oracle as O
from conftest import ROOT, rel_l2
from gsr import scenes
from gsr.renderer import Renderer

ADDON = os.path.join(ROOT, "home-measure-privacy_amd", "js", "gsr.node")


def test_addon_exports_every_operator():
    if not shutil.which("node") or not os.path.exists(ADDON):
        pytest.skip("node or gsr.node missing")
    out = subprocess.run(["node", "-e", f"console.log(Object.keys(require({json.dumps(ADDON)})).join(','))"],
                         capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    names = set(out.stdout.strip().split(","))
    for op in ("project", "binSort", "forward", "lossGrad", "backward", "shape", "adam", "stats", "alloc", "upload",
               "download", "read", "binSortAsync", "setCulling", "lossGradWindow", "adamSh", "batchCreate",
               "batchDestroy", "batchSetCameras", "batchStep", "batchEntries", "batchAutotune", "batchTuning",
               "batchSetTuning", "batchSetPasses", "batchSetSsimWindow", "batchSetCulling"):
        assert op in names, op


@pytest.mark.gpu
def test_js_shim_matches_python_path():
    assert shutil.which("node"), "node missing"
    sc = scenes.config(0, n=3000)
    cam = sc.cams[0]
    with tempfile.TemporaryDirectory() as d:
        camj = {"view": [float(x) for x in cam.view], "fx": cam.fx, "fy": cam.fy, "cx": cam.cx, "cy": cam.cy,
                "width": cam.width, "height": cam.height}
        json.dump({"cam": camj, "ssim": 0.0}, open(os.path.join(d, "scene.json"), "w"))
        sc.gs.astype(np.float32).tofile(os.path.join(d, "gs.f32"))
        sc.targets[0].astype(np.float32).tofile(os.path.join(d, "target.f32"))
        splats = O.project(cam, sc.gs, None, 0, np.float32)
        splats.tofile(os.path.join(d, "splats.f32"))
        print("node check", flush=True)
        r = subprocess.run(["node", os.path.join(ROOT, "tests", "node_check.js"), d], capture_output=True, text=True,
                           timeout=300)
        assert r.returncode == 0, r.stderr
        done = json.load(open(os.path.join(d, "done.json")))
        assert done["threwOnBadCamera"]
        ld = lambda name, dt: np.fromfile(os.path.join(d, name), dtype=dt)  # noqa: E731
        py = Renderer()
        ref = py.gradients(sc.gs, cam, sc.targets[0], 0.0)
        lg = py.lossGradientOf(ld("render.f32", np.float32).reshape(cam.height, cam.width, 3), sc.targets[0],
                               cam.width, cam.height, 0.2)
        py.close()
        np.testing.assert_array_equal(ld("image.f32", np.float32), ref["image"].ravel())
        np.testing.assert_array_equal(ld("render.f32", np.float32), ref["image"].ravel())
        assert rel_l2(ld("grads.f32", np.float32), ref["grads"].ravel()) < 1e-5
        assert rel_l2(ld("shape.f32", np.float32), ref["shape"].ravel()) < 1e-5
        np.testing.assert_array_equal(ld("lossgrad.f32", np.float32), lg)
        lst, ranges = O.tile_lists(splats, cam.width, cam.height)
        np.testing.assert_array_equal(ld("list.u32", np.uint32)[:lst.size], lst)
        np.testing.assert_array_equal(ld("ranges.u32", np.uint32), ranges)
