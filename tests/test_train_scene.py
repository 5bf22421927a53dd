"""trainScene through the JavaScript shim (gpu.js:890-959 on the N-API addon).

1. Without densification, the shim's run is replayed step by step through the C
   ABI from Python (same view order, same learning rates): the trained Gaussians
   agree to the nondeterminism of the backward's float atomics.
2. With densification, the splat count grows at the first densify step and stays
   within maxSplats, and the L1 over all views falls.  The densify rules
   themselves are checked on the CPU (test_formats.py).
"""
import json
import os
import random
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

import formats as OF
from conftest import ROOT, rel_l2
from gsr import scenes
from gsr.renderer import Renderer

LR = {"color": 0.025, "opacity": 0.05, "mean": 0.002, "scale": 0.005, "rotation": 0.001}  # gpu.js:891


def setup_scene(d, n=3000, k=12, size=96):
    rng = scenes.Rng(77)
    truth, _ = scenes.gaussians(rng, n, 0)
    cams = [scenes.look_at(3.0 * v, [0, 0, 0], size, size, 60.0) for v in scenes.fibonacci_sphere(k)]
    r = Renderer()
    images = [np.round(np.clip(r.render(truth, c), 0, 1) * 255).astype(np.uint8) for c in cams]
    r.close()
    start = truth.copy()
    start[:, 12:15] = rng.uniform(n * 3).reshape(n, 3)
    start[:, 7] = -1.0
    start.tofile(os.path.join(d, "start.f32"))
    for j, img in enumerate(images):
        img.tofile(os.path.join(d, f"view{j}.u8"))
    camj = [{"view": [float(x) for x in c.view], "fx": c.fx, "fy": c.fy, "cx": c.cx, "cy": c.cy,
             "width": c.width, "height": c.height} for c in cams]
    return start, cams, images, camj


def run_node(d, job):
    json.dump(job, open(os.path.join(d, "job.json"), "w"))
    print("node trainScene", job["options"], flush=True)
    res = subprocess.run(["node", os.path.join(ROOT, "tests", "node_train.js"), d], capture_output=True, text=True,
                         timeout=300)
    assert res.returncode == 0, res.stderr
    done = json.load(open(os.path.join(d, "done.json")))
    trained = np.fromfile(os.path.join(d, "trained.f32"), dtype=np.float32).reshape(-1, 16)
    return done, trained


def replay(start, cams, images, steps, rand):
    """The trainScene step sequence without densification, through the C ABI."""
    r = Renderer()
    ctx = r.ctx
    n = start.shape[0]
    gs = ctx.upload(start)
    moments, stats = ctx.alloc(n * 28 * 4), ctx.alloc(n * 8)
    order, seq = [], iter(rand)
    for t in range(1, steps + 1):
        if not order:
            order = OF.shuffled(len(cams), lambda: next(seq))
        v = order.pop()
        cam = cams[v]
        f = r.forward(gs, n, cam)
        tb = ctx.upload((images[v].astype(np.float64) / 255).astype(np.float32))
        grads, lg, shape = ctx.alloc(n * 64), ctx.alloc(cam.pixels * 12), ctx.alloc(n * 40)
        ctx.loss_grad(cam, f.image, tb, 0.0, lg)
        ctx.backward(cam, f.splats, f.order, f.ranges, f.image, lg, grads)
        ctx.shape(cam, gs, n, grads, shape)
        ctx.stats(n, grads, f.splats, stats)
        ctx.adam(float(t), n, LR, gs, grads, shape, moments)
        for b in (tb, grads, lg, shape):
            b.free()
        f.free()
    out = ctx.download(gs, np.float32, (n, 16))
    r.close()
    return out


@pytest.mark.gpu
def test_train_scene_matches_c_abi_replay():
    assert shutil.which("node"), "node missing"
    rr = random.Random(9)
    rand = [rr.random() for _ in range(400)]
    with tempfile.TemporaryDirectory() as d:
        start, cams, images, camj = setup_scene(d)
        steps = 60
        done, trained = run_node(d, {"cams": camj, "random": rand,
                                     "options": {"steps": steps, "densifyFrom": 10 ** 6}})
    assert done["count"] == start.shape[0]
    want = replay(start, cams, images, steps, rand)
    assert np.isfinite(trained).all()
    assert rel_l2(trained, want) < 1e-3
    moved = np.abs(trained - start).max()
    assert moved > 1e-3  # the run trained


@pytest.mark.gpu
def test_train_scene_densifies_and_learns():
    assert shutil.which("node"), "node missing"
    rr = random.Random(10)
    rand = [rr.random() for _ in range(200000)]
    with tempfile.TemporaryDirectory() as d:
        start, cams, images, camj = setup_scene(d)
        opts = {"steps": 200, "densifyFrom": 100, "densifyEvery": 50, "densifyUntil": 150, "maxSplats": 5000}
        done, trained = run_node(d, {"cams": camj, "random": rand, "options": opts})
    prog = done["progress"]
    assert [p["step"] for p in prog] == [50, 100, 150, 200]
    counts = [p["count"] for p in prog]
    assert counts[0] == start.shape[0] and counts[1] > counts[0]
    assert counts[-1] <= 5000 and counts[-1] == done["count"] == trained.shape[0]
    assert np.isfinite(trained).all()
    # photometric L1 over all views, rendered before and after
    r = Renderer()
    l1 = lambda gs: np.mean([np.abs(r.render(gs, c) - img / 255.0).mean() for c, img in zip(cams, images)])  # noqa: E731
    before, after = l1(start), l1(trained)
    r.close()
    assert after < 0.8 * before
    print(f"trainScene: {opts['steps'] / done['seconds']:.1f} steps/s at 96x96, counts {counts}, "
          f"L1 over all views {before:.4f} -> {after:.4f}", flush=True)
