"""The batched extension through the JavaScript shim (N-API addon): gradientsBatch
and trainScene({batchViews}) against the same computations through the C ABI
from Python, and the addon's asynchronous readback."""
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
from gsr import lib, scenes
from test_train_scene import LR, run_node, setup_scene

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)


def test_gradients_batch_matches_c_abi():
    """gradientsBatch (SH3, four views of uneven sizes) equals lib.Batch's step:
    images bit for bit, gradients and STATS to float-atomic summation order;
    read() resolves, and a context with a read pending refuses other calls;
    batchTuning / batchSetTuning carry the eight knobs, depth slices included."""
    assert shutil.which("node"), "node missing"
    sc = scenes.config(3, n=5000, views=4, seed=21)
    sizes = [(96, 64), (160, 120), (64, 48), (33, 17)]
    cams = []
    for c, (w, h) in zip(sc.cams, sizes):
        s = w / c.width
        cams.append(scenes.Camera(c.view, c.fx * s, c.fy * s, w / 2.0, h / 2.0, w, h))
    rng = scenes.Rng(5)
    targets = [rng.uniform(c.pixels * 3).astype(np.float32) for c in cams]
    with tempfile.TemporaryDirectory() as d:
        camj = [{"view": [float(x) for x in c.view], "fx": c.fx, "fy": c.fy, "cx": c.cx, "cy": c.cy,
                 "width": c.width, "height": c.height} for c in cams]
        json.dump({"cams": camj, "ssim": 0.0}, open(os.path.join(d, "scene.json"), "w"))
        sc.gs.astype(np.float32).tofile(os.path.join(d, "gs.f32"))
        sc.sh.astype(np.float32).tofile(os.path.join(d, "sh.f32"))
        for v, t in enumerate(targets):
            t.tofile(os.path.join(d, f"target{v}.f32"))
        log("node gradientsBatch")
        res = subprocess.run(["node", os.path.join(ROOT, "tests", "node_batch.js"), d], capture_output=True,
                             text=True, timeout=300)
        assert res.returncode == 0, res.stderr
        done = json.load(open(os.path.join(d, "done.json")))
        ld = lambda name: np.fromfile(os.path.join(d, name), dtype=np.float32)  # noqa: E731
        js = {"images": [ld(f"image{v}.f32") for v in range(len(cams))], "grads": ld("grads.f32").reshape(-1, 16),
              "shape": ld("shape.f32").reshape(-1, 10), "sh": ld("shgrad.f32").reshape(-1, 16, 3),
              "stats": ld("stats.f32").reshape(-1, 2)}
    assert done["busy"] and done["readBytes"] == 1 << 20
    # batchTuning: six knobs, then the depth slices (2 x 8 by default at 100k Gaussians)
    assert len(done["tuningDefault"]) == 8 and done["tuningDefault"][6:] == [2, 8]
    assert done["tuningSet"] == [1, 2048, 0, 0, 1024, 256, 3, 4]
    n = sc.n
    with lib.Context(0) as ctx:
        gsb, shb = ctx.upload(sc.gs), ctx.upload(sc.sh)
        tgt = ctx.upload(np.concatenate(targets))
        total = sum(c.pixels for c in cams)
        grads, shape, shg, st, img = (ctx.alloc(n * 64), ctx.alloc(n * 40), ctx.alloc(n * 192), ctx.alloc(n * 8),
                                      ctx.alloc(total * 12))
        b = lib.Batch(ctx, cams, n, 3)
        b.step(gsb, shb, tgt, 0.0, grads, shape, shg, img, False, stats=st)
        ref = {"images": ctx.download(img, np.float32, (total * 3,)), "grads": ctx.download(grads, np.float32, (n, 16)),
               "shape": ctx.download(shape, np.float32, (n, 10)), "sh": ctx.download(shg, np.float32, (n, 16, 3)),
               "stats": ctx.download(st, np.float32, (n, 2))}
        assert list(b.entries()) == done["entries"]
        b.close()
    np.testing.assert_array_equal(np.concatenate(js["images"]), ref["images"])
    for k in ("shape", "sh"):
        assert rel_l2(js[k], ref[k]) < 1e-5, k
    assert rel_l2(js["grads"][:, 3], ref["grads"][:, 3]) < 1e-5
    np.testing.assert_array_equal(js["stats"][:, 1], ref["stats"][:, 1])
    assert rel_l2(js["stats"][:, 0], ref["stats"][:, 0]) < 1e-5
    assert js["stats"][:, 1].sum() > n


def replay_batched(start, cams, images, steps, rand, k):
    """trainScene({batchViews: k}) without densification, through the C ABI: the
    same shuffled views, k per step in one batched step, Adam on their sums."""
    n = start.shape[0]
    with lib.Context(0) as ctx:
        gs = ctx.upload(start)
        moments, stats = ctx.alloc(n * 28 * 4), ctx.alloc(n * 8)
        grads, shape = ctx.alloc(n * 64), ctx.alloc(n * 40)
        order, seq, batch = [], iter(rand), None
        for t in range(1, steps + 1):
            chosen = []
            while len(chosen) < k:
                if not order:
                    order = OF.shuffled(len(cams), lambda: next(seq))
                chosen.append(order.pop())
            ccams = [cams[v] for v in chosen]
            if batch is None:
                batch = lib.Batch(ctx, ccams, n, 0)
            else:
                batch.set_cameras(ccams)
            tgt = ctx.upload(np.concatenate([(images[v].astype(np.float64).ravel() / 255).astype(np.float32)
                                             for v in chosen]))
            batch.step(gs, None, tgt, 0.0, grads, shape, None, None, True, stats=stats)
            ctx.adam(float(t), n, LR, gs, grads, shape, moments)
            tgt.free()
        out = ctx.download(gs, np.float32, (n, 16))
        batch.close()
    return out


def test_train_scene_batched_matches_c_abi_replay():
    """trainScene with three views per step (batched) equals the same loop through
    the C ABI from Python, to the float-atomic nondeterminism of the backward."""
    assert shutil.which("node"), "node missing"
    rr = random.Random(19)
    rand = [rr.random() for _ in range(400)]
    with tempfile.TemporaryDirectory() as d:
        start, cams, images, camj = setup_scene(d)
        steps, k = 40, 3
        done, trained = run_node(d, {"cams": camj, "random": rand,
                                     "options": {"steps": steps, "densifyFrom": 10 ** 6, "batchViews": k}})
    assert done["count"] == start.shape[0]
    want = replay_batched(start, cams, images, steps, rand, k)
    assert np.isfinite(trained).all()
    e = rel_l2(trained, want)
    log(f"  batched trainScene vs C ABI replay: rel L2 {e:.2e}")
    assert e < 1e-3
    assert np.abs(trained - start).max() > 1e-3


def test_train_scene_batched_densifies_and_learns():
    assert shutil.which("node"), "node missing"
    rr = random.Random(20)
    rand = [rr.random() for _ in range(200000)]
    with tempfile.TemporaryDirectory() as d:
        start, cams, images, camj = setup_scene(d)
        opts = {"steps": 100, "densifyFrom": 50, "densifyEvery": 50, "densifyUntil": 60, "maxSplats": 5000,
                "batchViews": 4}
        done, trained = run_node(d, {"cams": camj, "random": rand, "options": opts})
    counts = [p["count"] for p in done["progress"]]
    # densified at step 50, before that step's report
    assert [p["step"] for p in done["progress"]] == [50, 100]
    assert start.shape[0] < counts[0] <= 5000 and counts[-1] == done["count"] == trained.shape[0]
    from gsr.renderer import Renderer
    r = Renderer()
    l1 = lambda gs: np.mean([np.abs(r.render(gs, c) - img / 255.0).mean() for c, img in zip(cams, images)])  # noqa: E731
    before, after = l1(start), l1(trained)
    r.close()
    log(f"  batched trainScene (4 views per step): counts {counts}, L1 over all views {before:.4f} -> {after:.4f}, "
        f"{opts['steps'] * 4 / done['seconds']:.0f} views/s")
    assert after < 0.8 * before
