"""Host-side formats of the JavaScript shim (js/formats.js): .ply and .splat
I/O, densify, starting sizes, view shuffling and scene extent.

The oracle (oracle/formats.py) is pinned first: by tour/sample.ply's own bytes
and values, closed forms and hand-built cases.  The shim is then checked against
it, byte for byte and value for value, through Node.  No GPU.
"""
import base64
import json
import math
import os
import random
import shutil
import struct
import subprocess
import tempfile

import numpy as np
import pytest

import formats as OF
from conftest import ROOT

FIXTURE = os.path.join(ROOT, "tests", "golden", "sample_head.ply")


def gaussian(mean=(0, 0, 0), log_scale=(0, 0, 0), rotation=(1, 0, 0, 0), opacity_logit=0.0, color=(0.5, 0.5, 0.5),
             sh=None):
    g = {"mean": list(mean), "logScale": list(log_scale), "rotation": list(rotation),
         "opacityLogit": opacity_logit, "color": list(color)}
    if sh is not None:
        g["sh"] = list(sh)
    return g


# ---------------------------------------------------------------- oracle pins


def test_sh0_is_the_dc_basis_value():
    assert OF.SH0 == pytest.approx(1 / (2 * math.sqrt(math.pi)), abs=1e-16)


def test_sample_ply_values():
    gs = OF.from_ply(open(FIXTURE, "rb").read())
    assert len(gs) == 48
    for g in gs:  # every splat of tour/sample.ply: 6 cm, opacity logit 4, identity rotation
        assert g["logScale"] == [float(np.float32(math.log(0.06)))] * 3
        assert g["opacityLogit"] == 4.0
        assert g["rotation"] == [1.0, 0.0, 0.0, 0.0]
        assert all(0.0 <= c <= 1.0 for c in g["color"])
    raw = np.frombuffer(open(FIXTURE, "rb").read().split(b"end_header\n", 1)[1], dtype="<f4").reshape(48, 17)
    assert not raw[:, 3:6].any()  # normals
    np.testing.assert_allclose([g["color"] for g in gs], raw[:, 6:9].astype(np.float64) * OF.SH0 + 0.5, rtol=0, atol=1e-15)


def test_ply_rewrite_reproduces_sample_bytes():
    data = open(FIXTURE, "rb").read()
    assert OF.to_ply(OF.from_ply(data)) == data


def test_ply_sh_round_trip_and_layout():
    rng = np.random.default_rng(3)
    gs = [gaussian(mean=rng.normal(size=3), sh=rng.normal(size=48).astype(np.float32)) for _ in range(5)]
    data = OF.to_ply(gs)
    header = data.split(b"end_header\n")[0].decode()
    props = [ln.split()[-1] for ln in header.splitlines() if ln.startswith("property")]
    assert props[6:9] == ["f_dc_0", "f_dc_1", "f_dc_2"] and props[9] == "f_rest_0" and props[53] == "f_rest_44"
    assert props[54:] == ["opacity", "scale_0", "scale_1", "scale_2", "rot_0", "rot_1", "rot_2", "rot_3"]
    back = OF.from_ply(data)
    for g, b in zip(gs, back):
        np.testing.assert_array_equal(np.float32(b["sh"][3:]), np.float32(g["sh"][3:]))
        np.testing.assert_allclose(b["sh"][:3], g["sh"][:3], rtol=0, atol=1e-6)
    # f_rest is channel-major: f_rest_(c*15 + k-1) holds band coefficient k of channel c
    body = np.frombuffer(data.split(b"end_header\n", 1)[1], dtype="<f4").reshape(5, 62)
    assert body[0, 9 + 2 * 15 + 4] == np.float32(gs[0]["sh"][5 * 3 + 2])


def test_js_round_halves_up():
    assert [OF.js_round(x) for x in (2.5, 3.5, -0.5, 0.49999999999999994, 254.5)] == [3, 4, 0, 0, 255]


def test_splat_hand_computed():
    unit = gaussian()
    big = gaussian(mean=(1, 2, 3), log_scale=(math.log(2), 0, 0), rotation=(0, 0, 3, 4), opacity_logit=100.0,
                   color=(1.2, 0.0, 1.0))
    data = OF.to_splat([unit, big])
    assert len(data) == 64
    first, second = data[:32], data[32:]
    # the larger splat comes first; rotation (0,0,3,4)/5 -> (128, 128, 0.6*128+128, 0.8*128+128)
    assert struct.unpack("<6f", first[:24]) == (1.0, 2.0, 3.0, 2.0, 1.0, 1.0)
    assert list(first[24:]) == [255, 0, 255, 255, 128, 128, 205, 230]
    assert struct.unpack("<6f", second[:24]) == (0.0, 0.0, 0.0, 1.0, 1.0, 1.0)
    assert list(second[24:]) == [128, 128, 128, 128, 255, 128, 128, 128]
    poisoned = gaussian(opacity_logit=float("nan"))
    assert OF.to_splat([poisoned, unit]) == second


def test_splat_ties_keep_input_order():
    a, b = gaussian(mean=(1, 0, 0)), gaussian(mean=(2, 0, 0))
    data = OF.to_splat([a, b])
    assert struct.unpack("<f", data[:4])[0] == 1.0 and struct.unpack("<f", data[32:36])[0] == 2.0


def test_densify_rules():
    small = gaussian(log_scale=(math.log(0.05),) * 3)
    large = gaussian(mean=(1, 1, 1), log_scale=(math.log(0.5), math.log(0.2), math.log(0.3)))
    gs = [
        gaussian(opacity_logit=-10.0),                       # 0: opacity 4.5e-5 < 0.005: dropped
        gaussian(log_scale=(math.log(3.0), 0, 0)),           # 1: larger than maxSize 2: dropped
        gaussian(mean=(5, 0, 0)),                            # 2: gradient below threshold: kept
        small,                                               # 3: small, high gradient: cloned
        large,                                               # 4: large, high gradient: split
        gaussian(mean=(7, 0, 0)),                            # 5: high gradient but no room left: kept
    ]
    stats = [0, 0, 0, 0, 1e-5, 1, 3e-4, 1, 9e-4, 3, 1.0, 1]
    seq = iter([0.25, 0.75, 0.5, 1.0, 0.0, 0.5])
    out = OF.densify(gs, stats, grad_threshold=2e-4, size_threshold=0.1, max_count=8, max_size=2.0,
                     random=lambda: next(seq))
    # maxCount 8 with 6 inputs leaves room for two additions: the clone and the split
    assert len(out) == 6
    assert out[0] is gs[2] and out[1] is gs[3] and out[2] == gs[3] and out[2] is not gs[3] and out[5] is gs[5]
    shrink = math.log(1.6)
    c1, c2 = out[3], out[4]
    assert c1["mean"] == pytest.approx([1 - 0.25 * 0.5, 1 + 0.25 * 0.2, 1.0], abs=1e-15)
    assert c2["mean"] == pytest.approx([1 + 0.5 * 0.5, 1 - 0.5 * 0.2, 1.0], abs=1e-15)
    assert c1["logScale"] == [v - shrink for v in large["logScale"]] == c2["logScale"]
    # with room for one addition the large one is kept whole
    out2 = OF.densify(gs, stats, max_count=7, max_size=2.0, random=lambda: 0.5)
    assert [g["mean"] for g in out2] == [gs[2]["mean"], small["mean"], small["mean"], large["mean"], gs[5]["mean"]]


def test_neighbour_sizes_closed_forms():
    # a 4x4x4 lattice of spacing 0.03: every point has 3+ neighbours at 0.03
    grid = [(i * 0.03 + 0.001, j * 0.03 + 0.001, k * 0.03 + 0.001) for i in range(4) for j in range(4)
            for k in range(4)]
    far = [(5.0, 5.0, 5.0)]                     # alone: fallback 0.06
    pair = [(9.0, 9.0, 9.0), (9.0, 9.0, 9.001)]  # 1 mm apart: clamped to 0.005
    wide = [(-3.01, 0.0, 0.0), (-3.16, 0.0, 0.0)]  # 15 cm apart in adjacent cubes: clamped to 0.1
    pts = grid + far + pair + wide
    xyz = [c for p in pts for c in p]
    s = OF.neighbour_sizes(xyz, len(pts))
    np.testing.assert_allclose(s[:64], 0.03, rtol=1e-6)
    assert s[64] == np.float32(0.06) and s[65] == s[66] == np.float32(0.005) and s[67] == s[68] == np.float32(0.1)


def test_neighbour_sizes_match_brute_force_within_reach():
    rng = np.random.default_rng(5)
    pts = rng.uniform(0, 0.4, size=(150, 3))
    s = OF.neighbour_sizes(pts.ravel().tolist(), 150)
    d2 = ((pts[:, None, :] - pts[None, :, :]) ** 2).sum(-1)
    np.fill_diagonal(d2, np.inf)
    three = np.sort(d2, axis=1)[:, :3]
    reach = np.sqrt(three[:, 2]) < 0.1  # all three inside the 27 cubes
    want = np.clip(np.sqrt(three.mean(1)), 0.005, 0.1).astype(np.float32)
    assert reach.sum() > 100
    np.testing.assert_allclose(s[reach], want[reach], rtol=1e-6)


def test_shuffled_is_fisher_yates():
    seq = iter([0.9, 0.1, 0.5])
    # i=3: j=floor(0.9*4)=3; i=2: j=floor(0.1*3)=0 -> [2,1,0,3]; i=1: j=floor(0.5*2)=1
    assert OF.shuffled(4, lambda: next(seq)) == [2, 1, 0, 3]


def test_scene_extent_closed_form():
    # cameras at +-1 on x (view rows [I | -c]): mean 0, largest distance 1
    views = [[1, 0, 0, -c, 0, 1, 0, 0, 0, 0, 1, 0] for c in (1.0, -1.0)]
    assert OF.scene_extent(views) == pytest.approx(1.1)


# ---------------------------------------------------------------- the shim


def same(a, b, rel=1e-15):
    """Equal structure and values; floats to `rel` (V8's Math.log / Math.exp and
    glibc's may differ in the last ulp, which is all this allows)."""
    if isinstance(a, dict):
        return isinstance(b, dict) and a.keys() == b.keys() and all(same(a[k], b[k], rel) for k in a)
    if isinstance(a, (list, tuple)):
        return len(a) == len(b) and all(same(x, y, rel) for x, y in zip(a, b))
    if isinstance(a, float) or isinstance(b, float):
        return a == b or abs(a - b) <= rel * max(abs(a), abs(b))
    return a == b



def test_js_formats_match_oracle():
    if not shutil.which("node"):
        pytest.skip("node missing")
    rng = np.random.default_rng(11)
    r = random.Random(4)
    gs = []
    for i in range(40):
        gs.append(gaussian(mean=rng.normal(size=3).tolist(), log_scale=rng.normal(-2, 1, size=3).tolist(),
                           rotation=rng.normal(size=4).tolist(), opacity_logit=float(rng.normal(0, 3)),
                           color=rng.uniform(-0.1, 1.1, size=3).tolist()))
    gs[7]["color"] = [0.5, 2.5 / 255, 1.0]
    stats = [abs(float(x)) for x in rng.normal(0, 5e-4, size=80)]
    for i in range(0, 80, 2):
        stats[i + 1] = float(r.randint(0, 4))
    xyz = rng.uniform(0, 0.3, size=60 * 3).tolist()
    rgb = rng.uniform(0, 1, size=60 * 3).tolist()
    views = [rng.normal(size=12).tolist() for _ in range(5)]
    rand = [r.random() for _ in range(200)]
    shuffle_rand = [r.random() for _ in range(30)]
    opts = {"gradThreshold": 2e-4, "sizeThreshold": 0.1, "minOpacity": 0.005, "maxCount": 52, "maxSize": 1.0}
    sh_gs = [dict(g, sh=rng.normal(size=48).astype(np.float32).tolist()) for g in gs[:6]]
    with tempfile.TemporaryDirectory() as d:
        job = {"gaussians": gs + [dict(g) for g in gs[:2]], "poison": [40, 41], "stats": stats,
               "densifyOptions": opts, "random": rand, "xyz": xyz, "rgb": rgb, "fixture": FIXTURE,
               "shuffleN": 17, "shuffleRandom": shuffle_rand, "views": views}
        path = os.path.join(d, "job.json")
        json.dump(job, open(path, "w"))
        res = subprocess.run(["node", os.path.join(ROOT, "tests", "node_formats.js"), path], capture_output=True,
                             text=True, timeout=120)
        assert res.returncode == 0, res.stderr
        out = json.loads(res.stdout)
        job["gaussians"], job["poison"] = sh_gs, []
        json.dump(job, open(path, "w"))
        res = subprocess.run(["node", os.path.join(ROOT, "tests", "node_formats.js"), path], capture_output=True,
                             text=True, timeout=120)
        assert res.returncode == 0, res.stderr
        out_sh = json.loads(res.stdout)
    poisoned = [dict(g) for g in gs[:2]]
    for g in poisoned:
        g["mean"] = [float("nan")] + g["mean"][1:]
    assert base64.b64decode(out["splat"]) == OF.to_splat(gs + poisoned)
    assert base64.b64decode(out["ply"]) == OF.to_ply(gs)
    assert out["plyBack"] == OF.from_ply(OF.to_ply(gs))
    fixture = open(FIXTURE, "rb").read()
    assert out["fixture"] == OF.from_ply(fixture)
    assert base64.b64decode(out["fixtureAgain"]) == fixture
    seq = iter(rand)
    want = OF.densify(gs, stats, grad_threshold=2e-4, size_threshold=0.1, min_opacity=0.005, max_count=52,
                      max_size=1.0, random=lambda: next(seq))
    assert same(out["densify"], want)
    sizes = OF.neighbour_sizes(xyz, 60)
    assert out["sizes"] == [float(x) for x in sizes]
    assert same(out["init"], OF.init_from_points(xyz, rgb, 60))
    assert same(out["initFixed"], OF.init_from_points(xyz, rgb, 60, 0.02))
    seq2 = iter(shuffle_rand)
    assert out["shuffled"] == OF.shuffled(17, lambda: next(seq2))
    assert out["extent"] == OF.scene_extent(views)
    assert base64.b64decode(out_sh["ply"]) == OF.to_ply(sh_gs)
    assert out_sh["plyBack"] == OF.from_ply(OF.to_ply(sh_gs))
