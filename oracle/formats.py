"""CPU restatement of the host-side formats of trainer/gpu.js, in Python.

TEST INFRASTRUCTURE ONLY: what home-measure-privacy_amd/js/formats.js is tested against.
Imported by tests/ only; never by the product package.

Gaussians are dicts with the reference's keys (gpu.js:512-518); arithmetic is
float64 as in JavaScript, stored values are rounded to float32 where the
reference stores into a Float32Array.

Pinned by (tests/test_formats.py):
  to_ply / from_ply   the header and values of tour/sample.ply (first 48 records,
                      tests/golden/sample_head.ply), byte-exact round trips, the
                      SH0 constant of the real spherical harmonics (1 / (2 sqrt pi)).
  to_splat            hand-computed records (unit Gaussians, rounding at .5, the
                      255 clamp, weight order, non-finite records dropped).
  densify             hand-built cases of every rule (opacity and size pruning,
                      clone, split with a known random sequence, the maxCount cap).
  neighbour_sizes     closed forms on lattices and isolated points, and the brute
                      force three nearest over all points where the 27 cubes
                      cover every neighbour.
  shuffled            the Fisher-Yates permutation for a known random sequence.
"""
from __future__ import annotations

import math
import re
import struct

import numpy as np

SH0 = 0.28209479177387814  # gpu.js:499, 991


def f32(x: float) -> float:
    return float(np.float32(x))


def js_round(x: float) -> float:
    """Math.round: nearest integer, halves toward +infinity."""
    fl = math.floor(x)
    return fl + 1 if x - fl >= 0.5 else fl


def shuffled(n: int, random) -> list[int]:
    """neighbours.js:56-63."""
    order = list(range(n))
    for i in range(n - 1, 0, -1):
        j = math.floor(random() * (i + 1))
        order[i], order[j] = order[j], order[i]
    return order


def scene_extent(views) -> float:
    """gpu.js:564-568; views: list of 12-float view rows."""
    centres = [[-(m[j] * m[3] + m[4 + j] * m[7] + m[8 + j] * m[11]) for j in range(3)] for m in views]
    mean = [sum(c[j] for c in centres) / len(centres) for j in range(3)]
    far = max([1e-3] + [math.hypot(c[0] - mean[0], c[1] - mean[1], c[2] - mean[2]) for c in centres])
    return 1.1 * far


def neighbour_sizes(xyz, count: int, mn=0.005, mx=0.1, fallback=0.06) -> np.ndarray:
    """neighbours.js:13-53: RMS distance to the 3 nearest others in the 27
    surrounding cubes of edge 0.1, clamped; float32 out."""
    cell = 0.1
    xyz = [float(v) for v in xyz[: count * 3]]
    cube = lambda p: tuple(math.floor(xyz[p * 3 + a] / cell) for a in range(3))  # noqa: E731
    cubes: dict = {}
    for p in range(count):
        cubes.setdefault(cube(p), []).append(p)
    out = np.zeros(count, dtype=np.float32)
    for p in range(count):
        ci, cj, ck = cube(p)
        best = [math.inf, math.inf, math.inf]
        for di in (-1, 0, 1):
            for dj in (-1, 0, 1):
                for dk in (-1, 0, 1):
                    for q in cubes.get((ci + di, cj + dj, ck + dk), ()):
                        if q == p:
                            continue
                        dx, dy, dz = (xyz[q * 3 + a] - xyz[p * 3 + a] for a in range(3))
                        d2 = dx * dx + dy * dy + dz * dz
                        best = sorted(best + [d2])[:3]
        found = [d for d in best if math.isfinite(d)]
        size = math.sqrt(sum(found) / len(found)) if found else fallback
        out[p] = min(mx, max(mn, size))
    return out


def init_from_points(xyz, rgb, count: int, size: float | None = None) -> list[dict]:
    """gpu.js:507-521."""
    sizes = np.full(count, size, dtype=np.float32) if size else neighbour_sizes(xyz, count)
    lo = math.log(0.1 / (1 - 0.1))
    out = []
    for i in range(count):
        s = math.log(float(sizes[i]))
        out.append({"mean": [float(xyz[i * 3 + a]) for a in range(3)], "logScale": [s, s, s],
                    "rotation": [1, 0, 0, 0], "opacityLogit": lo,
                    "color": [float(rgb[i * 3 + a]) for a in range(3)]})
    return out


def densify(gaussians: list[dict], stats, grad_threshold=2e-4, size_threshold=0.1, min_opacity=0.005,
            max_count=400000, max_size=math.inf, random=None) -> list[dict]:
    """gpu.js:575-599."""
    out = []
    room = max_count - len(gaussians)
    for i, g in enumerate(gaussians):
        if 1 / (1 + math.exp(-g["opacityLogit"])) < min_opacity:
            continue
        size = max(math.exp(v) for v in g["logScale"])
        if size > max_size:
            continue
        avg = stats[i * 2] / stats[i * 2 + 1] if stats[i * 2 + 1] else 0
        if avg < grad_threshold or room <= 0:
            out.append(g)
            continue
        room -= 1
        if size <= size_threshold:
            out += [g, {k: (list(v) if isinstance(v, list) else v) for k, v in g.items()}]
            continue
        for _ in range(2):
            c = {k: (list(v) if isinstance(v, list) else v) for k, v in g.items()}
            c["mean"] = [g["mean"][j] + (random() - 0.5) * math.exp(g["logScale"][j]) for j in range(3)]
            c["logScale"] = [v - math.log(1.6) for v in g["logScale"]]
            out.append(c)
    return out


def _props(rest: int) -> list[str]:
    return (["x", "y", "z", "nx", "ny", "nz", "f_dc_0", "f_dc_1", "f_dc_2"] + [f"f_rest_{k}" for k in range(rest)] +
            ["opacity", "scale_0", "scale_1", "scale_2", "rot_0", "rot_1", "rot_2", "rot_3"])


def to_ply(gaussians: list[dict]) -> bytes:
    """gpu.js:524-533 (+ f_rest_0-44 from `sh` when present)."""
    rest = 45 if gaussians and "sh" in gaussians[0] else 0
    props = _props(rest)
    header = "\n".join(["ply", "format binary_little_endian 1.0", f"element vertex {len(gaussians)}"] +
                       [f"property float {p}" for p in props] + ["end_header", ""])
    rows = []
    for g in gaussians:
        dc = list(g["sh"][:3]) if rest else g["color"]
        row = list(g["mean"]) + [0.0, 0.0, 0.0] + [(float(c) - 0.5) / SH0 for c in dc]
        if rest:
            row += [g["sh"][k * 3 + c] for c in range(3) for k in range(1, 16)]
        row += [g["opacityLogit"]] + list(g["logScale"]) + list(g["rotation"])
        rows.append(row)
    body = np.asarray(rows, dtype=np.float64).astype("<f4").tobytes() if rows else b""
    return header.encode("latin1") + body


def from_ply(data: bytes) -> list[dict]:
    """gpu.js:983-999 (+ `sh` from f_rest_* when present)."""
    head = data[:4096].decode("latin1")
    end = head.index("end_header\n") + len("end_header\n")
    header = head[:end]
    count = int(re.search(r"element vertex (\d+)", header).group(1))
    props = re.findall(r"property float (\w+)", header)
    P = len(props)
    vals = np.frombuffer(data[end:end + count * P * 4], dtype="<f4").reshape(count, P).astype(np.float64)
    col = {p: k for k, p in enumerate(props)}
    rest = 0
    while f"f_rest_{rest}" in col:
        rest += 1
    per = rest // 3
    out = []
    for i in range(count):
        v = lambda name: float(vals[i, col[name]])  # noqa: E731
        dc = [v(f"f_dc_{k}") * SH0 + 0.5 for k in range(3)]
        g = {"mean": [v("x"), v("y"), v("z")], "logScale": [v("scale_0"), v("scale_1"), v("scale_2")],
             "rotation": [v(f"rot_{k}") for k in range(4)], "opacityLogit": v("opacity"),
             "color": [min(1.0, max(0.0, c)) for c in dc]}
        if per:
            sh = [0.0] * 48
            sh[:3] = [f32(c) for c in dc]
            for c in range(3):
                for k in range(1, min(per, 15) + 1):
                    sh[k * 3 + c] = float(vals[i, col["f_rest_0"] + c * per + k - 1])
            g["sh"] = sh
        out.append(g)
    return out


def to_splat(gaussians: list[dict]) -> bytes:
    """gpu.js:539-560."""
    keep = [g for g in gaussians if all(math.isfinite(x) for x in list(g["mean"]) + list(g["logScale"]) +
                                        [g["opacityLogit"]] + list(g["rotation"]) + list(g["color"]))]

    def weight(g):
        return math.exp(g["logScale"][0] + g["logScale"][1] + g["logScale"][2]) / (1 + math.exp(-g["opacityLogit"]))

    order = sorted(range(len(keep)), key=lambda i: -weight(keep[i]))  # stable: ties keep input order
    byte = lambda x: int(max(0, min(255, js_round(x))))  # noqa: E731
    out = bytearray()
    for i in order:
        g = keep[i]
        out += struct.pack("<6f", *[f32(x) for x in g["mean"]], *[f32(math.exp(x)) for x in g["logScale"]])
        q = g["rotation"]
        ln = math.hypot(*q) or 1
        out += bytes([byte(c * 255) for c in g["color"]] + [byte(255 / (1 + math.exp(-g["opacityLogit"])))] +
                     [byte((x / ln) * 128 + 128) for x in q])
    return bytes(out)
