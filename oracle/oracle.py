"""ctypes binding of the CPU oracle (oracle/gsr_oracle.cpp).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, __graft_entry__.smoke() and the
bench's cpu_baseline leg; never by the product package.  Each function takes a
``dtype`` of np.float32 (the reference-semantics restatement of trainer/gpu.js)
or np.float64 (the FP64 oracle).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libgsr_oracle.so")
SRC = os.path.join(HERE, "gsr_oracle.cpp")

_lib = None


def build(force: bool = False) -> str:
    """Compile the oracle with g++ (OpenMP, no FMA contraction)."""
    import subprocess

    os.makedirs(os.path.dirname(LIB_PATH), exist_ok=True)
    if force or not os.path.exists(LIB_PATH) or os.path.getmtime(LIB_PATH) < os.path.getmtime(SRC):
        tmp = LIB_PATH + ".tmp"
        subprocess.check_call(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-fopenmp",
                               "-o", tmp, SRC])
        os.replace(tmp, LIB_PATH)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        _lib.oracle_num_threads.restype = ctypes.c_int
        for suf in ("f32", "f64"):
            getattr(_lib, f"oracle_tile_lists_{suf}").restype = ctypes.c_uint32
            getattr(_lib, f"oracle_loss_value_{suf}").restype = ctypes.c_float if suf == "f32" else ctypes.c_double
            getattr(_lib, f"oracle_loss_grad_{suf}").argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                                                  ctypes.c_int,
                                                                  ctypes.c_float if suf == "f32" else ctypes.c_double,
                                                                  ctypes.c_void_p, ctypes.c_int]
            getattr(_lib, f"oracle_loss_value_{suf}").argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                                                   ctypes.c_int,
                                                                   ctypes.c_float if suf == "f32" else ctypes.c_double,
                                                                   ctypes.c_int]
    return _lib


def threads(n: int = 0) -> int:
    return lib().oracle_num_threads(n)


def _suf(dtype) -> str:
    return "f32" if np.dtype(dtype) == np.float32 else "f64"


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def cam_buf(cam, count: int) -> ctypes.Array:
    raw = cam.uniform(count)
    return ctypes.create_string_buffer(raw, len(raw))


def project(cam, gs, sh=None, deg=0, dtype=np.float64):
    """PREPROCESS (gpu.js:64-88).  Parameters are taken in the oracle's precision."""
    gs = _c(gs, dtype)
    n = gs.shape[0]
    shc = None if (sh is None or deg == 0) else _c(sh, dtype)
    out = np.zeros((n, 12), dtype=dtype)
    getattr(lib(), f"oracle_project_{_suf(dtype)}")(cam_buf(cam, n), _p(gs), _p(shc), ctypes.c_int(deg),
                                                    ctypes.c_uint32(n), _p(out))
    return out


def tile_lists(splats, width: int, height: int):
    """tileLists (gpu.js:450-481).  splats float32 or float64 [n, 12]."""
    dtype = np.float32 if splats.dtype == np.float32 else np.float64
    s = _c(splats, dtype)
    n = s.shape[0]
    T = ((width + 15) // 16) * ((height + 15) // 16)
    ranges = np.zeros(2 * T, dtype=np.uint32)
    f = getattr(lib(), f"oracle_tile_lists_{_suf(dtype)}")
    e = f(_p(s), ctypes.c_uint32(n), ctypes.c_int(width), ctypes.c_int(height), None, ctypes.c_uint64(0), _p(ranges))
    lst = np.zeros(max(1, e), dtype=np.uint32)
    f(_p(s), ctypes.c_uint32(n), ctypes.c_int(width), ctypes.c_int(height), _p(lst), ctypes.c_uint64(lst.size),
      _p(ranges))
    return lst[:e], ranges


def forward(splats, lst, ranges, width: int, height: int, dtype=np.float64):
    s = _c(splats, dtype)
    lst = _c(lst if lst.size else np.zeros(1, np.uint32), np.uint32)
    ranges = _c(ranges, np.uint32)
    image = np.zeros((height, width, 3), dtype=dtype)
    state = np.zeros((height, width, 2), dtype=dtype)
    getattr(lib(), f"oracle_forward_{_suf(dtype)}")(_p(s), _p(lst), _p(ranges), ctypes.c_int(width),
                                                    ctypes.c_int(height), _p(image), _p(state))
    return image, state


def loss_grad(image, target, lam: float, dtype=np.float64, window: int = 7):
    """dLoss/dpixel; window 7: the reference's 7x7 box SSIM, 11: 11x11 Gaussian."""
    image = _c(image, dtype)
    target = _c(target, dtype)
    H, W = image.shape[:2]
    g = np.zeros_like(image)
    getattr(lib(), f"oracle_loss_grad_{_suf(dtype)}")(_p(image), _p(target), W, H, lam, _p(g), window)
    return g


def loss_value(image, target, lam: float, dtype=np.float64, window: int = 7) -> float:
    image = _c(image, dtype)
    target = _c(target, dtype)
    H, W = image.shape[:2]
    return float(getattr(lib(), f"oracle_loss_value_{_suf(dtype)}")(_p(image), _p(target), W, H, lam, window))


def backward(splats, lst, ranges, state, dL, width: int, height: int, dtype=np.float64, threaded: bool = False):
    """BACKWARD (gpu.js:159-202).  threaded=True runs every host thread with
    atomic float adds (nondeterministic order; CPU baseline timing)."""
    s = _c(splats, dtype)
    n = s.shape[0]
    lst = _c(lst if lst.size else np.zeros(1, np.uint32), np.uint32)
    grads = np.zeros((n, 16), dtype=dtype)
    getattr(lib(), f"oracle_backward_{'mt_' if threaded else ''}{_suf(dtype)}")(_p(s), _p(lst), _p(_c(ranges, np.uint32)), _p(_c(state, dtype)),
                                                     _p(_c(dL, dtype)), ctypes.c_int(width), ctypes.c_int(height),
                                                     _p(grads))
    return grads


def shape_grad(cam, gs, grads, sh=None, deg=0, dtype=np.float64):
    gs = _c(gs, dtype)
    n = gs.shape[0]
    shc = None if (sh is None or deg == 0) else _c(sh, dtype)
    shape = np.zeros((n, 10), dtype=dtype)
    shg = np.zeros((n, 16, 3), dtype=dtype) if shc is not None else None
    getattr(lib(), f"oracle_shape_grad_{_suf(dtype)}")(cam_buf(cam, n), _p(gs), _p(shc), ctypes.c_int(deg),
                                                       ctypes.c_uint32(n), _p(_c(grads, dtype)), _p(shape), _p(shg))
    return shape, shg


def shape_fd(cam, gs, grads, dtype=np.float32):
    gs = _c(gs, dtype)
    n = gs.shape[0]
    shape = np.zeros((n, 10), dtype=dtype)
    getattr(lib(), f"oracle_shape_fd_{_suf(dtype)}")(cam_buf(cam, n), _p(gs), ctypes.c_uint32(n),
                                                     _p(_c(grads, dtype)), _p(shape))
    return shape


def step_buf(t: float, count: int, lr: dict) -> ctypes.Array:
    f = np.zeros(8, dtype=np.float32)
    f[0] = t
    f[2:7] = [lr["color"], lr["opacity"], lr["mean"], lr["scale"], lr["rotation"]]
    f.view(np.uint32)[1] = count
    raw = f.tobytes()
    return ctypes.create_string_buffer(raw, len(raw))


def adam(t, lr, gs, grads, shape, moments, dtype=np.float32):
    """ADAM (gpu.js:411-443) in place on copies; returns (gs, moments)."""
    gs = _c(gs, np.float32).copy()
    moments = _c(moments, dtype).copy()
    n = gs.shape[0]
    getattr(lib(), f"oracle_adam_{_suf(dtype)}")(step_buf(t, n, lr), _p(gs), _p(_c(grads, dtype)),
                                                 _p(_c(shape, dtype)), _p(moments))
    return gs, moments


def adam_sh(t, lr, lr_sh_rest, gs, sh, grads, shape, sh_grad, moments, dtype=np.float32):
    """Adam over the 59 numbers of an SH3 Gaussian; returns (gs, sh, moments)."""
    gs = _c(gs, np.float32).copy()
    sh = _c(sh, np.float32).copy()
    moments = _c(moments, dtype).copy()
    n = gs.shape[0]
    getattr(lib(), f"oracle_adam_sh_{_suf(dtype)}")(step_buf(t, n, lr), ctypes.c_float(lr_sh_rest), _p(gs), _p(sh),
                                                    _p(_c(grads, dtype)), _p(_c(shape, dtype)),
                                                    _p(_c(sh_grad, dtype)), _p(moments))
    return gs, sh, moments


def stats(grads, splats, st, dtype=np.float32):
    st = _c(st, dtype).copy()
    n = splats.shape[0]
    getattr(lib(), f"oracle_stats_{_suf(dtype)}")(ctypes.c_uint32(n), _p(_c(grads, dtype)), _p(_c(splats, dtype)),
                                                  _p(st))
    return st


def tile_reach(splats, tiles, gidx, width: int, height: int) -> np.ndarray:
    """Per (tile, Gaussian) pair: the largest alpha of alphaAt (gpu.js:92-103) over
    the tile's pixels in float64 (0 when the square or the power test rejects them
    all); below 1/255 the splat hits no pixel of the tile."""
    splats = np.ascontiguousarray(splats, dtype=np.float32)
    tiles = np.ascontiguousarray(tiles, dtype=np.uint32)
    gidx = np.ascontiguousarray(gidx, dtype=np.uint32)
    out = np.zeros(tiles.size, dtype=np.float64)
    lib().oracle_tile_reach(_p(splats), ctypes.c_uint64(tiles.size), _p(tiles), _p(gidx), ctypes.c_int(width),
                            ctypes.c_int(height), _p(out))
    return out


def sh_basis(deg: int, d, dtype=np.float64):
    d = _c(d, dtype)
    Y = np.zeros(16, dtype=dtype)
    getattr(lib(), f"oracle_sh_basis_{_suf(dtype)}")(ctypes.c_int(deg), _p(d), _p(Y))
    return Y


def project_shape(cam, q, dtype=np.float64):
    q = _c(q, dtype)
    out = np.zeros(7, dtype=dtype)
    ok = getattr(lib(), f"oracle_project_shape_{_suf(dtype)}")(cam_buf(cam, 1), _p(q), _p(out))
    return bool(ok), out


def render(cam, gs, sh=None, deg=0, dtype=np.float64):
    """project -> tileLists -> forward.  Returns (splats, list, ranges, image, state)."""
    splats = project(cam, gs, sh, deg, dtype)
    lst, ranges = tile_lists(splats, cam.width, cam.height)
    image, state = forward(splats, lst, ranges, cam.width, cam.height, dtype)
    return splats, lst, ranges, image, state


def gradients(cam, gs, dL, sh=None, deg=0, dtype=np.float64):
    """gradients() (gpu.js:826-839) with a given dLoss/dPixel: returns
    (image, grads[n,16], shape[n,10], sh_grad[n,16,3] or None), analytic chain."""
    splats, lst, ranges, image, state = render(cam, gs, sh, deg, dtype)
    grads = backward(splats, lst, ranges, state, dL, cam.width, cam.height, dtype)
    shape, shg = shape_grad(cam, gs, grads, sh, deg, dtype)
    return image, grads, shape, shg
