"""ctypes binding of libgsr.so (include/gsr.h) for Python hosts, tests and the bench.

The library must have been built (``python home-measure-privacy_amd/gsr/build.py``);
loading fails loudly otherwise: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libgsr.so")

GSR_OK = 0
GSR_E_INVALID = -1
GSR_E_HIP = -2
GSR_E_NOMEM = -3
GSR_E_CAPACITY = -4

# Every entry point declared in include/gsr.h.
EXPORTS = (
    "gsr_create", "gsr_destroy", "gsr_last_error", "gsr_alloc", "gsr_free", "gsr_upload", "gsr_download",
    "gsr_clear", "gsr_copy", "gsr_synchronize", "gsr_project", "gsr_bin_sort", "gsr_forward", "gsr_loss_grad", "gsr_backward",
    "gsr_shape", "gsr_adam", "gsr_adam_sh", "gsr_stats", "gsr_batch_create", "gsr_batch_destroy", "gsr_batch_set_cameras",
    "gsr_batch_step", "gsr_batch_entries", "gsr_batch_configure", "gsr_batch_autotune", "gsr_batch_profile",
    "gsr_batch_set_passes", "gsr_loss_grad_window", "gsr_batch_set_ssim_window", "gsr_batch_set_split",
    "gsr_batch_tuning", "gsr_batch_passes", "gsr_batch_set_emit_band", "gsr_set_culling", "gsr_batch_step_stats",
    "gsr_batch_set_culling", "gsr_batch_tile_lists", "gsr_bin_sort_noalloc",
)
SSIM_BOX7, SSIM_GAUSS11 = 7, 11
STAGES = ("project", "depth_sort", "bin_count", "bin_emit", "forward", "loss_grad", "backward", "chain")


class GsrError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"gsr error {code}: {msg}")
        self.code = code


class Camera(ctypes.Structure):
    _fields_ = [("view", ctypes.c_float * 12), ("fx", ctypes.c_float), ("fy", ctypes.c_float),
                ("cx", ctypes.c_float), ("cy", ctypes.c_float), ("width", ctypes.c_float),
                ("height", ctypes.c_float), ("count", ctypes.c_uint32), ("pixels", ctypes.c_uint32)]


class Step(ctypes.Structure):
    _fields_ = [("t", ctypes.c_float), ("count", ctypes.c_uint32), ("lr_color", ctypes.c_float),
                ("lr_opacity", ctypes.c_float), ("lr_mean", ctypes.c_float), ("lr_scale", ctypes.c_float),
                ("lr_rotation", ctypes.c_float), ("pad", ctypes.c_float)]


assert ctypes.sizeof(Camera) == 80 and ctypes.sizeof(Step) == 32

_lib = None
P = ctypes.c_void_p


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with home-measure-privacy_amd/gsr/build.py")
    lib = ctypes.CDLL(LIB_PATH)
    u32, u64, sz, i32, f32 = ctypes.c_uint32, ctypes.c_uint64, ctypes.c_size_t, ctypes.c_int, ctypes.c_float
    sig = {
        "gsr_create": (i32, [i32, ctypes.POINTER(P)]),
        "gsr_destroy": (None, [P]),
        "gsr_last_error": (ctypes.c_char_p, [P]),
        "gsr_alloc": (i32, [P, sz, ctypes.POINTER(P)]),
        "gsr_free": (i32, [P, P]),
        "gsr_upload": (i32, [P, P, P, sz]),
        "gsr_download": (i32, [P, P, P, sz]),
        "gsr_clear": (i32, [P, P, sz]),
        "gsr_copy": (i32, [P, P, P, sz]),
        "gsr_synchronize": (i32, [P]),
        "gsr_project": (i32, [P, ctypes.POINTER(Camera), P, P, i32, u32, P]),
        "gsr_bin_sort": (i32, [P, ctypes.POINTER(Camera), P, u32, P, u64, P, ctypes.POINTER(u64)]),
        "gsr_forward": (i32, [P, ctypes.POINTER(Camera), P, P, P, P, P]),
        "gsr_loss_grad": (i32, [P, ctypes.POINTER(Camera), P, P, f32, P]),
        "gsr_backward": (i32, [P, ctypes.POINTER(Camera), P, P, P, P, P, P]),
        "gsr_shape": (i32, [P, ctypes.POINTER(Camera), P, P, i32, u32, P, P, P]),
        "gsr_adam": (i32, [P, ctypes.POINTER(Step), P, P, P, P]),
        "gsr_adam_sh": (i32, [P, ctypes.POINTER(Step), f32, P, P, P, P, P, P]),
        "gsr_stats": (i32, [P, ctypes.POINTER(Step), P, P, P]),
        "gsr_batch_create": (i32, [P, u32, ctypes.POINTER(Camera), u32, i32, ctypes.POINTER(P)]),
        "gsr_batch_destroy": (None, [P]),
        "gsr_batch_set_cameras": (i32, [P, ctypes.POINTER(Camera)]),
        "gsr_batch_step": (i32, [P, P, P, P, f32, P, P, P, P, i32]),
        "gsr_batch_entries": (i32, [P, ctypes.POINTER(u64)]),
        "gsr_batch_configure": (i32, [P, u32, u32]),
        "gsr_batch_autotune": (i32, [P, P, P, P, f32, P, P, P, ctypes.POINTER(u32), ctypes.POINTER(u32)]),
        "gsr_batch_profile": (i32, [P, P, P, P, f32, P, P, P, P, P]),
        "gsr_batch_set_passes": (i32, [P, u32, u32]),
        "gsr_loss_grad_window": (i32, [P, ctypes.POINTER(Camera), P, P, f32, i32, P]),
        "gsr_batch_set_ssim_window": (i32, [P, i32]),
        "gsr_batch_tuning": (i32, [P, ctypes.POINTER(u32)]),
        "gsr_batch_passes": (i32, [P, ctypes.POINTER(u32), ctypes.POINTER(u32)]),
        "gsr_batch_set_emit_band": (i32, [P, u32]),
        "gsr_batch_set_split": (i32, [P, u32, u32, u32]),
        "gsr_set_culling": (i32, [P, i32]),
        "gsr_batch_step_stats": (i32, [P, P, P, P, f32, P, P, P, P, P, i32]),
        "gsr_batch_set_culling": (i32, [P, i32]),
        "gsr_batch_tile_lists": (i32, [P, u32, u32, P, u64, P, ctypes.POINTER(u64)]),
        "gsr_bin_sort_noalloc": (i32, [P, ctypes.POINTER(Camera), P, u32, P, u64, P, ctypes.POINTER(u64)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def camera(cam, count: int) -> Camera:
    """gsr_camera from a scenes.Camera (the 80-byte uniform of gpu.js:696-701)."""
    c = Camera()
    for k in range(12):
        c.view[k] = float(cam.view[k])
    c.fx, c.fy, c.cx, c.cy = float(cam.fx), float(cam.fy), float(cam.cx), float(cam.cy)
    c.width, c.height = float(cam.width), float(cam.height)
    c.count = count
    c.pixels = cam.width * cam.height
    return c


class Buffer:
    """A device buffer owned by a Context."""

    def __init__(self, ctx: "Context", ptr: int, nbytes: int):
        self.ctx, self.ptr, self.nbytes = ctx, ptr, nbytes

    @property
    def p(self):
        return ctypes.c_void_p(self.ptr)

    def free(self):
        if self.ptr:
            self.ctx._check(self.ctx.lib.gsr_free(self.ctx.h, self.p))
            self.ptr = 0


class Context:
    """One device context: stream, buffers and the operator entry points."""

    def __init__(self, device: int = 0):
        self.lib = load()
        h = P()
        rc = self.lib.gsr_create(device, ctypes.byref(h))
        if rc != GSR_OK:
            raise GsrError(rc, self.lib.gsr_last_error(None).decode())
        self.h = h

    def close(self):
        if self.h:
            self.lib.gsr_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc: int):
        if rc != GSR_OK:
            raise GsrError(rc, self.lib.gsr_last_error(self.h).decode())

    # ---- buffers
    def alloc(self, nbytes: int) -> Buffer:
        p = P()
        self._check(self.lib.gsr_alloc(self.h, nbytes, ctypes.byref(p)))
        return Buffer(self, p.value, max(16, nbytes))

    def upload(self, arr: np.ndarray, buf: Buffer | None = None) -> Buffer:
        arr = np.ascontiguousarray(arr)
        if buf is None:
            buf = self.alloc(arr.nbytes)
        if arr.nbytes:
            self._check(self.lib.gsr_upload(self.h, buf.p, arr.ctypes.data_as(P), arr.nbytes))
        return buf

    def download(self, buf: Buffer, dtype, shape) -> np.ndarray:
        out = np.empty(shape, dtype=dtype)
        if out.nbytes:
            self._check(self.lib.gsr_download(self.h, out.ctypes.data_as(P), buf.p, out.nbytes))
        return out

    def clear(self, buf: Buffer, nbytes: int | None = None, offset: int = 0):
        n = buf.nbytes - offset if nbytes is None else nbytes
        self._check(self.lib.gsr_clear(self.h, ctypes.c_void_p(buf.ptr + offset), n))

    def copy(self, dst: Buffer, src: Buffer, nbytes: int, dst_offset: int = 0, src_offset: int = 0):
        """Device-to-device copy queued on the context's stream."""
        self._check(self.lib.gsr_copy(self.h, ctypes.c_void_p(dst.ptr + dst_offset),
                                      ctypes.c_void_p(src.ptr + src_offset), nbytes))

    def sync(self):
        self._check(self.lib.gsr_synchronize(self.h))

    def set_culling(self, enabled: bool):
        """Hit-box culling of forward / backward (a verification knob: results do not change)."""
        self._check(self.lib.gsr_set_culling(self.h, 1 if enabled else 0))

    # ---- operators
    def project(self, cam, gs: Buffer, n: int, splats: Buffer, sh: Buffer | None = None, deg: int = 0):
        c = camera(cam, n)
        self._check(self.lib.gsr_project(self.h, ctypes.byref(c), gs.p, sh.p if sh else None, deg, n, splats.p))

    def bin_sort(self, cam, splats: Buffer, n: int, order: Buffer | None, ranges: Buffer,
                 noalloc: bool = False) -> int:
        """Returns the entry count; raises GsrError(GSR_E_CAPACITY) when order is too small.
        noalloc: gsr_bin_sort_noalloc (scratch in the library's static device array)."""
        c = camera(cam, n)
        e = ctypes.c_uint64(0)
        cap = 0 if order is None else order.nbytes // 4
        fn = self.lib.gsr_bin_sort_noalloc if noalloc else self.lib.gsr_bin_sort
        rc = fn(self.h, ctypes.byref(c), splats.p, n, order.p if order else None, cap, ranges.p, ctypes.byref(e))
        if rc == GSR_E_CAPACITY:
            err = GsrError(rc, self.lib.gsr_last_error(self.h).decode())
            err.entries = e.value
            raise err
        self._check(rc)
        return e.value

    def forward(self, cam, splats: Buffer, order: Buffer, ranges: Buffer, image: Buffer, state: Buffer | None = None):
        c = camera(cam, 0)
        self._check(self.lib.gsr_forward(self.h, ctypes.byref(c), splats.p, order.p, ranges.p, image.p,
                                         state.p if state else None))

    def loss_grad(self, cam, image: Buffer, target: Buffer, lam: float, out: Buffer, window: int = SSIM_BOX7):
        c = camera(cam, 0)
        if window == SSIM_BOX7:
            self._check(self.lib.gsr_loss_grad(self.h, ctypes.byref(c), image.p, target.p, lam, out.p))
        else:
            self._check(self.lib.gsr_loss_grad_window(self.h, ctypes.byref(c), image.p, target.p, lam, window,
                                                      out.p))

    def backward(self, cam, splats: Buffer, order: Buffer, ranges: Buffer, image: Buffer, dl: Buffer, grads: Buffer):
        c = camera(cam, 0)
        self._check(self.lib.gsr_backward(self.h, ctypes.byref(c), splats.p, order.p, ranges.p, image.p, dl.p,
                                          grads.p))

    def shape(self, cam, gs: Buffer, n: int, grads: Buffer, shape: Buffer, sh: Buffer | None = None, deg: int = 0,
              sh_grad: Buffer | None = None):
        c = camera(cam, n)
        self._check(self.lib.gsr_shape(self.h, ctypes.byref(c), gs.p, sh.p if sh else None, deg, n, grads.p, shape.p,
                                       sh_grad.p if sh_grad else None))

    def adam(self, t: float, n: int, lr: dict, gs: Buffer, grads: Buffer, shape: Buffer, moments: Buffer):
        st = Step(t, n, lr["color"], lr["opacity"], lr["mean"], lr["scale"], lr["rotation"], 0.0)
        self._check(self.lib.gsr_adam(self.h, ctypes.byref(st), gs.p, grads.p, shape.p, moments.p))

    def adam_sh(self, t: float, n: int, lr: dict, gs: Buffer, sh: Buffer, grads: Buffer, shape: Buffer,
                sh_grad: Buffer, moments: Buffer):
        """Adam over mean, log-scale, rotation, opacity and the 48 SH coefficients;
        lr["color"] is the SH DC rate and lr["sh_rest"] that of bands 1-3."""
        st = Step(t, n, lr["color"], lr["opacity"], lr["mean"], lr["scale"], lr["rotation"], 0.0)
        self._check(self.lib.gsr_adam_sh(self.h, ctypes.byref(st), lr["sh_rest"], gs.p, sh.p, grads.p, shape.p,
                                         sh_grad.p, moments.p))

    def stats(self, n: int, grads: Buffer, splats: Buffer, stats: Buffer):
        st = Step(0.0, n, 0, 0, 0, 0, 0, 0)
        self._check(self.lib.gsr_stats(self.h, ctypes.byref(st), grads.p, splats.p, stats.p))


class Batch:
    """Batched multi-view forward + backward (gsr_batch_*)."""

    def __init__(self, ctx: Context, cams, n: int, deg: int):
        self.ctx, self.lib = ctx, ctx.lib
        self.V, self.n, self.deg = len(cams), n, deg
        arr = (Camera * len(cams))(*[camera(c, n) for c in cams])
        self._tiles = [((c.width + 15) // 16) * ((c.height + 15) // 16) for c in cams]
        h = P()
        ctx._check(self.lib.gsr_batch_create(ctx.h, len(cams), arr, n, deg, ctypes.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            self.lib.gsr_batch_destroy(self.h)
            self.h = None

    def set_cameras(self, cams):
        arr = (Camera * len(cams))(*[camera(c, self.n) for c in cams])
        self.ctx._check(self.lib.gsr_batch_set_cameras(self.h, arr))

    def step(self, gs: Buffer, sh: Buffer | None, targets: Buffer, lam: float, grads: Buffer, shape: Buffer,
             sh_grad: Buffer | None = None, images: Buffer | None = None, graph: bool = True,
             stats: Buffer | None = None):
        """One forward + backward of every view; with ``stats`` ([n][2] floats)
        also adds each view's STATS (|(du, dv)|, 1) for validly projected Gaussians."""
        if stats is None:
            self.ctx._check(self.lib.gsr_batch_step(self.h, gs.p, sh.p if sh else None, targets.p, lam, grads.p,
                                                    shape.p, sh_grad.p if sh_grad else None,
                                                    images.p if images else None, 1 if graph else 0))
        else:
            self.ctx._check(self.lib.gsr_batch_step_stats(self.h, gs.p, sh.p if sh else None, targets.p, lam,
                                                          grads.p, shape.p, sh_grad.p if sh_grad else None, stats.p,
                                                          images.p if images else None, 1 if graph else 0))

    def set_culling(self, enabled: bool):
        """Hit-box binning and culling (a verification knob: results do not change)."""
        self.ctx._check(self.lib.gsr_batch_set_culling(self.h, 1 if enabled else 0))

    def tile_lists(self, view: int, pass_: int = 0):
        """The last step's (order, ranges) of one view and depth slice, as host arrays."""
        e = ctypes.c_uint64(0)
        rb = self.ctx.alloc(8 * self._tiles[view])
        rc = self.lib.gsr_batch_tile_lists(self.h, view, pass_, None, 0, rb.p, ctypes.byref(e))
        if rc not in (GSR_OK, GSR_E_CAPACITY):
            rb.free()
            self.ctx._check(rc)
        ob = self.ctx.alloc(max(1, e.value) * 4)
        self.ctx._check(self.lib.gsr_batch_tile_lists(self.h, view, pass_, ob.p, e.value, rb.p, ctypes.byref(e)))
        order = self.ctx.download(ob, np.uint32, (e.value,))
        ranges = self.ctx.download(rb, np.uint32, (2 * self._tiles[view],))
        ob.free()
        rb.free()
        return order, ranges

    def entries(self) -> np.ndarray:
        out = (ctypes.c_uint64 * self.V)()
        self.ctx._check(self.lib.gsr_batch_entries(self.h, out))
        return np.array(out[:], dtype=np.uint64)

    def set_passes(self, passes: int, growth: int = 2):
        self.ctx._check(self.lib.gsr_batch_set_passes(self.h, passes, growth))

    def set_split(self, forward_first: int, forward_later: int, backward: int):
        """Split thresholds (0 = never): the forward's first slice (by the previous
        step's walk), its later slices (by list length) and the backward."""
        self.ctx._check(self.lib.gsr_batch_set_split(self.h, forward_first, forward_later, backward))

    def tuning(self) -> tuple:
        """(tile_batch, chunk, split_first, split_later, split_backward, emit_band,
        depth-slice passes, growth) in effect."""
        out = (ctypes.c_uint32 * 6)()
        self.ctx._check(self.lib.gsr_batch_tuning(self.h, out))
        passes, growth = ctypes.c_uint32(), ctypes.c_uint32()
        self.ctx._check(self.lib.gsr_batch_passes(self.h, ctypes.byref(passes), ctypes.byref(growth)))
        return tuple(int(x) for x in out) + (passes.value, growth.value)

    def set_emit_band(self, tiles: int):
        self.ctx._check(self.lib.gsr_batch_set_emit_band(self.h, tiles))

    def set_tuning(self, tuning):
        """Sets what tuning() returned (the emit band may be left off: 128; the
        depth slices too: the batch's own)."""
        tb, ch, sf, sl, sb = tuning[:5]
        if len(tuning) > 6:
            self.set_passes(tuning[6], tuning[7])
        self.configure(tb, ch)
        self.set_split(sf, sl, sb)
        self.set_emit_band(tuning[5] if len(tuning) > 5 else 128)

    def set_ssim_window(self, window: int):
        self.ctx._check(self.lib.gsr_batch_set_ssim_window(self.h, window))

    def configure(self, tile_batch: int = 0, chunk: int = 0):
        self.ctx._check(self.lib.gsr_batch_configure(self.h, tile_batch, chunk))

    def profile(self, gs, sh, targets, lam, grads, shape, sh_grad=None):
        """One eager step timed per stage with HIP events: (ms per stage dict,
        counters [V, 4] of entries E, saturated-prefix entries F, touched
        Gaussians G, walked entries W)."""
        ms = (ctypes.c_float * len(STAGES))()
        cnt = (ctypes.c_uint64 * (4 * self.V))()
        self.ctx._check(self.lib.gsr_batch_profile(self.h, gs.p, sh.p if sh else None, targets.p, lam, grads.p,
                                                   shape.p, sh_grad.p if sh_grad else None, ms, cnt))
        return dict(zip(STAGES, [float(x) for x in ms])), np.array(cnt[:], dtype=np.uint64).reshape(self.V, 4)

    def autotune(self, gs, sh, targets, lam, grads, shape, sh_grad=None):
        tb, ch = ctypes.c_uint32(0), ctypes.c_uint32(0)
        self.ctx._check(self.lib.gsr_batch_autotune(self.h, gs.p, sh.p if sh else None, targets.p, lam, grads.p,
                                                    shape.p, sh_grad.p if sh_grad else None, ctypes.byref(tb),
                                                    ctypes.byref(ch)))
        return tb.value, ch.value
