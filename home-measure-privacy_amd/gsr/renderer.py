"""Host-side mirror of createRenderer (trainer/gpu.js:704-980) over libgsr.

Same entry points and argument meaning as the reference's renderer object:
  render(gaussians, cam)                      gpu.js:815-820
  gradients(gaussians, cam, target, ssim=0)   gpu.js:826-839
  lossGradientOf(image, target, w, h, lam)    gpu.js:962-977
  train(gaussians, views, steps, lr)          gpu.js:846-878
Gaussians are packed float32 [n, 16] records (packGaussians, gpu.js:688-694), with
optional SH coefficients [n, 16, 3]; cameras are gsr.scenes.Camera objects.
Errors raise lib.GsrError, as the reference throws.
"""
from __future__ import annotations

import numpy as np

from . import lib

DEFAULT_TRAIN_LR = {"color": 0.05, "opacity": 0.05, "mean": 0.002, "scale": 0.01, "rotation": 0.005}  # gpu.js:846


class Frame:
    """Device buffers of one forward pass (what forward() returns, gpu.js:783-812)."""

    def __init__(self, ctx: lib.Context, cam, n: int):
        self.ctx, self.cam, self.n = ctx, cam, n
        self.splats = ctx.alloc(max(1, n) * 48)
        self.ranges = ctx.alloc(cam.tiles * 8)
        self.image = ctx.alloc(cam.pixels * 12)
        self.order = None
        self.entries = 0

    def free(self):
        for b in (self.splats, self.ranges, self.image, self.order):
            if b is not None:
                b.free()


class Renderer:
    def __init__(self, device: int = 0):
        self.ctx = lib.Context(device)

    def close(self):
        self.ctx.close()

    # ---- forward (gpu.js:783-812)
    def forward(self, gs_buf, n: int, cam, sh_buf=None, deg: int = 0, state: bool = False) -> Frame:
        f = Frame(self.ctx, cam, n)
        self.ctx.project(cam, gs_buf, n, f.splats, sh_buf, deg)
        try:
            f.entries = self.ctx.bin_sort(cam, f.splats, n, None, f.ranges)
        except lib.GsrError as e:
            if e.code != lib.GSR_E_CAPACITY:
                raise
            f.entries = e.entries
        f.order = self.ctx.alloc(max(1, f.entries) * 4)
        if f.entries:
            self.ctx.bin_sort(cam, f.splats, n, f.order, f.ranges)
        f.state = self.ctx.alloc(cam.pixels * 8) if state else None
        self.ctx.forward(cam, f.splats, f.order, f.ranges, f.image, f.state)
        return f

    def render(self, gaussians: np.ndarray, cam, sh: np.ndarray | None = None, deg: int = 0) -> np.ndarray:
        gs = self.ctx.upload(np.ascontiguousarray(gaussians, dtype=np.float32))
        shb = self.ctx.upload(np.ascontiguousarray(sh, dtype=np.float32)) if deg > 0 else None
        f = self.forward(gs, gaussians.shape[0], cam, shb, deg)
        img = self.ctx.download(f.image, np.float32, (cam.height, cam.width, 3))
        f.free()
        gs.free()
        if shb:
            shb.free()
        return img

    def gradients(self, gaussians: np.ndarray, cam, target: np.ndarray, ssim: float = 0.0,
                  sh: np.ndarray | None = None, deg: int = 0, dl: np.ndarray | None = None) -> dict:
        """image, grads [n,16], shape [n,10] (and sh_grad [n,16,3] for deg > 0).
        ``dl`` replaces the loss gradient (dLoss/dPixel) when given."""
        ctx = self.ctx
        n = gaussians.shape[0]
        gs = ctx.upload(np.ascontiguousarray(gaussians, dtype=np.float32))
        shb = ctx.upload(np.ascontiguousarray(sh, dtype=np.float32)) if deg > 0 else None
        f = self.forward(gs, n, cam, shb, deg)
        lg = ctx.alloc(cam.pixels * 12)
        if dl is None:
            tb = ctx.upload(np.ascontiguousarray(target, dtype=np.float32))
            ctx.loss_grad(cam, f.image, tb, ssim, lg)
            tb.free()
        else:
            ctx.upload(np.ascontiguousarray(dl, dtype=np.float32), lg)
        grads = ctx.alloc(max(1, n) * 64)
        ctx.backward(cam, f.splats, f.order, f.ranges, f.image, lg, grads)
        shape = ctx.alloc(max(1, n) * 40)
        shg = ctx.alloc(max(1, n) * 192) if deg > 0 else None
        ctx.shape(cam, gs, n, grads, shape, shb, deg, shg)
        out = {
            "image": ctx.download(f.image, np.float32, (cam.height, cam.width, 3)),
            "grads": ctx.download(grads, np.float32, (n, 16)),
            "shape": ctx.download(shape, np.float32, (n, 10)),
            "entries": f.entries,
        }
        if shg is not None:
            out["sh_grad"] = ctx.download(shg, np.float32, (n, 16, 3))
            shg.free()
        for b in (lg, grads, shape, gs, shb):
            if b is not None:
                b.free()
        f.free()
        return out

    def lossGradientOf(self, image: np.ndarray, target: np.ndarray, width: int, height: int, lam: float) -> np.ndarray:
        from .scenes import Camera

        cam = Camera(np.zeros(12), 0.0, 0.0, 0.0, 0.0, width, height)
        ib = self.ctx.upload(np.ascontiguousarray(image, dtype=np.float32))
        tb = self.ctx.upload(np.ascontiguousarray(target, dtype=np.float32))
        out = self.ctx.alloc(width * height * 12)
        self.ctx.loss_grad(cam, ib, tb, lam, out)
        g = self.ctx.download(out, np.float32, (width * height * 3,))
        for b in (ib, tb, out):
            b.free()
        return g

    def train(self, gaussians: np.ndarray, views: list, steps: int, lr: dict | None = None):
        """train (gpu.js:846-878): view (t-1) % len per step, L1, Adam.  views are
        (camera, target[H, W, 3]) pairs.  Returns (gaussians, losses)."""
        lr = dict(DEFAULT_TRAIN_LR if lr is None else lr)
        ctx = self.ctx
        n = gaussians.shape[0]
        gs = ctx.upload(np.ascontiguousarray(gaussians, dtype=np.float32))
        moments = ctx.alloc(max(1, n) * 28 * 4)
        targets = [ctx.upload(np.ascontiguousarray(t, dtype=np.float32)) for _, t in views]
        losses = []
        for t in range(1, steps + 1):
            cam, target = views[(t - 1) % len(views)]
            f = self.forward(gs, n, cam)
            lg = ctx.alloc(cam.pixels * 12)
            ctx.loss_grad(cam, f.image, targets[(t - 1) % len(views)], 0.0, lg)
            grads = ctx.alloc(max(1, n) * 64)
            ctx.backward(cam, f.splats, f.order, f.ranges, f.image, lg, grads)
            shape = ctx.alloc(max(1, n) * 40)
            ctx.shape(cam, gs, n, grads, shape)
            ctx.adam(float(t), n, lr, gs, grads, shape, moments)
            image = ctx.download(f.image, np.float32, (cam.height, cam.width, 3))
            losses.append(float(np.abs(image - target).mean()))
            for b in (lg, grads, shape):
                b.free()
            f.free()
        out = ctx.download(gs, np.float32, (n, 16))
        for b in [gs, moments, *targets]:
            b.free()
        return out, losses


def createRenderer(device: int = 0) -> Renderer:
    """createRenderer (gpu.js:704): the renderer object of the reference."""
    return Renderer(device)
