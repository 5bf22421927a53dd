"""Training-step pieces on the GPU: STATS (trainer/gpu.js:483-497) single-view
and batched, and the batched training loop (batched step + SH3 Adam on the
summed gradients) of the clustered 64-view workload at reduced size, replayed
view by view through the single-view C ABI."""
import numpy as np
import pytest

import oracle as O
from conftest import rel_l2
from gsr import lib, scenes

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)


def test_stats_matches_oracle(ctx):
    """gsr_stats against the oracle's STATS over three accumulations, with valid
    and invalid splats and zero gradients."""
    rng = np.random.default_rng(8)
    n = 70000
    grads = (rng.standard_normal((n, 16)) * 1e-3).astype(np.float32)
    grads[::7, 4:6] = 0.0
    splats = rng.uniform(size=(n, 12)).astype(np.float32)
    splats[:, 11] = (rng.uniform(size=n) > 0.3).astype(np.float32)
    st0 = np.abs(rng.standard_normal((n, 2))).astype(np.float32)
    gb, sb, tb = ctx.upload(grads), ctx.upload(splats), ctx.upload(st0)
    ref = st0
    for _ in range(3):
        ctx.stats(n, gb, sb, tb)
        ref = O.stats(grads, splats, ref, np.float32)
    got = ctx.download(tb, np.float32, (n, 2))
    np.testing.assert_array_equal(got[:, 1], ref[:, 1])
    np.testing.assert_allclose(got[:, 0], ref[:, 0], rtol=2e-7, atol=0)
    for b in (gb, sb, tb):
        b.free()


def shrunk(sc, scale, sizes=None):
    """The scene's cameras scaled down by `scale` (intrinsics with them), or to
    the given (w, h) sizes in turn; new seeded targets."""
    cams = []
    for v, c in enumerate(sc.cams):
        w, h = sizes[v % len(sizes)] if sizes else (int(c.width / scale), int(c.height / scale))
        s = w / c.width
        cams.append(scenes.Camera(c.view, c.fx * s, c.fy * s, w / 2.0, h / 2.0, w, h))
    sc.cams = cams
    rng = scenes.Rng(1234)
    sc.targets = [rng.uniform(c.pixels * 3).astype(np.float32).reshape(c.height, c.width, 3) for c in cams]
    return sc


class PerView:
    """The single-view pipeline of the C ABI for one set of parameters: per view
    project, bin_sort, forward, L1 loss gradient, backward, shape (SH3) and STATS;
    gradients summed over views on the host."""

    def __init__(self, ctx, n):
        self.ctx, self.n = ctx, n
        self.splats, self.grads = ctx.alloc(n * 48), ctx.alloc(n * 64)
        self.shape, self.shg = ctx.alloc(n * 40), ctx.alloc(n * 192)
        self.order, self.cap = None, 0

    def run(self, sc, gsb, shb, stats=None, lam=0.0):
        ctx, n = self.ctx, self.n
        g_sum, s_sum, sh_sum = np.zeros((n, 16)), np.zeros((n, 10)), np.zeros((n, 16, 3))
        for cam, tgt in zip(sc.cams, sc.targets):
            ranges, image, dl = ctx.alloc(cam.tiles * 8), ctx.alloc(cam.pixels * 12), ctx.alloc(cam.pixels * 12)
            tb = ctx.upload(tgt)
            ctx.project(cam, gsb, n, self.splats, shb, 3)
            try:
                e = ctx.bin_sort(cam, self.splats, n, None, ranges)
            except lib.GsrError as err:
                assert err.code == lib.GSR_E_CAPACITY
                e = err.entries
            if e > self.cap:
                if self.order is not None:
                    self.order.free()
                self.cap = e
                self.order = ctx.alloc(max(1, e) * 4)
            ctx.bin_sort(cam, self.splats, n, self.order, ranges)
            ctx.forward(cam, self.splats, self.order, ranges, image)
            ctx.loss_grad(cam, image, tb, lam, dl)
            ctx.clear(self.grads)
            ctx.backward(cam, self.splats, self.order, ranges, image, dl, self.grads)
            ctx.shape(cam, gsb, n, self.grads, self.shape, shb, 3, self.shg)
            if stats is not None:
                ctx.stats(n, self.grads, self.splats, stats)
            g_sum += ctx.download(self.grads, np.float32, (n, 16))
            s_sum += ctx.download(self.shape, np.float32, (n, 10))
            sh_sum += ctx.download(self.shg, np.float32, (n, 16, 3))
            for b in (ranges, image, dl, tb):
                b.free()
        return g_sum, s_sum, sh_sum

    def free(self):
        for b in (self.splats, self.grads, self.shape, self.shg, self.order):
            if b is not None:
                b.free()


def test_batch_stats_match_per_view_stats(ctx):
    """Batched STATS over 5 uneven views and two steps equals one gsr_stats pass
    per view after each view's single-view backward: view counts exactly, the
    summed |(du, dv)| to float-atomic summation order."""
    sc = shrunk(scenes.config(3, n=20000, views=5, seed=3), 1.0,
                sizes=[(96, 64), (160, 120), (64, 48), (200, 100), (33, 17)])
    n = sc.n
    gsb, shb = ctx.upload(sc.gs), ctx.upload(sc.sh)
    tgt = ctx.upload(np.concatenate([t.ravel() for t in sc.targets]).astype(np.float32))
    grads, shape, shg = ctx.alloc(n * 64), ctx.alloc(n * 40), ctx.alloc(n * 192)
    st_b = ctx.alloc(n * 8)
    b = lib.Batch(ctx, sc.cams, n, 3)
    for graph in (False, True):
        b.step(gsb, shb, tgt, 0.0, grads, shape, shg, None, graph, stats=st_b)
    got = ctx.download(st_b, np.float32, (n, 2))
    st_v = ctx.alloc(n * 8)
    pv = PerView(ctx, n)
    for _ in range(2):
        pv.run(sc, gsb, shb, st_v)
    ref = ctx.download(st_v, np.float32, (n, 2))
    log(f"  stats: views counted {int(got[:, 1].sum())}, gradient norms rel L2 {rel_l2(got[:, 0], ref[:, 0]):.2e}")
    np.testing.assert_array_equal(got[:, 1], ref[:, 1])
    assert got[:, 1].sum() > n
    assert rel_l2(got[:, 0], ref[:, 0]) < 1e-5
    b.close()
    pv.free()
    for x in (gsb, shb, tgt, grads, shape, shg, st_b, st_v):
        x.free()


def test_batch_stats_counted_once_after_overflow(ctx):
    """A step whose entry storage overflows is run again with more storage: its
    statistics are added once, not twice."""
    sc = shrunk(scenes.config(3, n=3000, views=2, seed=3), 1.0, sizes=[(96, 64), (160, 120)])
    sc.gs[:, 4:7] += 2.5  # large Gaussians: far more entries than the first guess of storage
    n = sc.n
    gsb, shb = ctx.upload(sc.gs), ctx.upload(sc.sh)
    tgt = ctx.upload(np.concatenate([t.ravel() for t in sc.targets]).astype(np.float32))
    grads, shape, shg, st = ctx.alloc(n * 64), ctx.alloc(n * 40), ctx.alloc(n * 192), ctx.alloc(n * 8)
    b = lib.Batch(ctx, sc.cams, n, 3)
    b.step(gsb, shb, tgt, 0.0, grads, shape, shg, None, True, stats=st)
    assert int(b.entries().sum()) > 16 * n * 2  # the first guess (16 per Gaussian and view) overflowed
    got = ctx.download(st, np.float32, (n, 2))
    valid = np.zeros(n)
    sp = ctx.alloc(n * 48)
    for cam in sc.cams:
        ctx.project(cam, gsb, n, sp, shb, 3)
        valid += ctx.download(sp, np.float32, (n, 12))[:, 11] > 0.5
    np.testing.assert_array_equal(got[:, 1], valid)
    b.close()
    for x in (gsb, shb, tgt, grads, shape, shg, st, sp):
        x.free()


def test_training_loop_matches_per_view_replay(ctx):
    """The clustered 64-view training workload at reduced size (20k Gaussians, 6
    views at 1/8 of 1600x1066): four steps of batched step + STATS + SH3 Adam on
    the summed gradients, against the same loop run view by view through the
    single-view C ABI (per-view gradients summed on the host, then the same
    Adam).  Gradients agree each step, the parameters after four steps and the
    accumulated statistics too."""
    sc = shrunk(scenes.config(4, n=20000, views=6), 8.0)
    n = sc.n
    log(f"training replay: {n} Gaussians, views {[(c.width, c.height) for c in sc.cams]}")
    lr = {"color": 5e-3, "sh_rest": 2.5e-4, "mean": 1.6e-4, "opacity": 0.05, "scale": 0.005, "rotation": 0.001}
    # batched run
    gs_b, sh_b = ctx.upload(sc.gs), ctx.upload(sc.sh)
    tgt = ctx.upload(np.concatenate([t.ravel() for t in sc.targets]).astype(np.float32))
    grads, shape, shg = ctx.alloc(n * 64), ctx.alloc(n * 40), ctx.alloc(n * 192)
    mom_b, st_b = ctx.alloc(n * 480), ctx.alloc(n * 8)
    batch = lib.Batch(ctx, sc.cams, n, 3)
    # replay run
    gs_r, sh_r = ctx.upload(sc.gs), ctx.upload(sc.sh)
    mom_r, st_r = ctx.alloc(n * 480), ctx.alloc(n * 8)
    gr_u, sr_u, shr_u = ctx.alloc(n * 64), ctx.alloc(n * 40), ctx.alloc(n * 192)
    pv = PerView(ctx, n)
    for t in range(1, 5):
        batch.step(gs_b, sh_b, tgt, 0.0, grads, shape, shg, None, True, stats=st_b)
        gb_ = ctx.download(grads, np.float32, (n, 16))
        sb_ = ctx.download(shape, np.float32, (n, 10))
        shb_ = ctx.download(shg, np.float32, (n, 16, 3))
        ctx.adam_sh(float(t), n, lr, gs_b, sh_b, grads, shape, shg, mom_b)
        g_sum, s_sum, sh_sum = pv.run(sc, gs_r, sh_r, st_r)
        errs = (rel_l2(gb_[:, 3], g_sum[:, 3]), rel_l2(sb_, s_sum), rel_l2(shb_, sh_sum))
        log(f"  step {t}: gradient rel L2 opacity {errs[0]:.2e} shape {errs[1]:.2e} sh {errs[2]:.2e}")
        assert max(errs) < 1e-4
        ctx.upload(g_sum.astype(np.float32), gr_u)
        ctx.upload(s_sum.astype(np.float32), sr_u)
        ctx.upload(sh_sum.astype(np.float32), shr_u)
        ctx.adam_sh(float(t), n, lr, gs_r, sh_r, gr_u, sr_u, shr_u, mom_r)
    p_b = ctx.download(gs_b, np.float32, (n, 16)).astype(np.float64)
    p_r = ctx.download(gs_r, np.float32, (n, 16)).astype(np.float64)
    q_b = ctx.download(sh_b, np.float32, (n, 16, 3)).astype(np.float64)
    q_r = ctx.download(sh_r, np.float32, (n, 16, 3)).astype(np.float64)
    moved = rel_l2(p_b - sc.gs, p_r - sc.gs), rel_l2(q_b - sc.sh, q_r - sc.sh)
    log(f"  parameter updates after 4 steps, rel L2: gaussians {moved[0]:.2e} sh {moved[1]:.2e}")
    assert np.abs(p_b - sc.gs).max() > 0
    assert max(moved) < 1e-3
    s1, s2 = ctx.download(st_b, np.float32, (n, 2)), ctx.download(st_r, np.float32, (n, 2))
    np.testing.assert_array_equal(s1[:, 1], s2[:, 1])
    assert rel_l2(s1[:, 0], s2[:, 0]) < 1e-4
    batch.close()
    pv.free()
    for x in (gs_b, sh_b, tgt, grads, shape, shg, mom_b, st_b, gs_r, sh_r, mom_r, st_r, gr_u, sr_u, shr_u):
        x.free()
