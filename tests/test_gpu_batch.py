"""Batched multi-view step against the single-view path of the same library and
against the FP64 oracle summed over views."""
import numpy as np
import pytest

import oracle as O
from conftest import rel_l2
from gsr import lib, scenes
from gsr.renderer import Renderer

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)


def small_batch_scene(n=20000, views=5, seed=0, sizes=None):
    sc = scenes.config(3, n=n, views=views, seed=seed)
    # shrink the views so the FP64 oracle stays fast; keep them uneven
    sizes = sizes or [(96, 64), (160, 120), (64, 48), (200, 100), (33, 17), (128, 128), (48, 80)]
    cams = []
    for v, c in enumerate(sc.cams):
        w, h = sizes[v % len(sizes)]
        s = w / c.width
        cams.append(scenes.Camera(c.view, c.fx * s, c.fx * s, w / 2.0, h / 2.0, w, h))
    sc.cams = cams
    rng = scenes.Rng(99 + seed)
    sc.targets = [rng.uniform(c.pixels * 3).astype(np.float32).reshape(c.height, c.width, 3) for c in cams]
    return sc


def run_batch(ctx, sc, graph=True, tile_batch=0, chunk=0, steps=1, images=False, passes=None, growth=2, split=None,
              lam=0.0, emit_band=None):
    n = sc.n
    gsb, shb = ctx.upload(sc.gs), ctx.upload(sc.sh)
    tgt = ctx.upload(np.concatenate([t.ravel() for t in sc.targets]).astype(np.float32))
    grads, shape, shg = ctx.alloc(n * 64), ctx.alloc(n * 40), ctx.alloc(n * 192)
    total_px = sum(c.pixels for c in sc.cams)
    img = ctx.alloc(total_px * 12) if images else None
    b = lib.Batch(ctx, sc.cams, n, 3)
    b.configure(tile_batch, chunk)
    if passes is not None:
        b.set_passes(passes, growth)
    if split is not None:
        b.set_split(*split)
    if emit_band is not None:
        b.set_emit_band(emit_band)
    outs = []
    for _ in range(steps):
        b.step(gsb, shb, tgt, lam, grads, shape, shg, img, graph)
        outs.append((ctx.download(grads, np.float32, (n, 16)), ctx.download(shape, np.float32, (n, 10)),
                     ctx.download(shg, np.float32, (n, 16, 3)),
                     ctx.download(img, np.float32, (total_px * 3,)) if images else None))
    entries = b.entries()
    b.close()
    for x in (gsb, shb, tgt, grads, shape, shg, img):
        if x is not None:
            x.free()
    return outs, entries


# lambda 0: the backward forms the L1 gradient itself; 0.2: the SSIM term goes
# through the loss kernel, as on the single-view path.
@pytest.mark.parametrize("lam", [0.0, 0.2])
def test_batch_matches_single_view_path(ctx, lam):
    sc = small_batch_scene()
    log(f"batch vs single: lambda={lam} n={sc.n} views={[(c.width, c.height) for c in sc.cams]}")
    outs, entries = run_batch(ctx, sc, graph=lam > 0, images=True, lam=lam)
    g_b, s_b, sh_b, img_b = outs[0]
    r = Renderer()
    g_sum = np.zeros((sc.n, 16))
    s_sum = np.zeros((sc.n, 10))
    sh_sum = np.zeros((sc.n, 16, 3))
    imgs, ents = [], []
    for cam, tgt in zip(sc.cams, sc.targets):
        out = r.gradients(sc.gs, cam, tgt, lam, sh=sc.sh, deg=3)
        g_sum += out["grads"]
        s_sum += out["shape"]
        sh_sum += out["sh_grad"]
        imgs.append(out["image"].ravel())
        ents.append(out["entries"])
    r.close()
    # The batched step bins each splat to the tiles of its hit box (alpha can
    # reach 1/255 there) inside the reference square, so its lists are subsets of
    # the reference lists; the entries left out cannot hit any pixel.
    ents = np.array(ents, dtype=np.uint64)
    assert (entries <= ents).all() and entries.sum() < ents.sum()
    log(f"  entries batched {int(entries.sum())} reference {int(ents.sum())}")
    # same projection, depth order and blend arithmetic: identical images
    np.testing.assert_array_equal(img_b, np.concatenate(imgs))
    # gradients differ only by float atomic summation order
    assert rel_l2(g_b[:, :4], g_sum[:, :4]) < 1e-5
    assert rel_l2(s_b, s_sum) < 1e-5
    assert rel_l2(sh_b, sh_sum) < 1e-5


def test_batch_extreme_view_shapes_match_single_view_path(ctx):
    """A 1x1 view, a 4100-px-wide strip (257 tiles: binning bands of one tile
    row wider than the 256-tile emit budget), a 2000-px-tall strip and a view
    one pixel past a tile edge, batched with depth slices and tile splitting:
    the same images as view-by-view rendering, gradients to summation order."""
    sizes = [(1, 1), (4100, 20), (20, 2000), (17, 33)]
    sc = small_batch_scene(n=6000, views=4, seed=11, sizes=sizes)
    log(f"batch extreme shapes: {[(c.width, c.height) for c in sc.cams]}")
    outs, entries = run_batch(ctx, sc, graph=True, images=True, passes=2, growth=4, split=(64, 64, 64), steps=2)
    r = Renderer()
    g_sum = np.zeros((sc.n, 16))
    s_sum = np.zeros((sc.n, 10))
    imgs = []
    for cam, tgt in zip(sc.cams, sc.targets):
        out = r.gradients(sc.gs, cam, tgt, 0.0, sh=sc.sh, deg=3)
        g_sum += out["grads"]
        s_sum += out["shape"]
        imgs.append(out["image"].ravel())
    r.close()
    assert entries.sum() > 0
    for g_b, s_b, _, img_b in outs:
        np.testing.assert_array_equal(img_b, np.concatenate(imgs))
        assert rel_l2(g_b[:, :4], g_sum[:, :4]) < 1e-5
        assert rel_l2(s_b, s_sum) < 1e-5


@pytest.mark.parametrize("n", [0, 1, 63])
def test_batch_tiny_gaussian_counts(ctx, n):
    """No Gaussian, one, and fewer than a wave: the step runs (graph and eager);
    with Gaussians it matches view-by-view rendering."""
    sc = small_batch_scene(n=max(n, 1), views=3, seed=13)
    sc.gs, sc.sh = sc.gs[:n], sc.sh[:n]
    log(f"batch tiny n={n}")
    for graph in (False, True):
        outs, entries = run_batch(ctx, sc, graph=graph, images=True, steps=2)
        for g_b, s_b, _, img_b in outs:
            assert np.isfinite(img_b).all()
            if n == 0:
                assert not img_b.any() and int(entries.sum()) == 0
    if n == 0:
        return
    r = Renderer()
    imgs, g_sum = [], np.zeros((n, 16))
    for cam, tgt in zip(sc.cams, sc.targets):
        out = r.gradients(sc.gs, cam, tgt, 0.0, sh=sc.sh, deg=3)
        imgs.append(out["image"].ravel())
        g_sum += out["grads"]
    r.close()
    np.testing.assert_array_equal(outs[-1][3], np.concatenate(imgs))
    assert np.allclose(outs[-1][0][:, :4], g_sum[:, :4], rtol=1e-5, atol=1e-12)


def test_batch_tuning_reads_back_the_knobs(ctx):
    """gsr_batch_tuning / gsr_batch_passes return what configure, set_split,
    set_emit_band and set_passes set (0 keeps a knob); the depth slices of a
    batch under 65536 Gaussians default to one pass."""
    sc = small_batch_scene(n=3000, views=2, seed=4)
    b = lib.Batch(ctx, sc.cams, sc.n, 3)
    b.set_tuning((4, 2048, 100, 200, 300, 128))
    assert b.tuning() == (4, 2048, 100, 200, 300, 128, 1, 1)
    b.configure(0, 0)
    b.set_split(0, 0, 0)
    assert b.tuning() == (4, 2048, 0, 0, 0, 128, 1, 1)
    b.set_tuning((1, 1024, 0, 0, 1024))  # the emit band left off: 128
    assert b.tuning() == (1, 1024, 0, 0, 1024, 128, 1, 1)
    b.set_tuning((1, 1024, 0, 0, 1024, 256, 3, 4))
    assert b.tuning() == (1, 1024, 0, 0, 1024, 256, 3, 4)
    with pytest.raises(lib.GsrError):
        b.set_emit_band(32)
    with pytest.raises(lib.GsrError):
        b.set_passes(3, 64)  # growth above 16
    b.close()


def test_batch_graph_replay_and_knobs(ctx):
    """Graph replays (several steps) agree with the eager step; tile batch and
    chunk size do not change results beyond summation order."""
    sc = small_batch_scene(n=15000, views=4, seed=1)
    eager, e0 = run_batch(ctx, sc, graph=False)
    replay, e1 = run_batch(ctx, sc, graph=True, steps=3)
    tuned, e2 = run_batch(ctx, sc, graph=True, tile_batch=4, chunk=512, steps=2)
    banded, e3 = run_batch(ctx, sc, graph=True, emit_band=64, steps=2)  # 64-tile emit bands
    np.testing.assert_array_equal(e0, e1)
    np.testing.assert_array_equal(e0, e2)
    np.testing.assert_array_equal(e0, e3)
    for outs in (replay, tuned, banded):
        for g, s, sh, _ in outs:
            assert rel_l2(g[:, :4], eager[0][0][:, :4]) < 1e-5
            assert rel_l2(s, eager[0][1]) < 1e-5
            assert rel_l2(sh, eager[0][2]) < 1e-5


@pytest.mark.parametrize("passes,growth", [(2, 8), (3, 4), (5, 2), (8, 2)])
def test_depth_slice_passes_match_full_lists(ctx, passes, growth):
    """Binning in depth slices, each only for tiles still unsaturated, renders the
    same images bit for bit as the full tile lists, with the same gradients up to
    float atomic summation order; fewer entries are emitted."""
    sc = small_batch_scene(n=40000, views=4, seed=5)
    sc.gs[:, 7] += 2.0  # more opaque: tiles saturate early
    full, e_full = run_batch(ctx, sc, graph=False, images=True, passes=1)
    log(f"passes={passes} growth={growth}")
    sliced, e_sl = run_batch(ctx, sc, graph=True, images=True, passes=passes, growth=growth, steps=2)
    for g, s, sh, img in sliced:
        np.testing.assert_array_equal(img, full[0][3])
        assert rel_l2(g[:, :4], full[0][0][:, :4]) < 1e-5
        assert rel_l2(s, full[0][1]) < 1e-5
        assert rel_l2(sh, full[0][2]) < 1e-5
    assert e_sl.sum() < e_full.sum()
    log(f"  entries full {int(e_full.sum())} sliced {int(e_sl.sum())}")


@pytest.mark.parametrize("wide", [(4112, 64), (4096, 2048), (4112, 2176)])
def test_depth_slices_with_wide_views(ctx, wide):
    """Later depth slices are filtered to the Gaussians that reach an open tile
    through a 16-bit summed-area table of each view's open tiles in LDS, for
    batches whose views have at most 32768 tiles, and binned unfiltered beyond.
    A 4112 x 64 view (257 x 4 tiles: the 16-bit-per-coordinate rectangles, since
    257 columns do not fit the packed 8-bit form) is filtered, and so is a 4096 x
    2048 view (256 x 128 = 32768 tiles: the whole 64 KB of LDS for the table);
    with a 4112 x 2176 view (257 x 136 = 34952 tiles) the batch is not.  All
    render the full lists' images bit for bit."""
    sc = small_batch_scene(n=20000, views=3, seed=11, sizes=[wide, (96, 64), (160, 120)])
    sc.gs[:, 7] += 2.0
    log(f"views {[(c.width, c.height) for c in sc.cams]}")
    full, e_full = run_batch(ctx, sc, graph=False, images=True, passes=1)
    sliced, e_sl = run_batch(ctx, sc, graph=True, images=True, passes=3, growth=4, steps=2)
    for g, s, sh, img in sliced:
        np.testing.assert_array_equal(img, full[0][3])
        assert rel_l2(g[:, :4], full[0][0][:, :4]) < 1e-5
        assert rel_l2(s, full[0][1]) < 1e-5
        assert rel_l2(sh, full[0][2]) < 1e-5
    assert e_sl.sum() < e_full.sum()
    log(f"  entries full {int(e_full.sum())} sliced {int(e_sl.sum())}")


@pytest.mark.parametrize("split", [(64, 64, 64), (0, 0, 0), (0, 200, 0), (100, 0, 0)])
def test_tile_splitting_leaves_results_unchanged(ctx, split):
    """Blending long tiles as two half-tile items (in the forward's first depth
    slice by the previous step's walk, its later slices and the backward)
    renders the same images bit for bit, with the same gradients up to float
    atomic summation order.  Low thresholds split most tiles; the second step
    splits the first slice by the first step's walk."""
    sc = small_batch_scene(n=40000, views=4, seed=7)
    sc.gs[:, 7] += 1.0
    ref, _ = run_batch(ctx, sc, graph=False, images=True, passes=3, growth=4, split=(0, 0, 0))
    log(f"split forward/backward {split}")
    got, _ = run_batch(ctx, sc, graph=True, images=True, passes=3, growth=4, split=split, steps=2)
    for g, s, sh, img in got:
        np.testing.assert_array_equal(img, ref[0][3])
        assert rel_l2(g[:, :4], ref[0][0][:, :4]) < 1e-5
        assert rel_l2(s, ref[0][1]) < 1e-5
        assert rel_l2(sh, ref[0][2]) < 1e-5


def test_batch_matches_fp64_oracle(ctx):
    sc = small_batch_scene(n=6000, views=3, seed=2)
    outs, _ = run_batch(ctx, sc, graph=True)
    g_b, s_b, sh_b, _ = outs[0]
    g_sum = np.zeros((sc.n, 16))
    s_sum = np.zeros((sc.n, 10))
    sh_sum = np.zeros((sc.n, 16, 3))
    gs64, sh64 = sc.gs.astype(np.float64), sc.sh.astype(np.float64)
    for cam, tgt in zip(sc.cams, sc.targets):
        img = O.render(cam, gs64, sh64, 3)[3]
        dl = O.loss_grad(img, tgt.astype(np.float64), 0.0)
        _, grads, shape, shg = O.gradients(cam, gs64, dl, sh64, 3)
        g_sum += grads
        s_sum += shape
        sh_sum += shg
    # L1's sign(x - y) can flip on pixels where the float32 and float64 renders
    # straddle the target; the bar is the end-to-end 1e-4 of the differentiation check.
    for name, a, b in (("opacity", g_b[:, 3], g_sum[:, 3]), ("means", s_b[:, :3], s_sum[:, :3]),
                       ("log_scales", s_b[:, 3:6], s_sum[:, 3:6]), ("rotations", s_b[:, 6:], s_sum[:, 6:]),
                       ("sh", sh_b, sh_sum)):
        e = rel_l2(a, b)
        log(f"  batch vs fp64 {name}: {e:.2e}")
        assert e < 1e-4, name


def test_batch_entry_storage_grows(ctx):
    """The first guess of entry storage is exceeded by large footprints; the step
    detects the overflow, grows storage and reruns."""
    sc = small_batch_scene(n=3000, views=2, seed=3)
    sc.gs[:, 4:7] += 2.5  # much larger Gaussians: many entries per Gaussian
    outs, entries = run_batch(ctx, sc, graph=True, steps=2)
    assert int(entries.sum()) > 16 * sc.n * 2
    r = Renderer()
    for cam, e in zip(sc.cams, entries):
        out = r.render(sc.gs, cam, sc.sh, 3)
        assert out.shape == (cam.height, cam.width, 3)
    r.close()


def test_batch_wide_depth_range_sorts_with_four_passes(ctx):
    """The batched depth sort takes three passes while each view's depth keys'
    top 16 bits span at most 511 values (a depth range within a factor of about
    16).  A scene with Gaussians from depth 1.5 to about 100 exceeds that: the
    first step flags it, runs again with the four-pass sort, and renders as the
    single-view path (bit-identical images) with the same gradients."""
    sc = small_batch_scene(n=20000, views=3, seed=5)
    rng = np.random.default_rng(5)
    far = rng.choice(sc.n, 2000, replace=False)
    d = rng.standard_normal((far.size, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    sc.gs[far, 0:3] = (d * rng.uniform(60.0, 100.0, (far.size, 1))).astype(np.float32)
    sc.gs[far, 4:7] = np.log(1.5)  # a few pixels across at depth 80
    sc.gs[far, 7] = 3.0
    # The regime is exercised: some view's binned splats span more than 511
    # top-16-bit key values (the square rectangle of tileLists bounds the hit box).
    widest = 0
    for cam in sc.cams:
        sp = O.project(cam, sc.gs, sc.sh, 3, np.float32)
        u, v, r = (sp[:, k].astype(np.float64) for k in (0, 1, 6))
        on = (sp[:, 11] > 0.5) & (u + r > 0) & (u - r < cam.width) & (v + r > 0) & (v - r < cam.height)
        top = sp[on, 7].view(np.uint32) >> 16
        widest = max(widest, int(top.max()) - int(top.min()))
    log(f"  widest view: {widest} top-16-bit key values")
    assert widest > 511
    outs, entries = run_batch(ctx, sc, graph=True, steps=2, images=True)
    r = Renderer()
    imgs, g_sum = [], np.zeros((sc.n, 16))
    for cam, tgt in zip(sc.cams, sc.targets):
        out = r.gradients(sc.gs, cam, tgt, 0.0, sh=sc.sh, deg=3)
        imgs.append(out["image"].ravel())
        g_sum += out["grads"]
    r.close()
    for g_b, _, _, img_b in outs:
        np.testing.assert_array_equal(img_b, np.concatenate(imgs))
        assert rel_l2(g_b[:, :4], g_sum[:, :4]) < 1e-5
