"""gsr_bin_sort_noalloc, the tile lists of gsr_bin_sort with every scratch buffer
in the library's static device array (no allocation in the call), against
gsr_bin_sort on the same inputs: bit for bit, over calls of decreasing and
increasing size (the array keeps the last call's contents), and its errors."""
import numpy as np
import pytest

from gsr import lib, scenes

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)


def lists(ctx, cam, sp, n, noalloc):
    rb = ctx.alloc(cam.tiles * 8)
    try:
        e = ctx.bin_sort(cam, sp, n, None, rb, noalloc=noalloc)
    except lib.GsrError as err:
        assert err.code == lib.GSR_E_CAPACITY
        e = err.entries
    ob = ctx.alloc(max(1, e) * 4)
    assert ctx.bin_sort(cam, sp, n, ob, rb, noalloc=noalloc) == e
    order = ctx.download(ob, np.uint32, (e,))
    ranges = ctx.download(rb, np.uint32, (cam.tiles * 2,))
    ob.free()
    rb.free()
    return order, ranges


def test_bin_sort_noalloc_matches_bin_sort(ctx):
    # 300k SH3 at 1280x720, then a smaller input (20k at 640x480), then the large
    # one again: every call must equal gsr_bin_sort's lists.
    cases = [(1, None, None), (1, 20000, (640, 480)), (1, None, None), (0, None, None)]
    for idx, n, size in cases:
        sc = scenes.config(idx, n=n, with_targets=False)
        cam = sc.cams[0]
        if size:
            cam = scenes.look_at(3.0 * np.array([0.3, -0.4, 0.866]), [0, 0, 0], size[0], size[1], 60.0)
        gb = ctx.upload(sc.gs)
        shb = ctx.upload(sc.sh) if sc.sh_degree else None
        sp = ctx.alloc(sc.n * 48)
        ctx.project(cam, gb, sc.n, sp, shb, sc.sh_degree)
        log(f"noalloc: {sc.n} Gaussians at {cam.width}x{cam.height}")
        o_ref, r_ref = lists(ctx, cam, sp, sc.n, False)
        o_na, r_na = lists(ctx, cam, sp, sc.n, True)
        log(f"  entries {o_ref.size}")
        assert o_ref.size > 0
        np.testing.assert_array_equal(r_na, r_ref)
        np.testing.assert_array_equal(o_na, o_ref)
        for b in (gb, shb, sp):
            if b is not None:
                b.free()


def test_bin_sort_noalloc_errors(ctx):
    cam = scenes.look_at([0.0, 0.0, -3.0], [0, 0, 0], 1920, 1080, 60.0)
    # Above the static scratch (2M Gaussians on a 1080p view need about 85 MB):
    # refused before any launch.  The buffers are real and zero-filled (every
    # splat invalid), so the inputs are in bounds either way.
    n = 2_000_000
    sp, rb, ob = ctx.alloc(n * 48), ctx.alloc(cam.tiles * 8), ctx.alloc(1024)
    with pytest.raises(lib.GsrError) as ei:
        ctx.bin_sort(cam, sp, n, ob, rb, noalloc=True)
    assert ei.value.code == lib.GSR_E_CAPACITY and "scratch" in str(ei.value)
    # A capacity error reports the entry count, as gsr_bin_sort's does.
    sc = scenes.config(0, n=2000, with_targets=False)
    gb, sp2 = ctx.upload(sc.gs), ctx.alloc(2000 * 48)
    ctx.project(sc.cams[0], gb, 2000, sp2, None, 0)
    small = ctx.alloc(16)
    with pytest.raises(lib.GsrError) as ei:
        ctx.bin_sort(sc.cams[0], sp2, 2000, small, rb, noalloc=True)
    assert ei.value.code == lib.GSR_E_CAPACITY and ei.value.entries > 4
    for b in (sp, rb, ob, gb, sp2, small):
        b.free()
