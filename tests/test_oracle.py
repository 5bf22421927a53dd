"""Pins the CPU oracle (oracle/gsr_oracle.cpp) before the GPU results are compared with it.

Each check is independent of the oracle's own formulas: closed forms, invariants,
brute force written separately here, known answers from the reference's tests
(trainer/ssim-test.html) and central differences of the FP64 forward.
"""
import math

import numpy as np
import pytest

import oracle as O
from conftest import rel_l2
from gsr import scenes


def axis_camera(d=3.0, w=64, h=48, f=100.0):
    """Camera at (0, 0, -d) looking down +z with identity rotation."""
    view = np.array([1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, d], dtype=np.float64)
    return scenes.Camera(view, f, f, w / 2.0, h / 2.0, w, h)


def gaussian(mean, log_s, q=(1, 0, 0, 0), op_logit=0.0, col=(1, 1, 1)):
    g = np.zeros(16)
    g[0:3] = mean
    g[4:7] = log_s
    g[7] = op_logit
    g[8:12] = q
    g[12:15] = col
    return g


# ---------------------------------------------------------------- projection


def test_project_isotropic_closed_form():
    cam = axis_camera(d=3.0)
    sigma = 0.05
    gs = gaussian([0, 0, 0], [math.log(sigma)] * 3, op_logit=0.7)[None]
    sp = O.project(cam, gs)[0]
    var = (100.0 * sigma / 3.0) ** 2 + 0.3  # (f sigma / z)^2 + dilation
    assert sp[11] == 1.0
    np.testing.assert_allclose(sp[0:2], [32.0, 24.0], atol=1e-12)
    np.testing.assert_allclose(sp[2:5], [1 / var, 0.0, 1 / var], rtol=1e-12, atol=1e-15)
    assert sp[5] == pytest.approx(1 / (1 + math.exp(-0.7)), rel=1e-12)
    # reference radius: ceil(3 sqrt(mid + sqrt(max(0.1, mid^2 - det)))), mid^2 - det = 0 here
    assert sp[6] == math.ceil(3 * math.sqrt(var + math.sqrt(0.1)))
    assert sp[7] == pytest.approx(3.0, abs=1e-12)


def test_project_conic_inverts_ewa_covariance():
    """conic = inverse of J W (R S)(R S)^T W^T J^T + 0.3 I, computed here with numpy."""
    sc = scenes.config(0, n=300, with_targets=False)
    cam = sc.cams[0]
    gs = sc.gs.astype(np.float64)
    sp = O.project(cam, gs)
    # The camera travels as the float32 uniform (gpu.js:696-701).
    f32 = lambda x: np.float64(np.float32(x))  # noqa: E731
    V = cam.view.astype(np.float32).astype(np.float64).reshape(3, 4)
    cam = scenes.Camera(V.ravel(), f32(cam.fx), f32(cam.fy), f32(cam.cx), f32(cam.cy), cam.width, cam.height)
    checked = 0
    for i in range(gs.shape[0]):
        if sp[i, 11] < 0.5:
            continue
        q = gs[i, 8:12] / np.linalg.norm(gs[i, 8:12])
        w, x, y, z = q
        R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                      [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                      [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])
        assert np.allclose(R @ R.T, np.eye(3)) and np.linalg.det(R) == pytest.approx(1.0)
        M = R @ np.diag(np.exp(gs[i, 4:7]))
        p = V[:, :3] @ gs[i, :3] + V[:, 3]
        J = np.array([[cam.fx / p[2], 0, -cam.fx * p[0] / p[2] ** 2], [0, cam.fy / p[2], -cam.fy * p[1] / p[2] ** 2]])
        cov = J @ V[:, :3] @ M @ M.T @ V[:, :3].T @ J.T + 0.3 * np.eye(2)
        conic = np.array([[sp[i, 2], sp[i, 3]], [sp[i, 3], sp[i, 4]]])
        inv = np.linalg.inv(cov)
        np.testing.assert_allclose(conic, inv, rtol=1e-9, atol=1e-9 * np.abs(inv).max())
        np.testing.assert_allclose(sp[i, 0:2], [cam.fx * p[0] / p[2] + cam.cx, cam.fy * p[1] / p[2] + cam.cy])
        checked += 1
    assert checked > 250


def test_project_quaternion_invariance_and_rejects():
    sc = scenes.config(0, n=200, with_targets=False)
    cam = sc.cams[0]
    gs = sc.gs.astype(np.float64)
    base = O.project(cam, gs)
    for k in (-1.0, 2.5):
        g2 = gs.copy()
        g2[:, 8:12] *= k
        np.testing.assert_allclose(O.project(cam, g2), base, rtol=1e-12, atol=1e-12)
    # near plane: z < 0.2 rejects (gpu.js:36)
    cam2 = axis_camera(d=0.1)
    sp = O.project(cam2, gaussian([0, 0, 0.05], [-3] * 3)[None])[0]
    assert sp[11] == 0.0 and not sp.any()
    sp = O.project(cam2, gaussian([0, 0, 0.15], [-3] * 3)[None])[0]
    assert sp[11] == 1.0


def test_sh_basis_orthonormal_by_quadrature():
    # Gauss-Legendre in cos(theta) x uniform phi: exact for these polynomials.
    nt, nphi = 24, 48
    ct, wt = np.polynomial.legendre.leggauss(nt)
    phi = (np.arange(nphi) + 0.5) * 2 * np.pi / nphi
    G = np.zeros((16, 16))
    c0 = 0.28209479177387814  # Y_00; the oracle stores DC in colour units (Y[0] = 1)
    for a, wa in zip(ct, wt):
        st = math.sqrt(1 - a * a)
        for p in phi:
            d = np.array([st * math.cos(p), st * math.sin(p), a])
            Y = O.sh_basis(3, d)
            Y[0] = c0
            G += wa * (2 * np.pi / nphi) * np.outer(Y, Y)
    np.testing.assert_allclose(G, np.eye(16), atol=1e-12)


def test_sh_degree_one_sign_convention():
    # 3D Gaussian Splatting's degree-1 basis: -C1 y, C1 z, -C1 x with C1 = sqrt(3 / (4 pi)).
    c1 = math.sqrt(3 / (4 * math.pi))
    d = np.array([0.36, 0.48, 0.8])
    np.testing.assert_allclose(O.sh_basis(1, d)[1:4], [-c1 * 0.48, c1 * 0.8, -c1 * 0.36], rtol=1e-15)
    assert not O.sh_basis(1, d)[4:].any()


def test_sh_colour_reduces_to_dc():
    sc = scenes.config(1, n=500, with_targets=False)
    cam = sc.cams[0]
    sh = sc.sh.astype(np.float64).copy()
    sh[:, 1:, :] = 0.0
    a = O.project(cam, sc.gs.astype(np.float64), sh, 3)
    b = O.project(cam, sc.gs.astype(np.float64), None, 0)
    np.testing.assert_array_equal(a, b)  # DC == record colour in these scenes


# ---------------------------------------------------------------- tile lists


def brute_tile_lists(sp, W, H):
    tx, ty = (W + 15) // 16, (H + 15) // 16
    valid = [i for i in range(sp.shape[0]) if float(sp[i, 11]) > 0.5]
    valid.sort(key=lambda i: (float(sp[i, 7]), i))
    lists = [[] for _ in range(tx * ty)]
    for i in valid:
        u, v, r = float(sp[i, 0]), float(sp[i, 1]), float(sp[i, 6])
        x0 = max(0, math.floor((u - r - 0.5) / 16))
        x1 = min(tx - 1, math.floor((u + r - 0.5) / 16))
        y0 = max(0, math.floor((v - r - 0.5) / 16))
        y1 = min(ty - 1, math.floor((v + r - 0.5) / 16))
        for y in range(y0, y1 + 1):
            for x in range(x0, x1 + 1):
                lists[y * tx + x].append(i)
    flat, ranges, pos = [], [], 0
    for l in lists:
        ranges += [pos, pos + len(l)]
        flat += l
        pos += len(l)
    return np.array(flat, dtype=np.uint32), np.array(ranges, dtype=np.uint32)


@pytest.mark.parametrize("W,H,n", [(64, 48, 400), (100, 37, 300), (16, 16, 50)])
def test_tile_lists_brute_force(W, H, n):
    rng = np.random.default_rng(W + H)
    sp = np.zeros((n, 12), np.float32)
    sp[:, 0] = rng.uniform(-20, W + 20, n)
    sp[:, 1] = rng.uniform(-20, H + 20, n)
    sp[:, 6] = np.ceil(rng.uniform(0, 25, n))
    sp[:, 7] = rng.choice([0.5, 1.0, 1.5], n).astype(np.float32)  # ties
    sp[:, 7] += (rng.uniform(size=n) < 0.5) * rng.uniform(0, 1, n).astype(np.float32)
    sp[:, 11] = rng.uniform(size=n) > 0.1
    lst, ranges = O.tile_lists(sp, W, H)
    bl, br = brute_tile_lists(sp, W, H)
    np.testing.assert_array_equal(ranges, br)
    np.testing.assert_array_equal(lst, bl)


# ---------------------------------------------------------------- forward


def single_splat(u, v, A, B, C, op, r, col, depth=1.0):
    s = np.zeros(12)
    s[:] = [u, v, A, B, C, op, r, depth, *col, 1.0]
    return s


def test_forward_one_and_two_gaussians():
    W = H = 16
    a = single_splat(8.5, 8.5, 0.5, 0.0, 0.5, 0.6, 5, (0.2, 0.4, 0.8), depth=1.0)
    b = single_splat(8.5, 8.5, 0.5, 0.0, 0.5, 0.7, 5, (1.0, 0.5, 0.0), depth=2.0)
    sp = np.stack([b, a])  # index order != depth order
    lst, ranges = O.tile_lists(sp, W, H)
    img, st = O.forward(sp, lst, ranges, W, H)
    # pixel (8, 8) is at the centre: alpha = opacity
    expect = 0.6 * np.array([0.2, 0.4, 0.8]) + 0.7 * (1 - 0.6) * np.array([1.0, 0.5, 0.0])
    np.testing.assert_allclose(img[8, 8], expect, rtol=1e-14)
    assert st[8, 8, 0] == pytest.approx((1 - 0.6) * (1 - 0.7))
    assert st[8, 8, 1] == 2
    # a neighbour: alpha = op * exp(-0.5 * 0.5 * 1)
    ea, eb = 0.6 * math.exp(-0.25), 0.7 * math.exp(-0.25)
    np.testing.assert_allclose(img[8, 9], ea * np.array([0.2, 0.4, 0.8]) + eb * (1 - ea) * np.array([1.0, 0.5, 0.0]))
    # outside the square radius: nothing
    assert not img[8, 14].any() and not img[1, 8].any()


def test_forward_stop_and_thresholds():
    W = H = 16
    sps = [single_splat(8.5, 8.5, 1e-6, 0.0, 1e-6, 1.0, 8, (1, 1, 1), depth=1.0 + k) for k in range(5)]
    sp = np.stack(sps)
    lst, ranges = O.tile_lists(sp, W, H)
    img, st = O.forward(sp, lst, ranges, W, H)
    # alpha clamps at 0.99 (opacity 1); T = 0.01, ~1e-4 (just above it in float64),
    # 1e-6 -> the check before the fourth entry stops the pixel
    assert st[8, 8, 1] == 3
    assert st[8, 8, 0] == pytest.approx(1e-6)
    np.testing.assert_allclose(img[8, 8], 0.99 + 0.99 * 0.01 + 0.99 * 0.01 * 0.01)
    # alpha below 1/255 is skipped entirely
    faint = np.stack([single_splat(8.5, 8.5, 1e-6, 0.0, 1e-6, 1 / 256, 8, (1, 1, 1))])
    lst, ranges = O.tile_lists(faint, W, H)
    img, st = O.forward(faint, lst, ranges, W, H)
    assert not img.any() and st[8, 8, 1] == 0 and st[8, 8, 0] == 1.0


# ---------------------------------------------------------------- loss


def lcg_pair(W=20, H=14):
    seed = [7]

    def rand():
        seed[0] = (seed[0] * 16807) % 2147483647
        return seed[0] / 2147483647

    target = np.array([rand() for _ in range(W * H * 3)])
    image = np.array([min(1.0, max(0.0, v + (rand() - 0.5) * 0.3)) for v in target])
    return image.reshape(H, W, 3), target.reshape(H, W, 3), rand


def test_loss_l1_known_answer():
    image, target, _ = lcg_pair()
    g = O.loss_grad(image, target, 0.0)
    np.testing.assert_allclose(g, np.sign(image - target) / image.size, atol=1e-12, rtol=0)  # ssim-test.html:52-53


def test_loss_ssim_gradient_matches_central_differences():
    # ssim-test.html:41-51: 40 random pixels away from the L1 kink, h = 1e-3.
    image, target, rand = lcg_pair()
    lam = 0.2
    g = O.loss_grad(image, target, lam).ravel()
    flat = image.ravel()
    worst, checked = 0.0, 0
    for _ in range(40):
        i = int(rand() * flat.size)
        if abs(flat[i] - target.ravel()[i]) < 0.01:
            continue
        up, dn = flat.copy(), flat.copy()
        up[i] += 1e-3
        dn[i] -= 1e-3
        num = (O.loss_value(up.reshape(image.shape), target, lam) - O.loss_value(dn.reshape(image.shape), target, lam)) / 2e-3
        worst = max(worst, abs(num - g[i]) / max(1e-6, abs(num)))
        checked += 1
    assert checked > 30
    assert worst < 1e-4


def ssim_by_convolution(image, target, kernel1d):
    # Wang et al. 2004 SSIM averaged over the valid windows and the channels,
    # written as separable 2D correlations (scipy) rather than per-window loops.
    from scipy.signal import correlate2d

    k = np.outer(kernel1d, kernel1d)
    total, count = 0.0, 0
    for c in range(3):
        x, y = image[..., c], target[..., c]
        mx, my = correlate2d(x, k, "valid"), correlate2d(y, k, "valid")
        vx = correlate2d(x * x, k, "valid") - mx * mx
        vy = correlate2d(y * y, k, "valid") - my * my
        cxy = correlate2d(x * y, k, "valid") - mx * my
        s = (2 * mx * my + 1e-4) * (2 * cxy + 9e-4) / ((mx * mx + my * my + 1e-4) * (vx + vy + 9e-4))
        total += s.sum()
        count += s.size
    return total / count


@pytest.mark.parametrize("window", [7, 11])
def test_loss_value_matches_convolution_ssim(window):
    image, target, _ = lcg_pair(26, 19)
    if window == 7:
        kernel = np.full(7, 1 / 7)  # ssim-test.html:13-30: 7x7 box
    else:
        d = np.arange(11) - 5.0
        kernel = np.exp(-d * d / (2 * 1.5 ** 2))
        kernel /= kernel.sum()
        # the published taps of the normalised sigma-1.5 11-tap SSIM window
        np.testing.assert_allclose(kernel[:6], [0.00102838, 0.00759876, 0.03600077, 0.10936069, 0.21300554,
                                                0.26601172], atol=1e-8)
    lam = 0.2
    want = (1 - lam) * np.abs(image - target).mean() + lam * (1 - ssim_by_convolution(image, target, kernel))
    assert abs(O.loss_value(image, target, lam, window=window) - want) < 1e-12


@pytest.mark.parametrize("window", [7, 11])
def test_ssim_gradient_vanishes_at_identity(window):
    # SSIM(x, x) = 1 is its maximum: the SSIM term's gradient is zero there, and
    # L1's sign(0) is 0, so the whole gradient vanishes.
    image, _, _ = lcg_pair(26, 19)
    assert abs(O.loss_value(image, image, 1.0, window=window)) < 1e-12
    assert np.abs(O.loss_grad(image, image, 1.0, window=window)).max() < 1e-14


def test_loss_gauss11_gradient_matches_central_differences():
    image, target, rand = lcg_pair(26, 19)
    lam = 0.2
    g = O.loss_grad(image, target, lam, window=11).ravel()
    flat = image.ravel()
    worst, checked = 0.0, 0
    for _ in range(40):
        i = int(rand() * flat.size)
        if abs(flat[i] - target.ravel()[i]) < 0.01:
            continue
        up, dn = flat.copy(), flat.copy()
        up[i] += 1e-3
        dn[i] -= 1e-3
        num = (O.loss_value(up.reshape(image.shape), target, lam, window=11) -
               O.loss_value(dn.reshape(image.shape), target, lam, window=11)) / 2e-3
        worst = max(worst, abs(num - g[i]) / max(1e-6, abs(num)))
        checked += 1
    assert checked > 30
    assert worst < 1e-4


# ---------------------------------------------------------------- backward and chain


def fd_scene():
    sc = scenes.config(0, n=400)
    cam = sc.cams[0].crop(64, 64, 96, 80)
    return sc, cam


def test_backward_matches_central_differences():
    """Screen-space gradients of the FP64 blend against central differences of
    sum(w * image) with respect to each splat field."""
    sc, cam = fd_scene()
    sp, lst, ranges, img, st = O.render(cam, sc.gs.astype(np.float64))
    rng = np.random.default_rng(0)
    w = rng.standard_normal(img.shape)
    grads = O.backward(sp, lst, ranges, st, w, cam.width, cam.height)
    hits = np.argsort(-np.abs(grads[:, 4:9]).sum(1))[:12]

    def L(s):
        return float((O.forward(s, lst, ranges, cam.width, cam.height)[0] * w).sum())

    fields = {0: 4, 1: 5, 2: 6, 3: 7, 4: 8, 8: 0, 9: 1, 10: 2}  # splat field -> gradient slot
    h = 1e-7
    bad = 0
    for i in hits:
        for f, slot in list(fields.items()) + [(5, 3)]:
            up, dn = sp.copy(), sp.copy()
            up[i, f] += h
            dn[i, f] -= h
            num = (L(up) - L(dn)) / (2 * h)
            if f == 5:  # opacity -> logit
                op = sp[i, 5]
                num *= op * (1 - op)
            an = grads[i, slot]
            if abs(num - an) > 1e-6 * max(1.0, abs(an)) + 1e-5 * abs(an):
                bad += 1
    assert bad == 0


def test_shape_chain_matches_projection_differences():
    """Analytic chain of projectShape against central differences (FP64)."""
    sc = scenes.config(0, n=200, with_targets=False)
    cam = sc.cams[0]
    gs = sc.gs.astype(np.float64)
    rng = np.random.default_rng(1)
    grads = np.zeros((gs.shape[0], 16))
    grads[:, 4:9] = rng.standard_normal((gs.shape[0], 5))
    shape, _ = O.shape_grad(cam, gs, grads)
    idx = [0, 1, 2, 4, 5, 6, 8, 9, 10, 11]
    h = 1e-6
    for i in range(0, gs.shape[0], 7):
        q = gs[i, idx]
        ok, _ = O.project_shape(cam, q)
        if not ok:
            assert not shape[i].any()
            continue
        for k in range(10):
            up, dn = q.copy(), q.copy()
            up[k] += h
            dn[k] -= h
            _, pu = O.project_shape(cam, up)
            _, pd = O.project_shape(cam, dn)
            num = float(grads[i, 4:9] @ (pu[:5] - pd[:5]) / (2 * h))
            assert num == pytest.approx(shape[i, k], rel=1e-6, abs=1e-6 * (1 + abs(shape[i, k])))


@pytest.mark.parametrize("scene", [0, 2])
def test_full_chain_matches_central_differences(scene):
    """Every parameter group (mean, log-scale, rotation, opacity logit, colour or
    SH) through project -> tile lists -> blend, FP64, h = 1e-7."""
    if scene == 0:
        sc, cam = fd_scene()
    else:
        sc = scenes.config(2, n=4000)
        cam = sc.cams[0].crop(896, 476, 128, 128)
    gs = sc.gs.astype(np.float64)
    sh = None if sc.sh is None else sc.sh.astype(np.float64)
    deg = sc.sh_degree
    img = O.render(cam, gs, sh, deg)[3]
    w = np.random.default_rng(9).standard_normal(img.shape)
    _, grads, shape, shg = O.gradients(cam, gs, w, sh, deg)

    def L(g, s):
        return float((O.render(cam, g, s, deg)[3] * w).sum())

    order = np.argsort(-(np.abs(shape).sum(1) + np.abs(grads[:, :4]).sum(1)))[:6]
    slots = [0, 1, 2, 4, 5, 6, 8, 9, 10, 11, 7] + ([12, 13, 14] if deg == 0 else [])
    h = 1e-7
    for i in order:
        for k, s in enumerate(slots):
            gp, gm = gs.copy(), gs.copy()
            gp[i, s] += h
            gm[i, s] -= h
            num = (L(gp, sh) - L(gm, sh)) / (2 * h)
            an = shape[i, k] if k < 10 else (grads[i, 3] if s == 7 else grads[i, s - 12])
            assert num == pytest.approx(an, rel=2e-4, abs=1e-5), (i, s)
        if deg:
            for kk, c in ((0, 0), (3, 1), (8, 2), (15, 0)):
                sp_, sm_ = sh.copy(), sh.copy()
                sp_[i, kk, c] += h
                sm_[i, kk, c] -= h
                num = (L(gs, sp_) - L(gs, sm_)) / (2 * h)
                assert num == pytest.approx(shg[i, kk, c], rel=2e-4, abs=1e-5), (i, kk, c)


def test_reference_fd_shape_agrees_with_analytic_chain():
    """SHAPE's float32 central differences (gpu.js:315-327) estimate the analytic
    chain; they agree to FD accuracy (h = 1e-3 / 2e-3 in float32)."""
    sc = scenes.config(0, n=2000, with_targets=False)
    cam = sc.cams[0]
    rng = np.random.default_rng(2)
    grads = np.zeros((sc.n, 16), np.float32)
    grads[:, 4:9] = rng.standard_normal((sc.n, 5)).astype(np.float32)
    fd = O.shape_fd(cam, sc.gs, grads, np.float32)
    an, _ = O.shape_grad(cam, sc.gs.astype(np.float64), grads.astype(np.float64))
    assert rel_l2(fd, an) < 2e-2


# ---------------------------------------------------------------- Adam, stats


def test_adam_closed_form():
    rng = np.random.default_rng(3)
    n = 50
    gs = rng.standard_normal((n, 16)).astype(np.float32)
    gs[:, 12:15] = rng.uniform(0.2, 0.8, (n, 3))
    grads = rng.standard_normal((n, 16)).astype(np.float32)
    shape = rng.standard_normal((n, 10)).astype(np.float32)
    lr = {"color": 0.025, "opacity": 0.05, "mean": 0.002, "scale": 0.005, "rotation": 0.001}
    out, mom = O.adam(1.0, lr, gs, grads, shape, np.zeros((n, 28), np.float32), np.float64)
    # t = 1: m_hat = g, v_hat = g^2 -> step = lr * g / (|g| + 1e-8) (Kingma & Ba 2015)
    slot = [12, 13, 14, 7, 0, 1, 2, 4, 5, 6, 8, 9, 10, 11]
    g = np.concatenate([grads[:, :4], shape], axis=1).astype(np.float64)
    lrs = np.array([lr["color"]] * 3 + [lr["opacity"]] + [lr["mean"]] * 3 + [lr["scale"]] * 3 + [lr["rotation"]] * 4)
    expect = gs.astype(np.float64).copy()
    expect[:, slot] -= lrs * g / (np.abs(g) + 1e-8)
    expect[:, 12:15] = np.clip(expect[:, 12:15], 0, 1)
    expect[:, 8:12] /= np.linalg.norm(expect[:, 8:12], axis=1, keepdims=True)
    np.testing.assert_allclose(out, expect, atol=2e-6)
    np.testing.assert_allclose(mom[:, :14], 0.1 * g, rtol=1e-12)
    np.testing.assert_allclose(mom[:, 14:], 0.001 * g * g, rtol=1e-12)
    # t = 2 continues the moments
    out2, mom2 = O.adam(2.0, lr, out, grads, shape, mom, np.float64)
    m = 0.9 * mom[:, :14] + 0.1 * g
    v = 0.999 * mom[:, 14:] + 0.001 * g * g
    np.testing.assert_allclose(mom2[:, :14], m, rtol=1e-12)
    np.testing.assert_allclose(mom2[:, 14:], v, rtol=1e-12)


def test_adam_sh_closed_form():
    """SH3 Adam: at t = 1 every parameter moves by lr * g / (|g| + 1e-8) (Kingma & Ba
    with m = 0.1 g, v = 0.001 g^2 and bias corrections 0.1, 0.001), each group at
    its own rate; then the quaternion is renormalised."""
    rng = np.random.default_rng(7)
    n = 50
    gs = rng.standard_normal((n, 16)).astype(np.float32)
    sh = rng.standard_normal((n, 16, 3)).astype(np.float32)
    grads = rng.standard_normal((n, 16))
    shape = rng.standard_normal((n, 10))
    shg = rng.standard_normal((n, 16, 3))
    lr = {"color": 5e-3, "opacity": 0.05, "mean": 1.6e-4, "scale": 5e-3, "rotation": 1e-3}
    gs1, sh1, mom = O.adam_sh(1.0, lr, 2.5e-4, gs, sh, grads, shape, shg, np.zeros((n, 120)), np.float64)
    step = lambda g, r: r * g / (np.abs(g) + 1e-8)  # noqa: E731
    np.testing.assert_allclose(gs1[:, 0:3], gs[:, 0:3] - step(shape[:, 0:3], lr["mean"]), atol=1e-6)
    np.testing.assert_allclose(gs1[:, 4:7], gs[:, 4:7] - step(shape[:, 3:6], lr["scale"]), atol=1e-6)
    np.testing.assert_allclose(gs1[:, 7], gs[:, 7] - step(grads[:, 3], lr["opacity"]), atol=1e-6)
    q = gs[:, 8:12] - step(shape[:, 6:10], lr["rotation"])
    np.testing.assert_allclose(gs1[:, 8:12], q / np.linalg.norm(q, axis=1, keepdims=True), atol=1e-6)
    rates = np.full((16, 3), 2.5e-4)
    rates[0] = lr["color"]
    np.testing.assert_allclose(sh1, sh - step(shg, rates[None]), atol=1e-6)
    flat = mom.ravel()
    m_sh, v_sh = flat[:48 * n].reshape(n, 48), flat[48 * n:96 * n].reshape(n, 48)
    m_base, v_base = flat[96 * n:107 * n].reshape(11, n), flat[107 * n:118 * n].reshape(11, n)
    np.testing.assert_allclose(m_base[10], 0.1 * grads[:, 3], rtol=1e-12)
    np.testing.assert_allclose(v_base[0:3].T, 0.001 * shape[:, 0:3] ** 2, rtol=1e-12)
    np.testing.assert_allclose(m_sh, 0.1 * shg.reshape(n, 48), rtol=1e-12)
    np.testing.assert_allclose(v_sh, 0.001 * shg.reshape(n, 48) ** 2, rtol=1e-12)
    assert not flat[118 * n:].any()
    # untouched slots: pad 3, colour 12-14 (SH0 colour is not an SH3 parameter), 15
    np.testing.assert_array_equal(gs1[:, [3, 12, 13, 14, 15]], gs[:, [3, 12, 13, 14, 15]])


def test_stats_accumulate():
    sp = np.zeros((3, 12), np.float32)
    sp[:, 11] = [1, 0, 1]
    grads = np.zeros((3, 16), np.float32)
    grads[:, 4:6] = [[3, 4], [1, 1], [0, 2]]
    st = O.stats(grads, sp, np.zeros((3, 2), np.float32))
    st = O.stats(grads, sp, st)
    np.testing.assert_allclose(st, [[10, 2], [0, 0], [4, 2]])


def test_tile_reach_matches_single_splat_renders():
    """tile_reach (the largest alpha a splat reaches over a tile) against the
    oracle's own forward: each splat rendered alone into each tile of a small
    image (one-entry lists, T = 1, white colour), so a pixel's red channel is the
    splat's alpha there, or 0 where alphaAt misses.  Agreement where the forward
    hits; where tile_reach is below 1/255 no pixel of the tile is hit."""
    rng = np.random.default_rng(17)
    W, H, n = 67, 41, 300  # ragged tiles on both edges
    sp = np.zeros((n, 12), np.float32)
    sp[:, 0] = rng.uniform(-10, W + 10, n)
    sp[:, 1] = rng.uniform(-10, H + 10, n)
    a = rng.uniform(0.002, 0.3, n)
    c = rng.uniform(0.002, 0.3, n)
    b = rng.uniform(-0.9, 0.9, n) * np.sqrt(a * c)
    sp[:, 2], sp[:, 3], sp[:, 4] = a, b, c
    sp[:, 5] = rng.uniform(0.002, 1.0, n)
    sp[:, 6] = np.ceil(rng.uniform(1, 30, n))
    sp[:, 7] = 1.0
    sp[:, 8:11] = 1.0
    sp[:, 11] = 1.0
    tx, ty = (W + 15) // 16, (H + 15) // 16
    T = tx * ty
    tiles = np.repeat(np.arange(T, dtype=np.uint32), n)
    gidx = np.tile(np.arange(n, dtype=np.uint32), T)
    reach = O.tile_reach(sp, tiles, gidx, W, H).reshape(T, n)
    hits = misses = 0
    for i in range(n):
        lst = np.array([i] * T, dtype=np.uint32)
        ranges = np.stack([np.arange(T), np.arange(T) + 1], 1).astype(np.uint32).ravel()
        img, _ = O.forward(sp.astype(np.float64), lst, ranges, W, H)
        for t in range(T):
            y0, x0 = (t // tx) * 16, (t % tx) * 16
            peak = img[y0:y0 + 16, x0:x0 + 16, 0].max()
            if reach[t, i] >= 1 / 255:
                assert peak == pytest.approx(reach[t, i], rel=1e-12)
                hits += 1
            else:
                assert peak == 0.0
                misses += 1
    assert hits > 100 and misses > 100
