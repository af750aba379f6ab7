"""The float64 U-Net restatement used to check the GPU denoiser's gradients is
pinned by central finite differences of its own loss."""
import numpy as np

from oracle import unet


def make_case(seed=0, h=16, w=16):
    rng = np.random.default_rng(seed)
    lay, n = unet.param_layout()
    p = np.zeros(n)
    for name, cin, cout in unet.LAYERS:
        wo, bo, _, _ = lay[name]
        real = unet.IN_CH if name == "e0a" else cin
        W = rng.normal(scale=np.sqrt(2.0 / (9 * real)), size=(9, cout, cin))
        if name == "e0a":
            W[:, :, unet.IN_CH:] = 0
        p[wo:wo + W.size] = W.reshape(-1)
        p[bo:bo + cout] = rng.normal(scale=0.05, size=cout)
    p[lay["head_w"]:lay["head_w"] + 96] = rng.normal(scale=np.sqrt(1 / 32), size=96)
    feat = np.abs(rng.normal(size=(1, h, w, 10)))
    target = np.abs(rng.normal(size=(1, h, w, 3)))
    return p, feat, target


def test_unet_gradient_matches_finite_differences():
    p, feat, target = make_case()
    loss, g, out = unet.loss_and_grad(p, feat, target)
    assert np.isfinite(loss) and loss > 0
    # the loss's denominator is held constant in the gradient: differentiate
    # the loss with that denominator fixed at this point
    den = np.expm1(out) ** 2
    loss_f, g_f, _ = unet.loss_and_grad(p, feat, target, den)
    assert loss_f == loss and np.array_equal(g_f, g)
    lay, n = unet.param_layout()
    rng = np.random.default_rng(1)
    idx = []
    for name, cin, cout in unet.LAYERS:
        wo, bo, _, _ = lay[name]
        real = unet.IN_CH if name == "e0a" else cin
        for _ in range(2):
            t, co, ci = rng.integers(9), rng.integers(cout), rng.integers(real)
            idx.append(wo + (t * cout + co) * cin + ci)
        idx.append(bo + rng.integers(cout))
    idx += [lay["head_w"] + 5, lay["head_b"] + 1]
    h = 1e-6
    checked = 0
    for i in idx:
        q = p.copy()
        q[i] += h
        lp = unet.loss_only(q, feat, target, den)
        q[i] -= 2 * h
        lm = unet.loss_only(q, feat, target, den)
        fd = (lp - lm) / (2 * h)
        if abs(fd) < 1e-9 and abs(g[i]) < 1e-9:
            continue
        assert abs(fd - g[i]) <= 1e-4 * max(abs(fd), abs(g[i])) + 1e-9, (i, fd, g[i])
        checked += 1
    assert checked >= 25


def test_loss_is_relative_squared_error():
    """Closed form: with the head's output fixed, the loss is the mean of
    (expm1(out) - target)^2 / (expm1(out)^2 + 1e-2)."""
    p, feat, target = make_case(3, 8, 8)
    loss, _, out = unet.loss_and_grad(p, feat, target)
    pred = np.expm1(out)
    ref = np.mean((pred - target) ** 2 / (pred ** 2 + 1e-2))
    assert abs(loss - ref) <= 1e-12 * ref


def test_padded_input_channels_have_zero_gradient():
    p, feat, target = make_case(2)
    _, g, _ = unet.loss_and_grad(p, feat, target)
    lay, _ = unet.param_layout()
    wo, _, cin, cout = lay["e0a"]
    gw = g[wo:wo + 9 * cout * cin].reshape(9, cout, cin)
    assert np.all(gw[:, :, unet.IN_CH:] == 0) and np.any(gw[:, :, :unet.IN_CH] != 0)


def test_miss_pixels_do_not_enter_the_loss():
    """Pixels with depth guide 0 take no part: changing their target changes
    neither the loss nor the gradient, and the loss of the rest is unchanged
    by their presence."""
    p, feat, target = make_case(4, 8, 8)
    feat[0, 2:4, 3:6, 9] = 0.0
    l0, g0, _ = unet.loss_and_grad(p, feat, target)
    t2 = target.copy()
    t2[0, 2:4, 3:6] += 5.0
    l1, g1, _ = unet.loss_and_grad(p, feat, t2)
    assert l0 == l1 and np.array_equal(g0, g1)
    t3 = target.copy()
    t3[0, 0, 0] += 5.0
    l2, _, _ = unet.loss_and_grad(p, feat, t3)
    assert l2 != l0


def test_bf16_rounding_points():
    """bf16_round is round-to-nearest-even on the top 16 bits of float32
    (known answers, ties both ways), and the rounded restatement stays close
    to the exact one (its gradient differs by the bf16 storage error only)."""
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -8 + 2 ** -20, -2.5, 0.0, 2.0 - 2 ** -12,
                  -(1.0 + 2 ** -7 + 2 ** -9)])
    # halfway cases go to the even mantissa; 2 - 2^-12 carries into the exponent
    want = [1.0, 1.0, 1.0 + 2 ** -6, 1.0 + 2 ** -7, -2.5, 0.0, 2.0, -(1.0 + 2 ** -7)]
    assert unet.bf16_round(x).tolist() == want
    p, feat, target = make_case(2)
    l0, g0, _ = unet.loss_and_grad(p, feat, target)
    l1, g1, _ = unet.loss_and_grad(p, feat, target, bf16=True)
    assert l1 != l0 and abs(l1 - l0) < 0.02 * l0
    assert 0 < np.linalg.norm(g1 - g0) < 0.1 * np.linalg.norm(g0)
