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
    loss, g, _ = unet.loss_and_grad(p, feat, target)
    assert np.isfinite(loss) and loss > 0
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
        lp = unet.loss_only(q, feat, target)
        q[i] -= 2 * h
        lm = unet.loss_only(q, feat, target)
        fd = (lp - lm) / (2 * h)
        if abs(fd) < 1e-9 and abs(g[i]) < 1e-9:
            continue
        assert abs(fd - g[i]) <= 1e-4 * max(abs(fd), abs(g[i])) + 1e-9, (i, fd, g[i])
        checked += 1
    assert checked >= 25


def test_padded_input_channels_have_zero_gradient():
    p, feat, target = make_case(2)
    _, g, _ = unet.loss_and_grad(p, feat, target)
    lay, _ = unet.param_layout()
    wo, _, cin, cout = lay["e0a"]
    gw = g[wo:wo + 9 * cout * cin].reshape(9, cout, cin)
    assert np.all(gw[:, :, unet.IN_CH:] == 0) and np.any(gw[:, :, :unet.IN_CH] != 0)
