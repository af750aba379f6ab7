"""Float64 NumPy restatement of the U-Net denoiser: forward pass, loss and
hand-derived backward pass, for checking the GPU kernels' gradients.

TEST INFRASTRUCTURE ONLY.  The network is the one the C ABI's denoiser
documents (include/s3dg.h): 3x3 convolutions (zero padding) with ReLU,
2x2 average pooling, nearest upsampling, channel concatenation of skips,
a 1x1 head to log radiance, L1 loss on log(1 + radiance).  Parameters use
the same flat layout as s3dg_cnn_get_params.  This restatement is itself
pinned by central finite differences (tests/test_oracle_unet.py).
"""
from __future__ import annotations

import numpy as np

IN_CH, IN_PAD, OUT = 10, 32, 3
LAYERS = [("e0a", IN_PAD, 32), ("e0b", 32, 32), ("e1a", 32, 64), ("e1b", 64, 64), ("e2a", 64, 96), ("e2b", 96, 96),
          ("b0", 96, 128), ("b1", 128, 128), ("d2a", 128 + 96, 96), ("d2b", 96, 96), ("d1a", 96 + 64, 64),
          ("d1b", 64, 64), ("d0a", 64 + 32, 32), ("d0b", 32, 32)]


def param_layout():
    off, lay = 0, {}
    for name, cin, cout in LAYERS:
        lay[name] = (off, off + 9 * cout * cin, cin, cout)   # W offset, b offset
        off += 9 * cout * cin + cout
    lay["head_w"] = off
    off += OUT * 32
    lay["head_b"] = off
    off += OUT
    return lay, off


def unpack(params):
    lay, n = param_layout()
    assert params.size == n
    P = {}
    for name, cin, cout in LAYERS:
        wo, bo, _, _ = lay[name]
        P[name] = (params[wo:wo + 9 * cout * cin].reshape(9, cout, cin), params[bo:bo + cout])
    P["head"] = (params[lay["head_w"]:lay["head_w"] + OUT * 32].reshape(OUT, 32), params[lay["head_b"]:lay["head_b"] + OUT])
    return P


def pack_features(f):
    x = np.zeros(f.shape[:3] + (IN_PAD,))
    x[..., :IN_CH] = f
    x[..., 0:3] = np.log1p(np.maximum(f[..., 0:3], 0))
    x[..., 9] = np.log1p(np.maximum(f[..., 9], 0))
    return x


def im2col(x):
    n, h, w, c = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
    cols = np.empty((n, h, w, 9, c))
    for t in range(9):
        dy, dx = t // 3, t % 3
        cols[:, :, :, t, :] = xp[:, dy:dy + h, dx:dx + w, :]
    return cols


def col2im(dcols, shape):
    n, h, w, c = shape
    dxp = np.zeros((n, h + 2, w + 2, c))
    for t in range(9):
        dy, dx = t // 3, t % 3
        dxp[:, dy:dy + h, dx:dx + w, :] += dcols[:, :, :, t, :]
    return dxp[:, 1:h + 1, 1:w + 1, :]


def conv_fwd(x, W, b):
    cols = im2col(x)
    # out[p, co] = sum_{t, ci} cols[p, t, ci] * W[t, co, ci]
    y = np.einsum("nhwtc,toc->nhwo", cols, W) + b
    return y, cols


def conv_bwd(dy, cols, W, xshape):
    dW = np.einsum("nhwo,nhwtc->toc", dy, cols)
    db = dy.sum(axis=(0, 1, 2))
    dcols = np.einsum("nhwo,toc->nhwtc", dy, W)
    return col2im(dcols, xshape), dW, db


def pool(x):
    n, h, w, c = x.shape
    return x.reshape(n, h // 2, 2, w // 2, 2, c).mean(axis=(2, 4))


def pool_bwd(dy):
    return np.repeat(np.repeat(dy, 2, axis=1), 2, axis=2) / 4.0


def up(x):
    return np.repeat(np.repeat(x, 2, axis=1), 2, axis=2)


def up_bwd(dy):
    n, h, w, c = dy.shape
    return dy.reshape(n, h // 2, 2, w // 2, 2, c).sum(axis=(2, 4))


REL_EPS = 1e-2


def bf16_round(x):
    """Round-to-nearest-even to bfloat16 (through float32), returned as float64."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def loss_and_grad(params, features, target, den=None, bf16=False):
    """Returns (mean loss, gradient vector, head output in log space).

    The prediction is expm1 of the head output.  The loss per value is the
    relative squared error (pred - target)^2 / (den + REL_EPS) with den =
    pred^2 held constant (not differentiated), or the given array `den`
    (fixing it makes the loss a smooth function of the parameters, which is
    how the gradient is checked by finite differences).  Pixels whose depth
    guide is 0 (camera ray missed) are excluded from the loss: the network's
    output is replaced there by the input radiance.

    With bf16=True every tensor the GPU kernels store in bfloat16 is rounded
    to it at the same point: the packed input, each layer's and pooling's
    output, the head's input gradient, each data-gradient convolution's
    output, the 2x2 sums of the upsampled gradients and the skip gradients
    after the pooling path's contribution is added (sums in between stay
    float64)."""
    P = unpack(np.asarray(params, np.float64))
    lay, nparam = param_layout()
    cache = {}
    q = bf16_round if bf16 else (lambda v: v)

    def layer(name, x):
        W, b = P[name]
        z, cols = conv_fwd(x, W, b)
        y = q(np.maximum(z, 0))
        cache[name] = (x.shape, cols, y)
        return y

    def qpool(x):
        return q(pool(x))

    x = q(pack_features(np.asarray(features, np.float64)))
    e0 = layer("e0b", layer("e0a", x))
    e1 = layer("e1b", layer("e1a", qpool(e0)))
    e2 = layer("e2b", layer("e2a", qpool(e1)))
    bt = layer("b1", layer("b0", qpool(e2)))
    d2 = layer("d2b", layer("d2a", np.concatenate([up(bt), e2], -1)))
    d1 = layer("d1b", layer("d1a", np.concatenate([up(d2), e1], -1)))
    d0 = layer("d0b", layer("d0a", np.concatenate([up(d1), e0], -1)))
    Wh, bh = P["head"]
    out = d0 @ Wh.T + bh
    t = np.maximum(np.asarray(target, np.float64), 0)
    pred = np.expm1(out)
    d = pred * pred if den is None else np.asarray(den, np.float64)
    # pixels whose camera ray missed (depth guide 0) are passed through, not learned
    hit = (np.asarray(features)[..., 9] != 0)[..., None]
    diff = np.where(hit, pred - t, 0.0)
    count = diff.size
    loss = (diff * diff / (d + REL_EPS)).sum() / count
    g = np.zeros(nparam)
    dout = 2.0 * diff / (d + REL_EPS) * np.exp(out) / count
    g[lay["head_w"]:lay["head_w"] + OUT * 32] = np.einsum("nhwk,nhwc->kc", dout, d0).reshape(-1)
    g[lay["head_b"]:lay["head_b"] + OUT] = dout.sum(axis=(0, 1, 2))
    # the head's input gradient is stored masked by the ReLU (d0 > 0)
    dd0 = q((dout @ Wh) * (d0 > 0)) if bf16 else dout @ Wh

    def back(name, dy):
        xshape, cols, y = cache[name]
        dz = dy * (y > 0)
        W, _ = P[name]
        dx, dW, db = conv_bwd(dz, cols, W, xshape)
        wo, bo, cin, cout = lay[name]
        g[wo:wo + dW.size] = dW.reshape(-1)
        g[bo:bo + cout] = db
        return q(dx)

    dcat = back("d0a", back("d0b", dd0))
    dd1, de0 = q(up_bwd(dcat[..., :64])), dcat[..., 64:]
    dcat = back("d1a", back("d1b", dd1))
    dd2, de1 = q(up_bwd(dcat[..., :96])), dcat[..., 96:]
    dcat = back("d2a", back("d2b", dd2))
    dbt, de2 = q(up_bwd(dcat[..., :128])), dcat[..., 128:]
    dp2 = back("b0", back("b1", dbt))
    de2 = q(de2 + pool_bwd(dp2))
    dp1 = back("e2a", back("e2b", de2))
    de1 = q(de1 + pool_bwd(dp1))
    dp0 = back("e1a", back("e1b", de1))
    de0 = q(de0 + pool_bwd(dp0))
    back("e0a", back("e0b", de0))
    return loss, g, out


def loss_only(params, features, target, den=None):
    return loss_and_grad(params, features, target, den)[0]
