"""Learned-denoiser pipeline: training pairs rendered by the path tracer, U-Net
training on the GPU, and the time-to-relMSE comparison of (path tracing +
denoiser) against plain path tracing and NLM.

relMSE(x, ref) = mean over pixels and channels of (x - ref)^2 / (ref^2 + 1e-2).
"""
from __future__ import annotations

import os
import time

import numpy as np

from . import abi, scenes

REL_EPS = 1e-2
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "unet_cornell.npz")


def rel_mse(x: np.ndarray, ref: np.ndarray) -> float:
    return float(np.mean((x - ref) ** 2 / (ref ** 2 + REL_EPS)))


class FrameRenderer:
    """Renders means (sum / spp) and guide buffers of one scene with device buffers."""

    def __init__(self, ctx, desc, width, height, max_depth=8, seed=0):
        import salty_amd
        self.ctx, self.w, self.h, self.depth, self.seed = ctx, width, height, max_depth, seed
        self.scene = salty_amd.Scene(ctx, desc)
        n = width * height
        self.d_acc = ctx.malloc(n * 12)
        self.d_alb = ctx.malloc(n * 12)
        self.d_nrm = ctx.malloc(n * 12)
        self.d_dep = ctx.malloc(n * 4)
        self.zero = np.zeros(n * 3, np.float32)

    def close(self):
        for p in (self.d_acc, self.d_alb, self.d_nrm, self.d_dep):
            self.ctx.free(p)
        self.scene.close()

    def cfg(self, spp, first_pass=0):
        return abi.RenderConfig(self.w, self.h, self.depth, spp, self.seed, 0, first_pass, 0, 0)

    def mean(self, spp, first_pass=0):
        """(mean image HxWx3, device milliseconds of the render)."""
        self.ctx.copy_to_device(self.d_acc, self.zero)
        st = self.scene.render_device(self.cfg(spp, first_pass), self.d_acc)
        img = self.ctx.download(self.d_acc, (self.h, self.w, 3), np.float32) / np.float32(spp)
        return img, st.milliseconds

    def guides(self):
        t = time.perf_counter()
        self.scene.guides_device(self.cfg(1), self.d_alb, self.d_nrm, self.d_dep)
        self.ctx.synchronize()
        ms = (time.perf_counter() - t) * 1e3
        a = self.ctx.download(self.d_alb, (self.h, self.w, 3), np.float32)
        n = self.ctx.download(self.d_nrm, (self.h, self.w, 3), np.float32)
        d = self.ctx.download(self.d_dep, (self.h, self.w, 1), np.float32)
        return a, n, d, ms


def features(mean_rgb, albedo, normal, depth) -> np.ndarray:
    return np.concatenate([mean_rgb, albedo, normal, depth], -1).astype(np.float32)


def make_dataset(ctx, seeds, width=256, height=256, spp_in=4, spp_ref=1024, max_depth=8, log=None):
    """Noisy/reference pairs of Cornell-style scenes: features (N,H,W,10), targets (N,H,W,3)."""
    feats, tgts = [], []
    for i, s in enumerate(seeds):
        fr = FrameRenderer(ctx, scenes.cornell_variant(int(s)), width, height, max_depth, seed=int(s))
        noisy, _ = fr.mean(spp_in)
        ref, _ = fr.mean(spp_ref, first_pass=1 << 20)
        a, n, d, _ = fr.guides()
        fr.close()
        feats.append(features(noisy, a, n, d))
        tgts.append(ref)
        if log and (i + 1) % 32 == 0:
            log(f"dataset {i + 1}/{len(seeds)}")
    return np.stack(feats), np.stack(tgts)


def augment(f, t, rng):
    if rng.random() < 0.5:
        f, t = f[:, :, ::-1], t[:, :, ::-1]
        f = f.copy()
        f[..., 6] = -f[..., 6]          # mirror the normal's x
    return np.ascontiguousarray(f), np.ascontiguousarray(t)


def train(net, ctx, feats, tgts, steps=1500, batch=16, lr=1e-3, seed=0, log=None):
    rng = np.random.default_rng(seed)
    n, h, w, _ = feats.shape
    d_f = ctx.malloc(batch * h * w * 10 * 4)
    d_t = ctx.malloc(batch * h * w * 3 * 4)
    losses = []
    t0 = time.perf_counter()
    for s in range(steps):
        idx = rng.choice(n, size=batch, replace=n < batch)
        f, t = augment(feats[idx], tgts[idx], rng)
        ctx.copy_to_device(d_f, f)
        ctx.copy_to_device(d_t, t)
        cur_lr = lr * (0.5 * (1 + np.cos(np.pi * s / steps)) * 0.9 + 0.1)
        losses.append(net.train_step_device(batch, h, w, d_f, d_t, float(cur_lr)))
        if log and (s + 1) % 100 == 0:
            log(f"step {s + 1} loss {np.mean(losses[-100:]):.4f} ({time.perf_counter() - t0:.1f}s)")
    ctx.synchronize()
    secs = time.perf_counter() - t0
    ctx.free(d_f)
    ctx.free(d_t)
    return losses, secs


def save_weights(net, path=WEIGHTS):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.savez_compressed(path, params=net.get_params().astype(np.float32))


def load_weights(net, path=WEIGHTS) -> bool:
    if not os.path.exists(path):
        return False
    with np.load(path, allow_pickle=False) as z:
        net.set_params(z["params"])
    return True


def time_to_relmse(ctx, net, desc, target=2e-3, width=256, height=256, gt_spp=4096, max_depth=8, seed=1234,
                   nlm=True):
    """Wall-clock (device) time for plain path tracing to reach relMSE <= target
    against the gt_spp render, and for path tracing + denoiser."""
    fr = FrameRenderer(ctx, desc, width, height, max_depth, seed=seed)
    gt, gt_ms = fr.mean(gt_spp, first_pass=1 << 22)
    a, n, d, guide_ms = fr.guides()
    # plain: progressive spp ladder, time of the render alone
    plain = []
    spp = 1
    plain_spp = None
    while spp <= gt_spp // 2:
        img, ms = fr.mean(spp)
        r = rel_mse(img, gt)
        plain.append((spp, ms, r))
        if r <= target and plain_spp is None:
            plain_spp = (spp, ms)
            break
        spp *= 2
    # refine the crossing with a linear-in-1/spp fit of the last two points
    den = []
    d_f = ctx.malloc(width * height * 10 * 4)
    d_o = ctx.malloc(width * height * 3 * 4)
    best = None
    for spp in (1, 2, 4, 8, 16, 32, 64):
        img, ms = fr.mean(spp)
        f = features(img, a, n, d)[None]
        ctx.copy_to_device(d_f, f)
        ctx.synchronize()
        t = time.perf_counter()
        net.infer_device(1, height, width, d_f, d_o)
        ctx.synchronize()
        inf_ms = (time.perf_counter() - t) * 1e3
        out = ctx.download(d_o, (height, width, 3), np.float32)
        r = rel_mse(out, gt)
        total = ms + guide_ms + inf_ms
        den.append((spp, ms, inf_ms, guide_ms, r))
        if r <= target and best is None:
            best = (spp, total)
    nl = None
    if nlm:
        img, ms = fr.mean(16)
        rgba = np.concatenate([img, np.ones(img.shape[:2] + (1,), np.float32)], -1)
        coeffs = []
        for c in (0.05, 0.1, 0.2, 0.4):
            t = time.perf_counter()
            out = ctx.nlm(rgba, c, 7, 2)
            coeffs.append((c, rel_mse(out[..., :3], gt), (time.perf_counter() - t) * 1e3))
        nl = dict(spp=16, render_ms=ms, results=coeffs)
    ctx.free(d_f)
    ctx.free(d_o)
    fr.close()
    return dict(gt_spp=gt_spp, gt_ms=gt_ms, plain=plain, plain_hit=plain_spp, denoised=den, denoised_hit=best,
                guide_ms=guide_ms, nlm=nl)
