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

    def render_ms(self, spp, first_pass=0):
        """Device milliseconds of one render of spp passes (the second of two,
        so that the first has set up the launch graph)."""
        ms = 0.0
        for _ in range(2):
            self.ctx.copy_to_device(self.d_acc, self.zero)
            ms = self.scene.render_device(self.cfg(spp, first_pass), self.d_acc).milliseconds
        return ms

    def ladder(self, levels, first_pass=0):
        """Means after each of the increasing sample counts `levels` of one
        progressive render (passes first_pass, first_pass + 1, ...):
        [(spp, image, cumulative device ms)]."""
        self.ctx.copy_to_device(self.d_acc, self.zero)
        out, done, ms = [], 0, 0.0
        for spp in levels:
            st = self.scene.render_device(self.cfg(spp - done, first_pass + done), self.d_acc)
            ms += st.milliseconds
            done = spp
            img = self.ctx.download(self.d_acc, (self.h, self.w, 3), np.float32) / np.float32(spp)
            out.append((spp, img, ms))
        return out

    def guides(self):
        """(albedo, normal, depth) on the host, and the device milliseconds of
        the guide render (HIP events on the context stream)."""
        import salty_amd
        timer = salty_amd.DeviceTimer(self.ctx)
        timer.start()
        self.scene.guides_device(self.cfg(1), self.d_alb, self.d_nrm, self.d_dep)
        ms = timer.stop()
        timer.close()
        a = self.ctx.download(self.d_alb, (self.h, self.w, 3), np.float32)
        n = self.ctx.download(self.d_nrm, (self.h, self.w, 3), np.float32)
        d = self.ctx.download(self.d_dep, (self.h, self.w, 1), np.float32)
        return a, n, d, ms


def features(mean_rgb, albedo, normal, depth) -> np.ndarray:
    return np.concatenate([mean_rgb, albedo, normal, depth], -1).astype(np.float32)


# (name, input channels with real data, output channels, resolution level) of
# the 3x3 convolutions; the first layer's 10 input channels are padded to 32.
CONV_LAYERS = [("e0a", 10, 32, 0), ("e0b", 32, 32, 0), ("e1a", 32, 64, 1), ("e1b", 64, 64, 1), ("e2a", 64, 96, 2),
               ("e2b", 96, 96, 2), ("b0", 96, 128, 3), ("b1", 128, 128, 3), ("d2a", 224, 96, 2),
               ("d2b", 96, 96, 2), ("d1a", 160, 64, 1), ("d1b", 64, 64, 1), ("d0a", 96, 32, 0), ("d0b", 32, 32, 0)]


def forward_flops(n: int, h: int, w: int) -> float:
    """Multiply-adds x 2 of one forward pass: the 3x3 convolutions at their
    resolution levels plus the 1x1 head (32 -> 3)."""
    px = [n * (h >> l) * (w >> l) for l in range(4)]
    conv = sum(2.0 * 9 * ci * co * px[l] for _, ci, co, l in CONV_LAYERS)
    return conv + 2.0 * 32 * 3 * px[0]


INPUT_LEVELS = (1, 2, 4, 8, 16, 32, 64, 128, 256)


class Dataset:
    """Noisy inputs at several sample counts, guides and reference of each scene:
    noisy (N, L, H, W, 3), guides (N, H, W, 7), target (N, H, W, 3)."""

    def __init__(self, noisy, guides, target, levels):
        self.noisy, self.guides, self.target, self.levels = noisy, guides, target, tuple(levels)

    def __len__(self):
        return len(self.target)


def make_dataset(ctx, seeds, width=256, height=256, levels=INPUT_LEVELS, spp_ref=1024, max_depth=8, log=None):
    """Training pairs of Cornell-style scenes.  The reference uses passes
    disjoint from the inputs', so its noise is independent of theirs."""
    noisy, guides, tgts = [], [], []
    for i, s in enumerate(seeds):
        fr = FrameRenderer(ctx, scenes.cornell_variant(int(s)), width, height, max_depth, seed=int(s))
        noisy.append(np.stack([img for _, img, _ in fr.ladder(levels)]))
        ref, _ = fr.mean(spp_ref, first_pass=1 << 20)
        a, n, d, _ = fr.guides()
        fr.close()
        guides.append(np.concatenate([a, n, d], -1))
        tgts.append(ref)
        if log and (i + 1) % 64 == 0:
            log(f"dataset {i + 1}/{len(seeds)}")
    return Dataset(np.stack(noisy), np.stack(guides), np.stack(tgts), levels)


def make_dataset_batched(ctx, seeds, width=256, height=256, levels=INPUT_LEVELS, spp_ref=4096, max_depth=8,
                         group=64, log=None):
    """make_dataset's pairs, rendered `group` scenes at a time as one scene
    set: each level of the input ladder and the reference are one batched
    render over the group's frames (a frame's sum is the one its scene alone
    would give, so the data equal make_dataset's), the guides one call per
    scene.  Small frames are launch-bound one at a time; batched, the
    4,096-spp references of 2,000 scenes take minutes instead of half an
    hour."""
    import salty_amd
    seeds = [int(s) for s in seeds]
    n, npx = len(seeds), width * height
    noisy = np.empty((n, len(levels), height, width, 3), np.float32)
    guides = np.empty((n, height, width, 7), np.float32)
    target = np.empty((n, height, width, 3), np.float32)
    zero = np.zeros(npx * 3, np.float32)
    for g0 in range(0, n, group):
        gs = seeds[g0:g0 + group]
        scs = [salty_amd.Scene(ctx, scenes.cornell_variant(s)) for s in gs]
        sset = salty_amd.SceneSet(ctx, scs)
        acc = [ctx.malloc(npx * 12) for _ in gs]
        d_alb, d_nrm, d_dep = ctx.malloc(npx * 12), ctx.malloc(npx * 12), ctx.malloc(npx * 4)
        try:
            # the input ladder: passes 0 .. levels[-1] - 1 of one progressive render
            for a in acc:
                ctx.copy_to_device(a, zero)
            done = 0
            for li, spp in enumerate(levels):
                sset.render_batch_device([(j, width, height, spp - done, s, done, acc[j]) for j, s in enumerate(gs)],
                                         max_depth)
                done = spp
                for j in range(len(gs)):
                    noisy[g0 + j, li] = ctx.download(acc[j], (height, width, 3), np.float32) / np.float32(spp)
            # the reference: passes disjoint from the inputs'
            for a in acc:
                ctx.copy_to_device(a, zero)
            sset.render_batch_device([(j, width, height, spp_ref, s, 1 << 20, acc[j]) for j, s in enumerate(gs)],
                                     max_depth)
            for j, (sc, s) in enumerate(zip(scs, gs)):
                target[g0 + j] = ctx.download(acc[j], (height, width, 3), np.float32) / np.float32(spp_ref)
                sc.guides_device(abi.RenderConfig(width, height, max_depth, 1, s, 0, 0, 0, 0), d_alb, d_nrm, d_dep)
                guides[g0 + j, ..., 0:3] = ctx.download(d_alb, (height, width, 3), np.float32)
                guides[g0 + j, ..., 3:6] = ctx.download(d_nrm, (height, width, 3), np.float32)
                guides[g0 + j, ..., 6:7] = ctx.download(d_dep, (height, width, 1), np.float32)
        finally:
            for a in acc + [d_alb, d_nrm, d_dep]:
                ctx.free(a)
            sset.close()
            for sc in scs:
                sc.close()
        if log:
            log(f"dataset {min(g0 + group, n)}/{n}")
    return Dataset(noisy, guides, target, levels)


def train(net, ctx, data: Dataset, steps=3000, batch=16, lr=1e-3, crop=None, seed=0, log=None):
    """Adam with a cosine learning-rate schedule down to 10 %.  The dataset is
    held in device memory (through torch, which only moves and indexes it);
    each step takes `batch` random (scene, input level) pairs, a random crop of
    side `crop` (whole frames when None) and random mirrors in x and y (the
    normal's component changes sign)."""
    import torch
    dev = torch.device("cuda")
    noisy = torch.from_numpy(data.noisy).to(dev)
    guides = torch.from_numpy(data.guides).to(dev)
    target = torch.from_numpy(data.target).to(dev)
    n, L, h, w, _ = data.noisy.shape
    c = crop or h
    gen = torch.Generator().manual_seed(seed)
    sign_x = torch.ones(7, device=dev)
    sign_x[3] = -1.0
    sign_y = torch.ones(7, device=dev)
    sign_y[4] = -1.0
    losses = []
    t0 = time.perf_counter()
    for s in range(steps):
        idx = torch.randint(n, (batch,), generator=gen).to(dev)
        lev = torch.randint(L, (batch,), generator=gen).to(dev)
        y0 = int(torch.randint(h - c + 1, (1,), generator=gen)) if c < h else 0
        x0 = int(torch.randint(w - c + 1, (1,), generator=gen)) if c < w else 0
        f_rgb = noisy[idx, lev, y0:y0 + c, x0:x0 + c]
        f_g = guides[idx, y0:y0 + c, x0:x0 + c]
        t = target[idx, y0:y0 + c, x0:x0 + c]
        fx = (torch.rand(batch, generator=gen) < 0.5).to(dev).view(-1, 1, 1, 1)
        fy = (torch.rand(batch, generator=gen) < 0.5).to(dev).view(-1, 1, 1, 1)
        f_rgb = torch.where(fx, f_rgb.flip(2), f_rgb)
        f_g = torch.where(fx, f_g.flip(2) * sign_x, f_g)
        t = torch.where(fx, t.flip(2), t)
        f_rgb = torch.where(fy, f_rgb.flip(1), f_rgb)
        f_g = torch.where(fy, f_g.flip(1) * sign_y, f_g)
        t = torch.where(fy, t.flip(1), t)
        f = torch.cat([f_rgb, f_g], -1).contiguous()
        t = t.contiguous()
        torch.cuda.synchronize()
        cur_lr = lr * (0.5 * (1 + np.cos(np.pi * s / steps)) * 0.9 + 0.1)
        losses.append(net.train_step_device(batch, c, c, f.data_ptr(), t.data_ptr(), float(cur_lr)))
        if log and (s + 1) % 500 == 0:
            log(f"step {s + 1} loss {np.mean(losses[-500:]):.4f} ({time.perf_counter() - t0:.1f}s)")
    ctx.synchronize()
    secs = time.perf_counter() - t0
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


def crossing(points, target):
    """Interpolated cost at which relMSE reaches `target` from [(cost, relMSE)]
    measured at increasing cost: log-log between the bracketing points, or a
    1/cost extrapolation past the last point (None if never below and not
    decreasing)."""
    pts = [(c, r) for c, r in points if c > 0 and r > 0]
    for (c0, r0), (c1, r1) in zip(pts, pts[1:]):
        if r0 > target >= r1:
            k = np.log(target / r0) / np.log(r1 / r0)
            return float(np.exp(np.log(c0) + k * (np.log(c1) - np.log(c0))))
    if pts and pts[0][1] <= target:
        return float(pts[0][0])
    if len(pts) >= 2 and pts[-1][1] < pts[-2][1]:
        c, r = pts[-1]
        return float(c * r / target)
    return None


def time_to_relmse(ctx, net, desc, target=2e-3, width=256, height=256, gt_spp=65536, max_depth=8, seed=1234,
                   plain_max=32768, den_levels=INPUT_LEVELS, nlm=True):
    """Device time for plain path tracing to reach relMSE <= target against
    a gt_spp render (passes disjoint from the others), and for path tracing +
    guides + U-Net inference.  Images come from one progressive render; the
    time of a sample count is that of one render of that many samples (timed
    after an untimed one), and times are interpolated between the measured
    sample counts (relMSE of plain path tracing falls as 1/spp).  The guide
    render and the feature assembly + U-Net are timed with HIP events on the
    context stream, like the renders."""
    import salty_amd
    fr = FrameRenderer(ctx, desc, width, height, max_depth, seed=seed)
    gt, gt_ms = fr.mean(gt_spp, first_pass=1 << 24)
    a, n, d, guide_ms = fr.guides()
    plain = []
    spp = 1
    levels = []
    while spp <= plain_max:
        levels.append(spp)
        spp *= 2
    for spp, img, _ in fr.ladder(levels):
        r = rel_mse(img, gt)
        plain.append((spp, fr.render_ms(spp), r))
        if r < target / 2:
            break
    plain_ms = crossing([(ms, r) for _, ms, r in plain], target)
    d_f = ctx.malloc(width * height * 10 * 4)
    d_o = ctx.malloc(width * height * 3 * 4)
    # two untimed inferences on these buffers: the first runs the launches
    # directly, the second captures them into the graph the timed ones replay
    # (as the renders are timed after an untimed one)
    ctx.copy_to_device(d_f, np.zeros(width * height * 10, np.float32))
    for _ in range(2):
        net.infer_device(1, height, width, d_f, d_o)
    ctx.synchronize()
    den = []
    timer = salty_amd.DeviceTimer(ctx)
    for spp, img, _ in fr.ladder(den_levels):
        # render_ms leaves the sum of the spp passes (the ladder's image times
        # spp, bit for bit) in the device accumulation buffer; the features
        # are assembled there and the U-Net runs on them, device-timed
        ms = fr.render_ms(spp)
        timer.start()
        ctx.features_device(1, height, width, fr.d_acc, spp, fr.d_alb, fr.d_nrm, fr.d_dep, d_f)
        net.infer_device(1, height, width, d_f, d_o)
        inf_ms = timer.stop()
        out = ctx.download(d_o, (height, width, 3), np.float32)
        den.append((spp, ms, inf_ms, rel_mse(out, gt)))
    timer.close()
    best = [(ms + guide_ms + inf, r) for spp, ms, inf, r in den]
    den_ms = None
    for tot, r in sorted(best):
        if r <= target:
            den_ms = tot
            break
    nl = None
    if nlm:
        img, ms = fr.mean(16)
        best, results = nlm_sweep(ctx, img, gt)
        nl = dict(spp=16, render_ms=ms, input_relmse=rel_mse(img, gt), best=best, results=results)
    ctx.free(d_f)
    ctx.free(d_o)
    fr.close()
    return dict(gt_spp=gt_spp, gt_ms=gt_ms, plain=plain, plain_ms=plain_ms, denoised=den, denoised_ms=den_ms,
                guide_ms=guide_ms, best_denoised_relmse=min(r for *_, r in den), nlm=nl)


# NLM filter strengths tried on radiance (the reference's FilterNLM input; its
# patch distance sums squared differences over 5 x 5 pixels and 4 channels, so
# on path-traced radiance up to ~15 the useful h spans a wide range) and on
# log(1 + radiance).
NLM_H_LINEAR = (0.125, 0.25, 0.5, 1.0, 2.0, 4.0, 8.0, 16.0, 32.0)
NLM_H_LOG = (0.0125, 0.025, 0.05, 0.1, 0.2, 0.4, 0.8)


def nlm_sweep(ctx, img, gt, search_radius=7, patch_radius=2):
    """NLM (s3dg_nlm, FilterNLM's weights) of a noisy mean image at every
    strength of NLM_H_LINEAR on radiance and of NLM_H_LOG on log(1 +
    radiance); returns (best (space, h, relMSE), [(space, h, relMSE)])."""
    ones = np.ones(img.shape[:2] + (1,), np.float32)
    results = []
    rgba = np.concatenate([img, ones], -1)
    for h in NLM_H_LINEAR:
        out = ctx.nlm(rgba, h, search_radius, patch_radius)
        results.append(("linear", h, rel_mse(out[..., :3], gt)))
    lg = np.concatenate([np.log1p(np.maximum(img, 0)), ones], -1)
    for h in NLM_H_LOG:
        out = ctx.nlm(lg, h, search_radius, patch_radius)
        results.append(("log", h, rel_mse(np.expm1(out[..., :3]), gt)))
    best = min(results, key=lambda r: r[2])
    return best, results


def nlm_device_ms(ctx, width, height, search_radius=7, patch_radius=2, reps=3):
    """Device time of one NLM pass at this size (host clock around the call
    and a synchronisation, the best of reps after one untimed pass)."""
    d_in = ctx.malloc(width * height * 16)
    d_out = ctx.malloc(width * height * 16)
    try:
        ctx.copy_to_device(d_in, np.ones(width * height * 4, np.float32))
        ctx.nlm_device(width, height, 1.0, search_radius, patch_radius, d_in, d_out)
        ctx.synchronize()
        best = None
        for _ in range(reps):
            t = time.perf_counter()
            ctx.nlm_device(width, height, 1.0, search_radius, patch_radius, d_in, d_out)
            ctx.synchronize()
            ms = (time.perf_counter() - t) * 1e3
            best = ms if best is None else min(best, ms)
        return best
    finally:
        ctx.free(d_in)
        ctx.free(d_out)


def evaluate_heldout(ctx, net, seeds, width=256, height=256, spp_in=4, gt_spp=65536, plain_spp=256, max_depth=8,
                     log=None):
    """The learned denoiser against NLM and plain rendering on held-out
    Cornell-style scenes (seeds outside the training range): per scene the
    relMSE against a gt_spp render of the spp_in-sample input, of the U-Net
    on that input with the guides, of NLM on that input at its best strength
    (nlm_sweep), and of a plain plain_spp render, with the device times of
    each path."""
    import salty_amd
    d_f = ctx.malloc(width * height * 10 * 4)
    d_o = ctx.malloc(width * height * 3 * 4)
    nlm_ms = nlm_device_ms(ctx, width, height)
    timer = salty_amd.DeviceTimer(ctx)
    rows = []
    try:
        for s in seeds:
            fr = FrameRenderer(ctx, scenes.cornell_variant(int(s)), width, height, max_depth, seed=int(s))
            gt, _ = fr.mean(gt_spp, first_pass=1 << 24)
            a, n, d, guide_ms = fr.guides()
            img, _ = fr.mean(spp_in)
            # the sum of the spp_in passes stays in the device accumulation
            # buffer: features assembled there, then the U-Net (the untimed
            # run captures the graph the timed one replays)
            in_ms = fr.render_ms(spp_in)
            ctx.features_device(1, height, width, fr.d_acc, spp_in, fr.d_alb, fr.d_nrm, fr.d_dep, d_f)
            net.infer_device(1, height, width, d_f, d_o)
            timer.start()
            ctx.features_device(1, height, width, fr.d_acc, spp_in, fr.d_alb, fr.d_nrm, fr.d_dep, d_f)
            net.infer_device(1, height, width, d_f, d_o)
            inf_ms = timer.stop()
            den = ctx.download(d_o, (height, width, 3), np.float32)
            best, _ = nlm_sweep(ctx, img, gt)
            plain, _ = fr.mean(plain_spp, first_pass=1 << 16)
            plain_ms = fr.render_ms(plain_spp, first_pass=1 << 16)
            fr.close()
            row = dict(seed=int(s), input_relmse=rel_mse(img, gt), learned_relmse=rel_mse(den, gt),
                       nlm_relmse=best[2], nlm_space=best[0], nlm_h=best[1], plain_relmse=rel_mse(plain, gt),
                       input_ms=in_ms, guides_ms=guide_ms, inference_ms=inf_ms, nlm_ms=nlm_ms, plain_ms=plain_ms)
            rows.append(row)
            if log:
                log(row)
    finally:
        timer.close()
        ctx.free(d_f)
        ctx.free(d_o)
    return rows
