"""Benchmark of the rendering hot path on one MI355X.

Workload: the procedural mesh scene (~1.05 M triangles: displaced level-7
icosphere, 64 instanced level-3 icospheres, 640 k-triangle ground grid, one
spherical light) rendered at 1920x1080, 64 samples per pixel, max depth 12,
with the wavefront path tracer.  One step = one full 64-spp frame.

Prints one JSON line: Grays/s (extension + shadow rays traced per second) as
`value`, timed over whole frames whose bounce loops are replayed from the
captured HIP graph (the library's default path); paths/s beside it; the
roofline of the dominant kernel (closest-hit traversal) from HIP-event timings
of every launch of one more, instrumented frame; the time to relMSE 2e-3 of the
learned denoiser against plain path tracing; short runs of the other
configurations under config.legs (4 M-triangle soup x 16 M rays closest and
any hit, the 64-scene batch, the 1080p U-Net with NLM beside it); and the CPU
oracle (the restatement of the reference's CPU path) timed on a bounded sample
of the same frame on every host core the job may use.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E
NODE_BYTES = 128               # compressed 8-wide node, one line
TRI_BYTES = 36                 # v0, e0, e1 as Triangle::IsHit reads them
HIT_BYTES = 20                 # hit record written per extension ray
RAY_IN_BYTES = 4 + 24          # queue index, ray origin+direction (read from the
                               # second bounce on; the first bounce computes its
                               # camera rays from the path slot)


def ext_io_bytes(extension_rays, paths):
    """Average bytes per extension ray besides nodes and triangles."""
    later = max(extension_rays - paths, 0)
    return HIT_BYTES + RAY_IN_BYTES * later / max(extension_rays, 1)


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "k_extend_mesh_scene_hbm_traffic.md")


def extend_traffic_bytes():
    """HBM bytes per k_extend launch of the default workload, from the committed
    counter profile (FETCH_SIZE and WRITE_SIZE scaled by a plain copy), or None."""
    rows = []
    try:
        with open(TRAFFIC_PROFILE) as f:
            rows = [line for line in f if line.startswith("| `k_extend<false")]
    except OSError:
        pass
    # the instrumented frame's row (the frame the kernel split is timed on)
    rows.sort(key=lambda line: "instrumented frame" not in line)
    return float(rows[0].strip().strip("|").split("|")[-1]) * 1e6 if rows else None


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=3)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--width", type=int, default=1920)
    p.add_argument("--height", type=int, default=1080)
    p.add_argument("--spp", type=int, default=64)
    p.add_argument("--depth", type=int, default=12)
    p.add_argument("--seed", type=int, default=0)
    p.add_argument("--cpu-seconds", type=float, default=12.0, help="target duration of the CPU baseline sample")
    p.add_argument("--passes-in-flight", type=int, default=0,
                   help="passes traced together per wavefront (0: the library's automatic choice)")
    p.add_argument("--no-autotune", action="store_true", help="keep the library's default kernel settings")
    p.add_argument("--tuning", default=None,
                   help="mesh workload: kernel settings 'variant,refill,tri_batch,fetch_chunk,tail_paths' (as a "
                        "default run's config.tuning reports them) instead of autotuning, e.g. for profiling")
    p.add_argument("--materials", choices=("c2", "all"), default="c2",
                   help="mesh workload materials: Lambert/mirror/glass, or all five (Phong and Plastic too)")
    p.add_argument("--lens-radius", type=float, default=0.0, help="mesh workload: thin-lens camera when > 0")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--no-legs", action="store_true",
                   help="default workload only: skip the short triangle-soup, 64-scene batch and U-Net runs "
                        "reported beside the headline")
    p.add_argument("--no-denoise", action="store_true",
                   help="skip the time-to-relMSE comparison (learned denoiser vs plain path tracing)")
    p.add_argument("--workload", choices=("mesh", "soup", "unet", "batch", "denoise"), default="mesh",
                   help="mesh: the path-traced mesh scene (default); soup: closest/any-hit rays against a "
                        "triangle soup; unet: denoiser inference at 1920x1080 and training steps; batch: 64 "
                        "uneven scenes rendered together; denoise: only the time-to-relMSE comparison on the "
                        "Cornell box (part of the default line)")
    p.add_argument("--heldout", type=int, default=32,
                   help="denoise workload: held-out scenes evaluated against NLM and plain 256-spp rendering")
    p.add_argument("--gt-spp", type=int, default=65536, help="denoise workload: ground-truth samples per pixel")
    p.add_argument("--soup-tris", type=int, default=4 << 20)
    p.add_argument("--soup-rays", type=int, default=16 << 20)
    p.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                   help="only run a plain device copy of MIB MiB (reference point for HBM byte counters) and exit")
    return p.parse_args()


def copy_calibration(mib: int) -> None:
    """One device-to-device copy kernel of a known size: FETCH_SIZE / WRITE_SIZE
    of this dispatch against its mib MiB give the counters' scale."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_mib": mib, "checksum": float(b[::4096].sum().item())}), flush=True)


def host_cpu():
    """Host cores this process may run on and the CPU model.  The reference
    runs its rows on every core it has (s3d_pt.cpp:311-321); the baseline uses
    every core the job is given: the CPUs of its affinity mask, further
    limited by OMP_NUM_THREADS when the environment sets it (a GPU box grants
    one GPU's job a share of the host)."""
    nproc = os.cpu_count() or 1
    try:
        allowed = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        allowed = nproc
    threads = allowed
    omp = os.environ.get("OMP_NUM_THREADS", "")
    if omp.isdigit() and int(omp) > 0:
        threads = min(threads, int(omp))
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"threads": max(1, threads), "nproc": nproc, "affinity_cpus": allowed,
            "omp_num_threads": omp or None, "cpu_model": model}


def cpu_baseline(desc, args):
    """Times the oracle (C++ restatement of PathTracer::Radiance with the
    reference traversal, OpenMP over rows as s3d_pt.cpp:311-313) on a band of
    rows at 1 spp, sized to about args.cpu_seconds of work."""
    from oracle import oracle as O
    hc = host_cpu()
    threads = hc["threads"]
    sc = O.OracleScene(desc)
    W, H = args.width, args.height
    mid = H // 2
    # probe 8 centre rows at 1 spp, then size the timed sample to the target
    t = time.perf_counter()
    sc.render(W, H, 1, args.depth, seed=args.seed, rows=(mid - 4, mid + 4), threads=threads)
    per_row = max((time.perf_counter() - t) / 8, 1e-6)
    rows_needed = args.cpu_seconds / per_row
    if rows_needed < H:
        rows, spp = max(8, int(rows_needed)), 1
    else:
        rows, spp = H, max(1, int(round(rows_needed / H)))
    r0 = mid - rows // 2
    t = time.perf_counter()
    _, counts = sc.render(W, H, spp, args.depth, seed=args.seed, rows=(r0, r0 + rows), threads=threads)
    total_t = time.perf_counter() - t
    total_rays = int(counts[0] + counts[1])
    total_paths = int(counts[2])
    return {
        "value": total_rays / total_t / 1e9,
        "unit": "Grays/s",
        "paths_per_s": total_paths / total_t,
        "cores": threads,
        "kind": "port",
        "sample": f"rows {r0}-{r0 + rows - 1} of the {W}x{H} frame, {spp} spp, depth {args.depth} "
                  f"({total_paths} paths, {total_rays} rays, {total_t:.1f} s on {threads} threads)",
        "host": hc,
    }


def soup_cpu_baseline(desc, rays, seconds):
    """The oracle's reference traversal (no clipping, exact slab tests) on a
    prefix of the ray batch sized to about `seconds`, on the host's cores."""
    from oracle import oracle as O
    hc = host_cpu()
    threads = hc["threads"]
    mesh = O.OracleMesh(desc.meshes[0].vertices, desc.meshes[0].bvh_kind)
    t = time.perf_counter()
    mesh.trace(rays[:256], threads=threads)
    per_ray = max((time.perf_counter() - t) / 256, 1e-9)
    m = int(min(len(rays), max(1024, seconds / per_ray)))
    t = time.perf_counter()
    mesh.trace(rays[:m], threads=threads)
    el = time.perf_counter() - t
    return {"value": m / el / 1e9, "unit": "Grays/s", "cores": threads, "kind": "port",
            "sample": f"first {m} rays of the batch, closest hit ({el:.1f} s on {threads} threads)",
            "host": hc}


def soup_main(args):
    print(json.dumps(soup_run(args, args.steps, args.warmup, None if args.no_cpu_baseline else args.cpu_seconds)),
          flush=True)


def soup_run(args, steps, warmup, cpu_seconds):
    """Closest-hit and any-hit throughput of the batch traversal kernel on a
    triangle soup (v0 ~ U[0,1]^3, edges of length N(0.01, 0.003) in uniform
    directions) with rays from U[-0.1, 1.1]^3 in uniform directions,
    tmin 1e-4, tmax infinity.  Returns the JSON line as a dict."""
    import salty_amd
    from salty_amd import abi, scenes
    t = time.perf_counter()
    desc = scenes.triangle_soup(args.soup_tris, seed=args.seed)
    rays = scenes.soup_rays(args.soup_rays, seed=args.seed + 1)
    gen_s = time.perf_counter() - t
    ctx = salty_amd.Context(0)
    t = time.perf_counter()
    scene = salty_amd.Scene(ctx, desc)
    build_s = time.perf_counter() - t
    n = len(rays)
    d_rays = ctx.upload(rays)
    d_hits = ctx.malloc(n * 16)
    d_cnt = ctx.upload(np.zeros(2, np.uint64))
    scene.trace_counted_device(d_rays, n, abi.TRACE_CLOSEST, d_hits, d_cnt)
    cnt = ctx.download(d_cnt, (2,), np.uint64)
    nodes, tris = float(cnt[0]) / n, float(cnt[1]) / n
    res, picked = {}, {}
    base = ctx.get_tuning()
    for mode, name in ((abi.TRACE_CLOSEST, "closest"), (abi.TRACE_ANY, "any")):
        # the refill threshold is the traversal setting that matters for a
        # batch of incoherent rays: one timed trace per candidate, the
        # fastest kept (results do not depend on it)
        if not args.no_autotune:
            best = None
            for refill in (16, 24, 32, 48):
                ctx.set_tuning(abi.Tuning(base.traversal_variant, refill, base.tri_batch, base.passes_in_flight,
                                          base.fetch_chunk, base.leaf_size, base.tail_paths))
                scene.trace_device(d_rays, n, mode, d_hits)
                ctx.synchronize()
                t = time.perf_counter()
                scene.trace_device(d_rays, n, mode, d_hits)
                ctx.synchronize()
                dt = time.perf_counter() - t
                if best is None or dt < best[0]:
                    best = (dt, refill)
            ctx.set_tuning(abi.Tuning(base.traversal_variant, best[1], base.tri_batch, base.passes_in_flight,
                                      base.fetch_chunk, base.leaf_size, base.tail_paths))
            picked[name] = best[1]
        for _ in range(warmup):
            scene.trace_device(d_rays, n, mode, d_hits)
        ctx.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            scene.trace_device(d_rays, n, mode, d_hits)
        ctx.synchronize()
        res[name] = (time.perf_counter() - t) / steps
    bytes_per_ray = nodes * NODE_BYTES + tris * TRI_BYTES + 32 + 16
    achieved = bytes_per_ray * n / res["closest"] / 1e9
    # the same closest-hit trace through the host-buffer entry point (rays
    # in and hits out over PCIe): reported beside the line, never its value
    t = time.perf_counter()
    scene.trace(rays, abi.TRACE_CLOSEST)
    host_s = time.perf_counter() - t
    line = {
        "metric": "Grays/s, closest-hit batch traversal (any-hit beside it)",
        "value": n / res["closest"] / 1e9,
        "unit": "Grays/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": res["closest"] * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded triangle soup and rays)",
        "config": {"workload": f"triangle soup, {args.soup_tris} triangles, {n} rays, tmin 1e-4, tmax inf",
                   "any_hit_grays_per_s": n / res["any"] / 1e9, "any_hit_ms": res["any"] * 1e3,
                   "host_buffer_grays_per_s": n / host_s / 1e9, "host_buffer_ms": host_s * 1e3,
                   "scene_build_s": build_s, "generate_s": gen_s,
                   "device_scene_bytes": scene.stats().device_bytes,
                   "refill": picked or None},
        "roofline": {"kernel": "k_trace<closest> (batch traversal of the compressed BVH8)", "bound": "hbm",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": None, "bytes_per_ray": bytes_per_ray, "nodes_per_ray": nodes,
                     "tris_per_ray": tris},
    }
    if cpu_seconds:
        line["cpu_baseline"] = soup_cpu_baseline(desc, rays, cpu_seconds)
    for p in (d_rays, d_hits, d_cnt):
        ctx.free(p)
    scene.close()
    ctx.close()
    return line


BF16_PEAK_TFLOPS = 2500.0      # MI355X dense BF16 MFMA


def unet_cpu_baseline(seconds):
    """The float64 NumPy U-Net forward of the oracle on a 64x64 crop, repeated
    to about `seconds`, in FLOP/s."""
    from oracle import unet
    from salty_amd import denoise
    rng = np.random.default_rng(0)
    _, nparam = unet.param_layout()
    p = rng.normal(scale=0.05, size=nparam)
    f = np.abs(rng.normal(size=(1, 64, 64, 10)))
    t = np.abs(rng.normal(size=(1, 64, 64, 3)))
    reps, t0 = 0, time.perf_counter()
    while True:
        unet.loss_and_grad(p, f, t)   # forward + backward; counted as 3 forwards
        reps += 1
        el = time.perf_counter() - t0
        if el >= seconds:
            break
    flops = reps * 3 * denoise.forward_flops(1, 64, 64)
    return {"value": flops / el / 1e12, "unit": "TFLOP/s", "cores": 1, "kind": "port",
            "sample": f"{reps} forward+backward passes of a 64x64 crop in float64 NumPy ({el:.1f} s)"}


def unet_main(args):
    print(json.dumps(unet_run(args, args.steps, args.warmup, None if args.no_cpu_baseline else args.cpu_seconds)),
          flush=True)


def unet_run(args, steps, warmup, cpu_seconds):
    """Denoiser inference on one 1920x1080 10-channel frame, training steps of
    16 x 256^2 crops, and NLM (search radius 7, patch radius 2) on the frame."""
    import salty_amd
    from salty_amd import denoise
    ctx = salty_amd.Context(0)
    net = salty_amd.Denoiser(ctx, seed=0)
    denoise.load_weights(net)
    rng = np.random.default_rng(args.seed)
    W, H = args.width, args.height
    feat = np.concatenate([rng.gamma(1.0, 0.5, (1, H, W, 3)), rng.uniform(0, 1, (1, H, W, 3)),
                           rng.normal(size=(1, H, W, 3)), rng.uniform(1, 5, (1, H, W, 1))], -1).astype(np.float32)
    d_f = ctx.upload(feat)
    d_o = ctx.malloc(H * W * 3 * 4)
    for _ in range(warmup):
        net.infer_device(1, H, W, d_f, d_o)
    ctx.synchronize()
    t = time.perf_counter()
    for _ in range(steps):
        net.infer_device(1, H, W, d_f, d_o)
    ctx.synchronize()
    inf_s = (time.perf_counter() - t) / steps
    # NLM on the same frame (RGBA)
    rgba = np.concatenate([feat[0, ..., :3], np.ones((H, W, 1), np.float32)], -1)
    d_in = ctx.upload(rgba)
    d_nl = ctx.malloc(rgba.nbytes)
    ctx.nlm_device(W, H, 0.2, 7, 2, d_in, d_nl)
    ctx.synchronize()
    t = time.perf_counter()
    for _ in range(steps):
        ctx.nlm_device(W, H, 0.2, 7, 2, d_in, d_nl)
    ctx.synchronize()
    nlm_s = (time.perf_counter() - t) / steps
    # training steps at 16 x 256^2
    B, S = 16, 256
    tf = np.ascontiguousarray(feat[0, :S, :S][None].repeat(B, 0))
    d_tf = ctx.upload(tf)
    d_tt = ctx.upload(np.ascontiguousarray(tf[..., :3]))
    net.train_step_device(B, S, S, d_tf, d_tt, 1e-6)
    ctx.synchronize()
    t = time.perf_counter()
    for _ in range(steps):
        net.train_step_device(B, S, S, d_tf, d_tt, 1e-6)
    ctx.synchronize()
    train_s = (time.perf_counter() - t) / steps
    fl = denoise.forward_flops(1, H, W)
    achieved = fl / inf_s / 1e12
    line = {
        "metric": "U-Net denoiser inference TFLOP/s at 1920x1080 (training steps and NLM beside it)",
        "value": achieved,
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": inf_s * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16",
        "data": "synthetic (random 10-channel frame)",
        "config": {"workload": f"4-level U-Net 32/64/96/128, {W}x{H} 10-channel input, BF16 MFMA, FP32 accumulation",
                   "inference_gflop": fl / 1e9,
                   "train_step_ms": train_s * 1e3, "train_step_gflop": 3 * denoise.forward_flops(B, S, S) / 1e9,
                   "train_tflops": 3 * denoise.forward_flops(B, S, S) / train_s / 1e12,
                   "nlm_ms": nlm_s * 1e3, "nlm_radii": [7, 2]},
        "roofline": {"kernel": "whole inference (3x3 convolutions dominate)", "bound": "mfma",
                     "achieved": achieved, "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / BF16_PEAK_TFLOPS, "traffic": None},
    }
    if cpu_seconds:
        line["cpu_baseline"] = unet_cpu_baseline(min(cpu_seconds, 10.0))
    for p in (d_f, d_o, d_in, d_nl, d_tf, d_tt):
        ctx.free(p)
    net.close()
    ctx.close()
    return line


def time_to_relmse_leg(ctx):
    """Device time to relMSE <= 2e-3 against a 65,536-spp ground truth on the
    36-triangle Cornell box (256x256, depth 8): plain path tracing against path
    tracing + guides + the trained U-Net (weights in salty_amd/weights)."""
    import salty_amd
    from salty_amd import denoise, scenes
    net = salty_amd.Denoiser(ctx, seed=0)
    if not denoise.load_weights(net):
        net.close()
        return {"error": "no trained weights"}
    r = denoise.time_to_relmse(ctx, net, scenes.cornell_box(), target=2e-3, width=256, height=256,
                               gt_spp=65536, max_depth=8, nlm=True)
    net.close()
    pick = None
    for spp, ms, inf, rel in r["denoised"]:
        if rel <= 2e-3 and (pick is None or ms + inf < pick[1] + pick[2]):
            pick = (spp, ms, inf, rel)
    out = {"scene": "Cornell box, 36 triangles, 256x256, depth 8", "target_relmse": 2e-3,
           "ground_truth_spp": r["gt_spp"], "plain_ms": r["plain_ms"], "denoised_ms": r["denoised_ms"],
           "best_denoised_relmse": r["best_denoised_relmse"]}
    if pick:
        out.update(denoised_input_spp=pick[0], render_ms=pick[1], inference_ms=pick[2], guides_ms=r["guide_ms"],
                   denoised_relmse=pick[3])
    if r["plain_ms"] and r["denoised_ms"]:
        out["speedup"] = r["plain_ms"] / r["denoised_ms"]
    if r["nlm"]:
        out["nlm_16spp"] = {"input_relmse": r["nlm"]["input_relmse"], "best_space": r["nlm"]["best"][0],
                            "best_h": r["nlm"]["best"][1], "best_relmse": r["nlm"]["best"][2]}
    return out


def heldout_leg(ctx, count, gt_spp):
    """The learned denoiser on `count` held-out scenes (seeds 3000..) from 4
    spp, against NLM (search radius 7, patch radius 2, best strength) on the
    same input and plain 256-spp rendering, all scored against gt_spp renders."""
    import salty_amd
    from salty_amd import denoise
    net = salty_amd.Denoiser(ctx, seed=0)
    if not denoise.load_weights(net):
        net.close()
        return {"error": "no trained weights"}
    rows = denoise.evaluate_heldout(ctx, net, range(3000, 3000 + count), gt_spp=gt_spp,
                                    log=lambda r: print(json.dumps(r), file=sys.stderr, flush=True))
    net.close()
    keys = ("input_relmse", "learned_relmse", "nlm_relmse", "plain_relmse", "input_ms", "guides_ms",
            "inference_ms", "nlm_ms", "plain_ms")
    mean = {k: float(np.mean([r[k] for r in rows])) for k in keys}
    return {"scenes": count, "seeds": [3000, 3000 + count - 1], "resolution": [256, 256], "input_spp": 4,
            "plain_spp": 256, "ground_truth_spp": gt_spp, "nlm_radii": [7, 2], "mean": mean,
            "learned_at_or_below_2e-3": sum(r["learned_relmse"] <= 2e-3 for r in rows),
            "learned_beats_nlm": sum(r["learned_relmse"] < r["nlm_relmse"] for r in rows),
            "learned_beats_plain_256spp": sum(r["learned_relmse"] < r["plain_relmse"] for r in rows),
            "per_scene": [{k: (round(v, 6) if isinstance(v, float) else v) for k, v in r.items()} for r in rows]}


def denoise_main(args):
    """The time-to-relMSE comparison (speedup of path tracing + guides +
    U-Net over plain path tracing in reaching relMSE 2e-3, device time) and
    the held-out evaluation against NLM and plain 256-spp rendering."""
    import salty_amd
    ctx = salty_amd.Context(0)
    t = time_to_relmse_leg(ctx)
    if args.heldout:
        t["heldout"] = heldout_leg(ctx, args.heldout, args.gt_spp)
    line = {
        "metric": "time to relMSE 2e-3, learned denoiser vs plain path tracing (speedup), 1 GPU",
        "value": t.get("speedup"),
        "unit": "x",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 1,
        "ms_per_step": t.get("denoised_ms"),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+bf16",
        "data": "synthetic (procedural scene, trained weights in salty_amd/weights)",
        "config": {"workload": t.get("scene"), "time_to_relmse": t},
    }
    print(json.dumps(line), flush=True)
    ctx.close()


def batch_main(args):
    print(json.dumps(batch_run(args, args.steps, args.warmup)), flush=True)


def batch_run(args, steps, warmup):
    """64 seeded scenes (1k-500k triangles, log-uniform), each at a resolution
    from {320x240, 640x480, 1280x720, 1920x1080} and 4, 16 or 64 spp, depth 8,
    rendered as one batch: all frames' paths share every launch."""
    import salty_amd
    from salty_amd import abi, scenes
    t = time.perf_counter()
    cfg = scenes.batch_config(args.seed)
    gen_s = time.perf_counter() - t
    ctx = salty_amd.Context(0)
    t = time.perf_counter()
    scs = [salty_amd.Scene(ctx, d) for d, _, _, _ in cfg]
    sset = salty_amd.SceneSet(ctx, scs)
    build_s = time.perf_counter() - t
    tris = sum(sc.stats().triangles for sc in scs)
    bufs = [ctx.malloc(w * h * 12) for _, w, h, _ in cfg]
    frames = [(k, w, h, spp, k, 0, bufs[k]) for k, (_, w, h, spp) in enumerate(cfg)]
    depth = 8
    # work counts at 1 spp, outside the timed region
    cst = sset.render_batch_device([(k, w, h, 1, k, 1000, d) for k, w, h, _, _, _, d in frames], depth,
                                   flags=abi.RENDER_COUNT_WORK)
    ext_nodes = cst.extend_nodes / max(cst.extension_rays, 1)
    ext_tris = cst.extend_tris / max(cst.extension_rays, 1)
    for _ in range(warmup):
        sset.render_batch_device(frames, depth)
    ctx.synchronize()
    # timed steps: the graph-replayed bounce loops, as a caller runs them
    rays_ext = rays_sh = paths = 0
    t0 = time.perf_counter()
    for _ in range(steps):
        st = sset.render_batch_device(frames, depth)
        rays_ext += st.extension_rays
        rays_sh += st.shadow_rays
        paths += st.paths
    ctx.synchronize()
    elapsed = time.perf_counter() - t0
    # kernel split from one more step with an event pair around every launch
    kst = sset.render_batch_device(frames, depth, flags=abi.RENDER_TIME_KERNELS)
    launches = kst.extend_launches
    avg_ext_ms = kst.extend_ms / max(launches, 1)
    bytes_per_ray = ext_nodes * NODE_BYTES + ext_tris * TRI_BYTES + ext_io_bytes(kst.extension_rays, kst.paths)
    achieved = (bytes_per_ray * kst.extension_rays / max(launches, 1) / (avg_ext_ms * 1e-3) / 1e9
                if avg_ext_ms else 0.0)
    line = {
        "metric": "Grays/s + paths/s, 64 uneven scenes rendered as one batch, 1 GPU",
        "value": (rays_ext + rays_sh) / elapsed / 1e9,
        "unit": "Grays/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": elapsed * 1e3 / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded procedural scenes)",
        "config": {"workload": f"{len(cfg)} scenes, {tris} triangles in all, resolutions 320x240..1920x1080, "
                               f"4/16/64 spp, depth 8, one scene set",
                   "paths_per_s": paths / elapsed, "paths_per_step": paths // max(steps, 1),
                   "scene_build_s": build_s, "scene_generate_s": gen_s},
        "roofline": {"kernel": "k_extend (closest-hit traversal of the compressed BVH8)", "bound": "hbm",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": None, "avg_launch_ms": avg_ext_ms, "launches": launches,
                     "bytes_per_ray": bytes_per_ray, "nodes_per_ray": ext_nodes, "tris_per_ray": ext_tris},
    }
    for d in bufs:
        ctx.free(d)
    sset.close()
    for sc in scs:
        sc.close()
    ctx.close()
    return line


def main():
    args = parse()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.workload == "soup":
        if args.gpus != 1:
            print("bench.py runs on one GPU only", file=sys.stderr)
            sys.exit(2)
        soup_main(args)
        return
    if args.workload == "batch":
        if args.gpus != 1:
            print("bench.py runs on one GPU only", file=sys.stderr)
            sys.exit(2)
        batch_main(args)
        return
    if args.workload == "unet":
        if args.gpus != 1:
            print("bench.py runs on one GPU only", file=sys.stderr)
            sys.exit(2)
        unet_main(args)
        return
    if args.workload == "denoise":
        if args.gpus != 1:
            print("bench.py runs on one GPU only", file=sys.stderr)
            sys.exit(2)
        denoise_main(args)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the path shards by pixel tiles x "
              f"sample ranges with a replicated BVH, and only the one-GPU share is implemented.", file=sys.stderr)
        sys.exit(2)

    import salty_amd
    from salty_amd import abi, scenes

    t = time.perf_counter()
    desc = scenes.mesh_scene(seed=args.seed, materials=args.materials, lens_radius=args.lens_radius)
    gen_s = time.perf_counter() - t
    ctx = salty_amd.Context(0)
    t = time.perf_counter()
    scene = salty_amd.Scene(ctx, desc)
    build_s = time.perf_counter() - t
    sstats = scene.stats()
    W, H = args.width, args.height
    d_accum = ctx.malloc(W * H * 3 * 4)
    zero = np.zeros(W * H * 3, np.float32)
    ctx.copy_to_device(d_accum, zero)

    def cfg(spp, flags=0, first_pass=0):
        return abi.RenderConfig(W, H, args.depth, spp, args.seed, 0, first_pass, flags, args.passes_in_flight)

    # work counts (nodes visited / triangles tested per ray), outside the timed region
    cst = scene.render_device(cfg(2, abi.RENDER_COUNT_WORK, first_pass=1000), d_accum)
    ext_nodes = cst.extend_nodes / max(cst.extension_rays, 1)
    ext_tris = cst.extend_tris / max(cst.extension_rays, 1)
    ext_spills = cst.extend_stack_spills / max(cst.extension_rays, 1)

    tuning = None
    if args.tuning:
        v, r, tb, fc, tail = (int(x) for x in args.tuning.split(","))
        base = ctx.get_tuning()
        ctx.set_tuning(abi.Tuning(v, r, tb, base.passes_in_flight, fc, base.leaf_size, tail))
        tuning = {"traversal_variant": v, "refill": r, "tri_batch": tb, "fetch_chunk": fc, "tail_paths": tail,
                  "source": "--tuning"}
    elif not args.no_autotune:
        t, _ = scene.autotune(cfg(args.spp))
        tuning = {"traversal_variant": t.traversal_variant, "refill": t.refill, "tri_batch": t.tri_batch,
                  "fetch_chunk": t.fetch_chunk, "tail_paths": t.tail_paths}
    for w in range(args.warmup):
        scene.render_device(cfg(args.spp, first_pass=0), d_accum)
    ctx.synchronize()

    # timed steps: each a whole frame, its bounce loop replayed from the
    # captured HIP graph as a caller runs it
    rays_ext = rays_sh = paths = 0
    device_ms = 0.0
    t0 = time.perf_counter()
    for s in range(args.steps):
        st = scene.render_device(cfg(args.spp, first_pass=0), d_accum)
        device_ms += st.milliseconds
        rays_ext += st.extension_rays
        rays_sh += st.shadow_rays
        paths += st.paths
    ctx.synchronize()
    elapsed = time.perf_counter() - t0
    # the kernel split: one more frame with a HIP event pair around every
    # launch (on the context's stream, where the kernels run)
    kst = scene.render_device(cfg(args.spp, abi.RENDER_TIME_KERNELS, first_pass=0), d_accum)
    launches = kst.extend_launches
    # the same frame through the host-buffer entry point (s3dg_render copies
    # the accumulation buffer in and out over PCIe): reported beside the
    # line, never as its value
    host_acc = np.zeros((H, W, 3), np.float32)
    t = time.perf_counter()
    _, hst = scene.render(W, H, args.spp, args.depth, seed=args.seed, accum=host_acc,
                          passes_in_flight=args.passes_in_flight)
    pcie_s = time.perf_counter() - t
    pcie_grays = (hst.extension_rays + hst.shadow_rays) / pcie_s / 1e9

    rays = rays_ext + rays_sh
    grays = rays / elapsed / 1e9
    default_workload = (W, H, args.spp, args.depth, args.seed) == (1920, 1080, 64, 12, 0)
    traffic = extend_traffic_bytes() if default_workload else None
    avg_ext_ms = kst.extend_ms / max(launches, 1)
    bytes_per_ray = ext_nodes * NODE_BYTES + ext_tris * TRI_BYTES + ext_io_bytes(kst.extension_rays, kst.paths)
    bytes_per_launch = bytes_per_ray * kst.extension_rays / max(launches, 1)
    achieved = bytes_per_launch / (avg_ext_ms * 1e-3) / 1e9 if avg_ext_ms > 0 else 0.0

    line = {
        "metric": "Grays/s + paths/s; time to relMSE 2e-3 (learned denoiser vs plain), 1 GPU",
        "value": grays,
        "unit": "Grays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded procedural scene)",
        "config": {
            "workload": f"procedural mesh scene, {sstats.triangles} triangles (displaced L7 icosphere + 64 "
                        f"instanced L3 icospheres + 640k-triangle ground), {W}x{H}, {args.spp} spp, depth "
                        f"{args.depth}, wavefront path tracer with NEE"
                        + ("" if args.materials == "c2" else ", all five materials (Phong and Plastic included)")
                        + ("" if args.lens_radius <= 0 else f", thin-lens camera (lens radius {args.lens_radius})"),
            "paths_per_s": paths / elapsed,
            "extension_rays": rays_ext,
            "shadow_rays": rays_sh,
            "scene_build_s": build_s,
            "scene_generate_s": gen_s,
            "device_scene_bytes": sstats.device_bytes,
            "tuning": tuning,
            "timed_path": "HIP-graph bounce loop (default render flags)",
            "device_ms_per_step": device_ms / args.steps,
            "instrumented_frame_ms": kst.milliseconds,
            "host_buffer_grays_per_s": pcie_grays,
            "host_buffer_ms": pcie_s * 1e3,
        },
        "roofline": {
            "kernel": "k_extend (closest-hit traversal of the compressed BVH8)",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_source": "profiles/k_extend_mesh_scene_hbm_traffic.md" if traffic else None,
            "avg_launch_ms": avg_ext_ms,
            "launches": launches,
            "bytes_per_ray": bytes_per_ray,
            "nodes_per_ray": ext_nodes,
            "tris_per_ray": ext_tris,
            "global_stack_pushes_per_ray": ext_spills,
            "shadow_kernel_ms_per_launch": kst.shadow_ms / max(launches, 1),
            "timing": "HIP events around each k_extend launch of one instrumented frame",
        },
    }
    if not args.no_denoise:
        line["config"]["time_to_relmse"] = time_to_relmse_leg(ctx)
    ctx.free(d_accum)
    scene.close()
    ctx.close()
    if not args.no_legs:
        # short runs of the other configurations, reported beside the headline
        cpu = None if args.no_cpu_baseline else 5.0
        legs = {}
        for name, run in (("soup", lambda: soup_run(args, 2, 1, cpu)),
                          ("batch", lambda: batch_run(args, 2, 1)),
                          ("unet", lambda: unet_run(args, 5, 2, cpu))):
            t = time.perf_counter()
            leg = run()
            legs[name] = {k: leg[k] for k in ("metric", "value", "unit", "ms_per_step", "steps", "config",
                                              "roofline", "cpu_baseline") if k in leg}
            legs[name]["wall_s"] = time.perf_counter() - t
        line["config"]["legs"] = legs
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(desc, args)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
