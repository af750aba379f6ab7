"""Benchmark of the rendering hot path on one MI355X.

Workload: the procedural mesh scene (~1.05 M triangles: displaced level-7
icosphere, 64 instanced level-3 icospheres, 640 k-triangle ground grid, one
spherical light) rendered at 1920x1080, 64 samples per pixel, max depth 12,
with the wavefront path tracer.  One step = one full 64-spp frame.

Prints one JSON line: Grays/s (extension + shadow rays traced per second) as
`value`, paths/s beside it, the roofline of the dominant kernel (closest-hit
traversal) from HIP-event timings of every launch in the timed region, and the
CPU oracle (the restatement of the reference's CPU path) timed on a bounded
sample of the same frame on the host's cores.
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
EXT_IO_BYTES = 4 + 24 + 20     # queue index, ray origin+direction, hit record


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
    p.add_argument("--no-cpu-baseline", action="store_true")
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


def cpu_baseline(desc, args):
    """Times the oracle (C++ restatement of PathTracer::Radiance with the
    reference traversal, OpenMP over rows as s3d_pt.cpp:311-313) on a band of
    rows at 1 spp, sized to about args.cpu_seconds of work."""
    from oracle import oracle as O
    threads = max(1, min(16, os.cpu_count() or 1))
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
    }


def main():
    args = parse()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the path shards by pixel tiles x "
              f"sample ranges with a replicated BVH, which this build does not launch.", file=sys.stderr)
        sys.exit(2)

    import salty_amd
    from salty_amd import abi, scenes

    t = time.perf_counter()
    desc = scenes.mesh_scene(seed=args.seed)
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
        return abi.RenderConfig(W, H, args.depth, spp, args.seed, 0, first_pass, flags, 0)

    # work counts (nodes visited / triangles tested per ray), outside the timed region
    cst = scene.render_device(cfg(2, abi.RENDER_COUNT_WORK, first_pass=1000), d_accum)
    ext_nodes = cst.extend_nodes / max(cst.extension_rays, 1)
    ext_tris = cst.extend_tris / max(cst.extension_rays, 1)

    for w in range(args.warmup):
        scene.render_device(cfg(args.spp, first_pass=0), d_accum)
    ctx.synchronize()

    ext_ms = sh_ms = 0.0
    launches = rays_ext = rays_sh = paths = 0
    t0 = time.perf_counter()
    for s in range(args.steps):
        st = scene.render_device(cfg(args.spp, abi.RENDER_TIME_KERNELS, first_pass=0), d_accum)
        ext_ms += st.extend_ms
        sh_ms += st.shadow_ms
        launches += st.extend_launches
        rays_ext += st.extension_rays
        rays_sh += st.shadow_rays
        paths += st.paths
    ctx.synchronize()
    elapsed = time.perf_counter() - t0

    rays = rays_ext + rays_sh
    grays = rays / elapsed / 1e9
    avg_ext_ms = ext_ms / max(launches, 1)
    bytes_per_ray = ext_nodes * NODE_BYTES + ext_tris * TRI_BYTES + EXT_IO_BYTES
    bytes_per_launch = bytes_per_ray * rays_ext / max(launches, 1)
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
                        f"{args.depth}, wavefront path tracer with NEE",
            "paths_per_s": paths / elapsed,
            "extension_rays": rays_ext,
            "shadow_rays": rays_sh,
            "scene_build_s": build_s,
            "scene_generate_s": gen_s,
            "device_scene_bytes": sstats.device_bytes,
        },
        "roofline": {
            "kernel": "k_extend (closest-hit traversal of the compressed BVH8)",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "avg_launch_ms": avg_ext_ms,
            "launches": launches,
            "bytes_per_ray": bytes_per_ray,
            "nodes_per_ray": ext_nodes,
            "tris_per_ray": ext_tris,
            "shadow_kernel_ms_per_launch": sh_ms / max(launches, 1),
        },
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(desc, args)
    print(json.dumps(line), flush=True)
    ctx.free(d_accum)


if __name__ == "__main__":
    main()
