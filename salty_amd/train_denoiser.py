"""Trains the U-Net denoiser on path-traced Cornell-style scenes and reports
the time-to-relMSE comparison on held-out scenes.

usage: python -m salty_amd.train_denoiser [--train-scenes N] [--spp-ref S] [--steps S] [--out PATH]

Training scenes use seeds 1000.., held-out scenes seeds 0..(held-1) (seed 0 is
the canonical 36-triangle Cornell box).  Inputs are the means after 1, 2, 4,
..., 256 samples per pixel of one progressive render; the reference is an
independent spp-ref render (2,000 scenes with 4,096-spp references by
default, rendered 64 scenes at a time as scene sets).  Prints progress lines
and one final JSON line; writes the weights to --out (default
salty_amd/weights/unet_cornell.npz).
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--train-scenes", type=int, default=2000)
    ap.add_argument("--held-out", type=int, default=4)
    ap.add_argument("--spp-ref", type=int, default=4096)
    ap.add_argument("--per-scene", action="store_true",
                    help="render the training pairs scene by scene instead of 64 scenes per batch")
    ap.add_argument("--steps", type=int, default=60000)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--crop", type=int, default=0, help="random crop side (0: whole frames)")
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--gt-spp", type=int, default=65536)
    ap.add_argument("--out", default=None)
    args = ap.parse_args()

    # the training loop indexes the device-resident dataset with torch; torch's
    # HIP runtime has to be initialised before the library's context
    import torch
    torch.cuda.init()
    import salty_amd
    from salty_amd import denoise, scenes

    def log(msg):
        print(msg, flush=True)

    ctx = salty_amd.Context(0)
    t = time.perf_counter()
    make = denoise.make_dataset if args.per_scene else denoise.make_dataset_batched
    data = make(ctx, range(1000, 1000 + args.train_scenes), args.size, args.size, spp_ref=args.spp_ref, log=log)
    data_s = time.perf_counter() - t
    log(f"dataset: {len(data)} scenes in {data_s:.1f}s")
    net = salty_amd.Denoiser(ctx, seed=0)
    losses, train_s = denoise.train(net, ctx, data, steps=args.steps, batch=args.batch, lr=args.lr,
                                    crop=args.crop or None, log=log)
    out = args.out or denoise.WEIGHTS
    denoise.save_weights(net, out)
    log(f"trained {args.steps} steps in {train_s:.1f}s; final loss {np.mean(losses[-100:]):.4f}; weights {out}")
    results = []
    for s in range(args.held_out):
        r = denoise.time_to_relmse(ctx, net, scenes.cornell_variant(s), width=args.size, height=args.size,
                                   gt_spp=args.gt_spp, nlm=(s == 0))
        results.append(r)
        log(json.dumps({"held_out_seed": s, "plain_ms": r["plain_ms"], "denoised_ms": r["denoised_ms"],
                        "best_denoised_relmse": r["best_denoised_relmse"],
                        "denoised": [(spp, round(ms, 2), round(inf, 2), round(rr, 5))
                                     for spp, ms, inf, rr in r["denoised"]],
                        "plain": [(spp, round(ms, 2), round(rr, 5)) for spp, ms, rr in r["plain"]]}))
    print(json.dumps({"train_scenes": args.train_scenes, "steps": args.steps, "train_seconds": train_s,
                      "dataset_seconds": data_s, "final_loss": float(np.mean(losses[-100:])),
                      "held_out": [{k: v for k, v in r.items() if k in ("plain_ms", "denoised_ms",
                                                                        "best_denoised_relmse", "gt_ms")}
                                   for r in results]}), flush=True)


if __name__ == "__main__":
    sys.exit(main())
