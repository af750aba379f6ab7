"""Trains the U-Net denoiser on path-traced Cornell-style pairs and reports the
time-to-relMSE comparison on held-out scenes.

usage: python -m salty_amd.train_denoiser [--train-scenes N] [--steps S] [--out PATH]

Training scenes use seeds 1000.., held-out scenes seeds 0..(held-1) (seed 0 is
the canonical 36-triangle Cornell box).  Prints progress lines and one final
JSON line; writes the weights to salty_amd/weights/unet_cornell.npz.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--train-scenes", type=int, default=192)
    ap.add_argument("--held-out", type=int, default=8)
    ap.add_argument("--spp-in", type=int, default=4)
    ap.add_argument("--spp-ref", type=int, default=1024)
    ap.add_argument("--steps", type=int, default=1500)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--out", default=None)
    args = ap.parse_args()

    import salty_amd
    from salty_amd import denoise, scenes

    def log(msg):
        print(msg, flush=True)

    ctx = salty_amd.Context(0)
    t = time.perf_counter()
    feats, tgts = denoise.make_dataset(ctx, range(1000, 1000 + args.train_scenes), args.size, args.size, args.spp_in,
                                       args.spp_ref, log=log)
    data_s = time.perf_counter() - t
    log(f"dataset: {len(feats)} pairs in {data_s:.1f}s")
    net = salty_amd.Denoiser(ctx, seed=0)
    losses, train_s = denoise.train(net, ctx, feats, tgts, steps=args.steps, batch=args.batch, lr=args.lr, log=log)
    denoise.save_weights(net, args.out or denoise.WEIGHTS)
    log(f"trained {args.steps} steps in {train_s:.1f}s; final loss {np.mean(losses[-50:]):.4f}")
    results = []
    for s in range(args.held_out):
        r = denoise.time_to_relmse(ctx, net, scenes.cornell_variant(s), width=args.size, height=args.size,
                                   nlm=(s == 0))
        results.append(r)
        log(json.dumps({"held_out_seed": s, "plain_hit": r["plain_hit"], "denoised_hit": r["denoised_hit"],
                        "denoised": r["denoised"]}))
    print(json.dumps({"train_scenes": args.train_scenes, "steps": args.steps, "train_seconds": train_s,
                      "dataset_seconds": data_s, "final_loss": float(np.mean(losses[-50:])),
                      "held_out": results}), flush=True)


if __name__ == "__main__":
    sys.exit(main())
