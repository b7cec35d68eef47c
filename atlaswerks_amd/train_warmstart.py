"""Train the atlas warm-start U-Net on the device and save its weights.

    python -m atlaswerks_amd.train_warmstart [--pairs 2000] [--steps 6000]
        [--size 128] [--out atlaswerks_amd/weights/unet_atlas_128.npy]

Training pairs are built as in ``warmstart.make_atlas_training_set``: groups of 16
subjects at size^3 (a common 8-blob template, own sinusoidal displacement up to 5
voxels, Rician noise 0.02), each group's plain greedy atlas run over scale levels
1/4, 1/2, 1 with 100 / 50 / 25 iterations; the network learns (subject, initial
mean) -> the subject's final displacement.  Forward, backward and Adam run in this
library's HIP kernels (``awk_unet_trainer_*``).  Prints one progress line per 10
groups and per epoch, then a JSON summary; the weights are a flat float32 blob in
``unet.LAYERS`` order (numpy .npy).
"""
import argparse
import json
import os
import sys
import time


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--pairs", type=int, default=2000)
    ap.add_argument("--steps", type=int, default=6000)
    ap.add_argument("--size", type=int, default=128)
    ap.add_argument("--lr", type=float, default=3e-4)
    ap.add_argument("--lr-final", type=float, default=3e-5)
    ap.add_argument("--seed", type=int, default=4)
    ap.add_argument("--out", default=os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                                  "weights", "unet_atlas_128.npy"))
    args = ap.parse_args(argv)

    import numpy as np
    import torch

    from . import warmstart
    from .unet import UNetWarmStart

    def log(msg):
        print(msg, flush=True)

    n = args.size
    t0 = time.time()
    triples, info = warmstart.make_atlas_training_set(args.pairs, shape=(n, n, n), seed=args.seed,
                                                      log=log)
    log("training set ready: %s" % json.dumps(info))
    net = UNetWarmStart((n, n, n), seed=args.seed)
    ts = time.time()
    tr, losses = warmstart.train(net, triples, args.steps, lr=args.lr, lr_final=args.lr_final,
                                 seed=args.seed, log=log)
    torch.cuda.synchronize()
    train_s = time.time() - ts
    w = tr.weights()
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    np.save(args.out, w.astype(np.float32))
    summary = {"pairs": info["pairs"], "groups": info["groups"], "size": [n, n, n],
               "steps": args.steps, "lr": [args.lr, args.lr_final],
               "generation_s": info["generation_s"], "train_s": round(train_s, 2),
               "ms_per_step": round(1000.0 * train_s / max(args.steps, 1), 3),
               "epoch_losses": [round(v, 6) for v in losses],
               "plain_final_energy_mean": info["plain_final_energy_mean"],
               "total_s": round(time.time() - t0, 2), "out": os.path.relpath(args.out)}
    print(json.dumps(summary), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
