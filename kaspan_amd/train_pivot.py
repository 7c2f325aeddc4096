"""Train the learned pivot scorer on the GPU and save its parameters.

    python -m kaspan_amd.train_pivot [--out kaspan_amd/pivot_mlp_params.npy] [--target size|shrinkage]

Training graphs are planted power-law SCC graphs (the generator of the
planted benchmark configuration, exponent 2.1) with seeds disjoint from the
benchmark's. The outcome of a pivot is what a FW-BW from it decides: one
solve of each graph gives every vertex's SCC size, and the default target is
log1p(SCC size) for a sample of live vertices. With --target shrinkage the
samples come from the multi-pivot rounds' own outcome log with random pivots
and the target is the partition's shrinkage. Forward, gradient and Adam are
libkspn kernels.
"""
import argparse
import json
import time

import numpy as np
import torch

from . import generate
from . import pivot as P


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=P.PARAMS_PATH)
    ap.add_argument("--n", type=int, default=2_000_000)
    ap.add_argument("--max-size", type=int, default=400_000)
    ap.add_argument("--graphs", type=int, default=6)
    ap.add_argument("--rounds", type=int, default=24)
    ap.add_argument("--epochs", type=int, default=4000)
    ap.add_argument("--lr", type=float, default=1e-2)
    ap.add_argument("--target", choices=["size", "shrinkage"], default="size")
    args = ap.parse_args()
    t0 = time.time()
    xs, ts = [], []
    for i in range(args.graphs):
        g = generate("planted", 0x7452414E00000000 + i, n=args.n, max_size=args.max_size)
        for rep in range(1 if args.target == "size" else 3):
            x, t = P.outcome_samples(g, rounds=args.rounds, seed=17 * i + rep + 1, target=args.target)
            xs.append(x)
            ts.append(t)
        del g
    x, t = torch.cat(xs), torch.cat(ts)
    params, losses = P.train(x, t, epochs=args.epochs, lr=args.lr)
    # fit on a held-out graph: R^2 of the (BF16 MFMA) forward pass against the target
    gh = generate("planted", 0x7452414E00000000 + 1000, n=args.n, max_size=args.max_size)
    xh, th = P.outcome_samples(gh, seed=999, target=args.target)
    ph = P.forward(params, xh)
    r2 = 1.0 - float(((ph - th) ** 2).mean()) / float(th.var())
    np.save(args.out, params.cpu().numpy().astype(np.float32), allow_pickle=False)
    print(json.dumps({"target": args.target, "heldout_r2": round(r2, 4), "samples": int(x.shape[0]),
                      "graphs": args.graphs, "n": args.n, "epochs": args.epochs,
                      "loss_first": losses[0], "loss_last": losses[-1], "seconds": round(time.time() - t0, 1),
                      "out": args.out}))


if __name__ == "__main__":
    main()
