"""Train the learned pivot scorer on the GPU and save its parameters.

    python -m kaspan_amd.train_pivot [--out kaspan_amd/pivot_mlp_params.npy]

Training graphs are planted power-law SCC graphs (the generator of the
planted benchmark configuration, exponent 2.1) with seeds disjoint from the
benchmark's. For each graph the method's own outcome log is taken: one solve
gives every vertex's SCC size, i.e. what a FW-BW from that vertex would
decide... the multi-pivot FW-BW log of seeded random pivots: per partition
and round the pivot, partition size, SCC size and the forward-only /
backward-only / unreached parts; the target is the residual shrinkage.
Forward, gradient and Adam are libkspn kernels.
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
    ap.add_argument("--epochs", type=int, default=1500)
    ap.add_argument("--lr", type=float, default=3e-3)
    args = ap.parse_args()
    t0 = time.time()
    xs, ts = [], []
    for i in range(args.graphs):
        g = generate("planted", 0x7452414E00000000 + i, n=args.n, max_size=args.max_size)
        for rep in range(3):
            x, t = P.outcome_samples(g, rounds=args.rounds, seed=17 * i + rep + 1)
            xs.append(x)
            ts.append(t)
        del g
    x, t = torch.cat(xs), torch.cat(ts)
    params, losses = P.train(x, t, epochs=args.epochs, lr=args.lr)
    np.save(args.out, params.cpu().numpy().astype(np.float32), allow_pickle=False)
    print(json.dumps({"samples": int(x.shape[0]), "graphs": args.graphs, "n": args.n, "epochs": args.epochs,
                      "loss_first": losses[0], "loss_last": losses[-1], "seconds": round(time.time() - t0, 1),
                      "out": args.out}))


if __name__ == "__main__":
    main()
