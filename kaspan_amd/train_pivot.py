"""Train the learned pivot scorer on the GPU and save its parameters.

    python -m kaspan_amd.train_pivot [--out kaspan_amd/pivot_mlp_params.npy] [--target log|size|shrinkage]

Training graphs are planted power-law SCC graphs (the generator of the
planted benchmark configuration, exponent 2.1) with seeds disjoint from the
benchmark's. The data are the pivot outcomes the multi-pivot FW-BW rounds
log: each round draws one pivot per partition at random among the
partition's candidates (the plain rule's top half-octave of in*out, the set
the learned rule chooses from), and the default target is log1p of the SCC
that pivot decided. --target size uses every vertex's SCC size from a full
solve instead; --target shrinkage the partition's shrinkage. Forward (BF16
MFMA), gradient (BF16 MFMA, FP32 accumulation) and Adam are libkspn kernels.
At the end the scorer is compared with the plain rule on a held-out graph:
vertices decided by the first rounds and the solve time.
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
    ap.add_argument("--rounds", type=int, default=48)
    ap.add_argument("--epochs", type=int, default=4000)
    ap.add_argument("--lr", type=float, default=1e-2)
    ap.add_argument("--target", choices=["log", "size", "shrinkage"], default="log")
    ap.add_argument("--gamma", type=float, default=0.0,
                    help="log target: minus gamma * log1p(forward set), favouring searches that decide what they traverse")
    args = ap.parse_args()
    t0 = time.time()
    xs, ts = [], []
    for i in range(args.graphs):
        g = generate("planted", 0x7452414E00000000 + i, n=args.n, max_size=args.max_size)
        for rep in range(1 if args.target == "size" else 4):
            x, t = P.outcome_samples(g, rounds=args.rounds, seed=17 * i + rep + 1, target=args.target, gamma=args.gamma)
            xs.append(x)
            ts.append(t)
        del g
    x, t = torch.cat(xs), torch.cat(ts)
    params, losses = P.train(x, t, epochs=args.epochs, lr=args.lr)
    # fit on a held-out graph: R^2 of the (BF16 MFMA) forward pass against the target
    gh = generate("planted", 0x7452414E00000000 + 1000, n=args.n, max_size=args.max_size)
    xh, th = P.outcome_samples(gh, seed=999, target=args.target, gamma=args.gamma)
    ph = P.forward(params, xh)
    r2 = 1.0 - float(((ph - th) ** 2).mean()) / float(th.var())
    np.save(args.out, params.cpu().numpy().astype(np.float32), allow_pickle=False)
    # held-out comparison with the plain rule: decided vertices at a fixed round budget, solve time
    from . import scc

    lab = torch.empty(gh.n, dtype=torch.int32, device="cuda")
    gain = {}
    for k in (4, 8, 16):
        _, sp = scc(gh, lab, fwbw_rounds=k)
        _, sl = scc(gh, lab, fwbw_rounds=k, pivot=1, mlp_params=params)
        gain[f"decided_{k}_rounds"] = {"plain": sp.decided_fwbw, "learned": sl.decided_fwbw}
    print(json.dumps({"target": args.target, "gamma": args.gamma, "heldout_r2": round(r2, 4), "samples": int(x.shape[0]),
                      "graphs": args.graphs, "n": args.n, "epochs": args.epochs,
                      "loss_first": losses[0], "loss_last": losses[-1], "seconds": round(time.time() - t0, 1),
                      "heldout": gain, "out": args.out}))


if __name__ == "__main__":
    main()
