"""Autotuner for the solver's performance knobs.

Coordinate descent over the options that change time but never labels:

* launch shapes: vertices per thread of the pull sweep and edges per thread
  of the push expansion (compiled instantiations), and the workgroups per
  CU of pull sweeps and list levels;
* the push/pull switch (Beamer alpha, beta);
* the frontier bound and workgroup count of the push levels run on the
  device, and whether they run as chained launches or one persistent launch;
* FW-BW rounds and the tail-solver threshold;
* the pipeline method.

Each candidate is timed on the GPU (median of ``reps`` solves, HIP events on
the solve's stream). Every candidate's labels must equal the first
candidate's, otherwise the tuner raises: a tuning run doubles as a parity
check of the option space on the caller's own graph.
"""
from __future__ import annotations

import json
import os
import statistics

import torch

import kaspan_amd as ka

TUNED_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuned")

# candidate values per knob, in the order they are tried
SPACE = {
    # KSPN_FLAG_FULL_BW (64): three-way split of the multi-pivot rounds; KSPN_FLAG_PERSISTENT (1024): device
    # push levels in one persistent launch instead of chained launches; KSPN_FLAG_NARROW (2048): chained
    # narrow levels for frontiers of degree <= 64; KSPN_FLAG_PRUNE (4096): edges found to join two partitions
    # are skipped from then on; KSPN_FLAG_CHUNKED (8192): frontiers of degree <= 64 without a degree scan.
    # First: they change what the frontier bounds below mean.
    "flags": [0, 64, 1024, 1088, 2048, 2112, 4096, 6144, 6208, 8192, 14336, 14400],
    "fwbw_rounds": [1, 4, 16, 24, 32, 48, 64],
    "alpha": [14.0, 4.0, 40.0, 1e-9],
    "beta": [24.0, 6.0, 96.0],
    "coop_edges": [1 << 20, 1 << 16, 1 << 23, 1],
    "coop_workgroups": [128, 64, 256],
    "tail_threshold": [4096, 1024, 16384],
    "method": [0, 1],
}
# launch shapes: compiled instantiations (pull sweep vertices per thread, push tile edges per thread) and
# grid sizes (workgroups per CU) the library selects at run time; 0 is the library's default
SHAPES = {
    "pull_ilp": [2, 1, 4],
    "expand_ept": [0, 8, 4, 2, 16],
    "sweep_wgs": [0, 4, 16],
    "list_wgs": [0, 2, 8],
    # launches per chain of device push levels (0: the library's default, 16, or 8 for narrow levels)
    "chain_levels": [0, 4, 8, 16],
}
SPACE.update(SHAPES)
DEFAULTS = {"pull_ilp": 2, "expand_ept": 0, "sweep_wgs": 0, "list_wgs": 0, "chain_levels": 0, "alpha": 14.0, "beta": 24.0,
            "coop_edges": 1 << 20, "coop_workgroups": 128, "fwbw_rounds": 1, "tail_threshold": 4096, "method": 0,
            "flags": 0, "learned_rounds": 0}
# the learned rule's schedule: rounds with the learned pivot before the plain rule takes over (0: all)
LEARNED_SPACE = {"learned_rounds": [0, 2, 4, 6, 8, 12, 16], "fwbw_rounds": [20, 24, 28, 32, 40, 48]}


def _solver(target):
    if isinstance(target, ka.Batch):
        return lambda lab, **o: ka.scc_batch(target, lab, stats=False, **o), target.graph.n
    return lambda lab, **o: ka.scc(target, lab, stats=False, **o), target.n


def time_opts(target, opts, reps=3, labels=None):
    """median milliseconds of ``reps`` solves with ``opts`` (after one warm-up) and the labels"""
    solve, n = _solver(target)
    lab = labels if labels is not None else torch.empty(n, dtype=torch.int32, device="cuda")
    solve(lab, **opts)
    stream = torch.cuda.current_stream()
    times = []
    for _ in range(reps):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        solve(lab, **opts)
        e1.record(stream)
        torch.cuda.synchronize()
        times.append(e0.elapsed_time(e1))
    return statistics.median(times), lab


def tune(target, space=None, start=None, reps=3, passes=1, log=None, fixed=None):
    """Coordinate descent from ``start`` (library defaults): returns (best opts, best ms, trials).

    ``target`` is a BidiGraph or a Batch. ``fixed`` options (the pivot rule and
    its parameters) go into every solve unchanged. ``trials`` lists (opts, ms)
    for every candidate timed.
    """
    space = space or SPACE
    fixed = dict(fixed or {})
    best = dict(DEFAULTS if start is None else start)
    best_ms, ref = time_opts(target, {**best, **fixed}, reps)
    ref = ref.clone()
    trials = [(dict(best), best_ms)]
    for _ in range(passes):
        for knob, values in space.items():
            for v in values:
                if best.get(knob) == v:
                    continue
                cand = dict(best)
                cand[knob] = v
                ms, lab = time_opts(target, {**cand, **fixed}, reps)
                if not torch.equal(lab, ref):
                    raise ka.KspnError(-1, f"autotune: labels differ with {cand}")
                trials.append((cand, ms))
                if log:
                    log(f"[autotune] {knob}={v}: {ms:.3f} ms (best {best_ms:.3f})")
                if ms < best_ms:
                    best, best_ms = cand, ms
    return best, best_ms, trials


def path_for(name):
    return os.path.join(TUNED_DIR, f"{name}.json")


def save(name, opts, ms=None, learned_opts=None, note=None):
    """store tuned options for a workload (kaspan_amd/tuned/<name>.json)"""
    os.makedirs(TUNED_DIR, exist_ok=True)
    rec = {"workload": name, "opts": {k: v for k, v in opts.items() if k in DEFAULTS or k in SHAPES}}
    if learned_opts:
        rec["learned_opts"] = learned_opts
    if ms is not None:
        rec["ms"] = ms
    if note:
        rec["note"] = note
    with open(path_for(name), "w") as f:
        json.dump(rec, f, indent=1, sort_keys=True)
    return path_for(name)


def load(name):
    """the stored record for a workload with its path, or None"""
    p = path_for(name)
    if not os.path.exists(p):
        return None
    rec = json.load(open(p))
    rec["path"] = os.path.relpath(p, os.path.dirname(TUNED_DIR))
    return rec


def main():
    """python -m kaspan_amd.autotune --config NAME: tune the plain rule (and, on the planted workload,
    the learned one) on the GPU and store the result in kaspan_amd/tuned/NAME.json, which bench.py
    reads. Each rule starts from its stored options (--fresh: from the library defaults)."""
    import argparse
    import time

    from . import pivot as P
    from .workloads import CONFIGS, Workload

    ap = argparse.ArgumentParser()
    ap.add_argument("--config", required=True, choices=sorted(CONFIGS))
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--passes", type=int, default=1)
    ap.add_argument("--learned", action="store_true", help="also tune the learned pivot rule")
    ap.add_argument("--fresh", action="store_true", help="start from the library defaults, not the stored options")
    args = ap.parse_args()
    stored = None if args.fresh else load(args.config)
    t0 = time.time()
    wl = Workload(ka, args.config)
    target = wl.batch if wl.batch is not None else wl.g
    log = lambda m: print(m, flush=True)  # noqa: E731
    space = dict(SPACE)
    if wl.batch is not None:
        space.pop("method")  # the batch runs the multi-pivot method
    start = dict(DEFAULTS, **stored["opts"]) if stored else None
    best, best_ms, trials = tune(target, space=space, start=start, reps=args.reps, passes=args.passes, log=log)
    learned, learned_ms = None, None
    if args.learned:
        mlp = P.load_params()
        if mlp is None:
            raise SystemExit("no trained scorer (python -m kaspan_amd.train_pivot)")
        lspace = {k: v for k, v in space.items() if k != "method"}
        lspace.update(LEARNED_SPACE)
        lstart = dict(DEFAULTS, **stored["learned_opts"]) if stored and stored.get("learned_opts") else dict(best)
        learned, learned_ms, ltrials = tune(target, space=lspace, start=lstart, reps=args.reps,
                                            passes=args.passes, log=log, fixed={"pivot": 1, "mlp_params": mlp})
        trials += ltrials
    path = save(args.config, best, ms=best_ms, learned_opts=learned,
                note=f"coordinate descent, {len(trials)} trials, median of {args.reps} solves each, "
                     f"{time.time() - t0:.0f}s" + (f"; learned rule {learned_ms:.3f} ms" if learned else ""))
    print(json.dumps({"config": args.config, "plain_ms": best_ms, "opts": best, "learned_ms": learned_ms,
                      "learned_opts": learned, "path": path}))


if __name__ == "__main__":
    main()
