"""Autotuner for the solver's performance knobs.

Coordinate descent over the options that change time but never labels:

* the push/pull switch (Beamer alpha, beta);
* the persistent multi-level kernel's frontier bound and workgroup count;
* FW-BW rounds and the tail-solver threshold;
* the pipeline method.

Each candidate is timed on the GPU (median of ``reps`` solves, HIP events on
the solve's stream). Every candidate's labels must equal the first
candidate's, otherwise the tuner raises: a tuning run doubles as a parity
check of the option space on the caller's own graph.
"""
from __future__ import annotations

import statistics

import torch

import kaspan_amd as ka

# candidate values per knob, in the order they are tried
SPACE = {
    "alpha": [14.0, 4.0, 40.0, 1e-9],
    "beta": [24.0, 6.0, 96.0],
    "coop_edges": [1 << 20, 1 << 16, 1 << 23, 1],
    "coop_workgroups": [128, 64, 256],
    "fwbw_rounds": [1, 4, 16, 64],
    "tail_threshold": [4096, 1024, 16384],
    "method": [0, 1],
}
DEFAULTS = {"alpha": 14.0, "beta": 24.0, "coop_edges": 1 << 20, "coop_workgroups": 128, "fwbw_rounds": 1,
            "tail_threshold": 4096, "method": 0}


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


def tune(target, space=None, start=None, reps=3, passes=1, log=None):
    """Coordinate descent from ``start`` (library defaults): returns (best opts, best ms, trials).

    ``target`` is a BidiGraph or a Batch. ``trials`` lists (opts, ms) for every
    candidate timed.
    """
    space = space or SPACE
    best = dict(DEFAULTS if start is None else start)
    best_ms, ref = time_opts(target, best, reps)
    ref = ref.clone()
    trials = [(dict(best), best_ms)]
    for _ in range(passes):
        for knob, values in space.items():
            for v in values:
                if best.get(knob) == v:
                    continue
                cand = dict(best)
                cand[knob] = v
                ms, lab = time_opts(target, cand, reps)
                if not torch.equal(lab, ref):
                    raise ka.KspnError(-1, f"autotune: labels differ with {cand}")
                trials.append((cand, ms))
                if log:
                    log(f"[autotune] {knob}={v}: {ms:.3f} ms (best {best_ms:.3f})")
                if ms < best_ms:
                    best, best_ms = cand, ms
    return best, best_ms, trials
