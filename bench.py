"""Benchmark: edges/sec to full SCC labels on one MI355X.

Workload (default): directed R-MAT scale 24, edge factor 16 (268,435,456
draws), a=0.57 b=0.19 c=0.19 d=0.05, ids randomly permuted, self-loops and
duplicates removed, uint32 ids / uint64 offsets, CSR + reverse CSR resident
in HBM. A step is one full solve (CSR in HBM -> canonical labels in HBM).
value = m_final / seconds per step.

The batch workload (--config batch4096) packs 4096 independent graphs (G(n,m)
or R-MAT with probability 1/2, n log-uniform in [1e3, 1e6], average
out-degree log-uniform in [2, 16]) into one union graph solved by one
kspn_scc_batch call per step; it also reports graphs/sec and the p50/p99
latency of solving sampled graphs one at a time.

Prints one JSON line. `roofline` is measured by timing every kernel launch
of the solve with HIP events recorded on the solve's stream (the library's
profile mode) over --steps further solves; `cpu_baseline` is the oracle's
single-thread restatement of kaspan::scc on the same graph, copied to the
host.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# library profile names -> kernel symbols as rocprofv3 reports them
ROCPROF_NAMES = {"pull_fw": ["kspn::k_pull<0>"], "pull_bw": ["kspn::k_pull<1>"], "trim_first": ["kspn::k_trim_first"],
                 "expand_fw": ["kspn::k_expand<0>"], "expand_bw": ["kspn::k_expand<1>", "kspn::k_expand<5>"],
                 "label_scc": ["kspn::k_label_bits", "kspn::k_label_prep", "kspn::k_part_apply"], "tail_solve": ["kspn::k_tail_solve"],
                 "trim_sweep": ["kspn::k_trim_sweep<1>"],
                 "reach_coop": ["kspn::k_reach_coop<0>", "kspn::k_reach_coop<1>", "kspn::k_reach_coop<5>"],
                 "bitmap_ops": ["kspn::k_part_classify"], "pivot_reduce": ["kspn::k_part_key"],
                 "pull_list": ["kspn::k_pull_list<0>", "kspn::k_pull_list<1>"]}


def pmc_traffic(config, kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE
    summary of this workload (copy-calibrated, launch-weighted over the kernel's
    instantiations), or None."""
    path = os.path.join(ROOT, "profiles", f"{config}_pmc_traffic.json")
    if not os.path.exists(path):
        return None
    data = json.load(open(path)).get("kernels", {})
    found = [data[k] for k in ROCPROF_NAMES.get(kernel, []) if k in data]
    launches = sum(k["launches"] for k in found)
    if not launches:
        return None
    return sum(k["bytes_per_launch"] * k["launches"] for k in found) / launches


CONFIGS = {
    # name: (generator kind, kwargs, seed, description)
    "gnm64k": ("gnm", dict(n=65536, draws=524288), 0x6B5350AA00000000,
               "G(n,m) n=65536 m=524288 draws, self-loops/duplicates removed"),
    "rmat24": ("rmat", dict(scale=24, draws=16 << 24, a=0.57, b=0.19, c=0.19, permute=True), 0x6B5350AA00000001,
               "R-MAT scale 24, edge factor 16, a=.57 b=.19 c=.19 d=.05, permuted ids, self-loops/duplicates removed"),
    "planted50m": ("planted", dict(n=50_000_000, max_size=10_000_000, expo=2.1, intra_deg=4.0, inter_deg=3.0),
                   0x6B5350AA00000002,
                   "planted SCCs, n=5e7, power-law sizes (exponent 2.1, 1..1e7), Hamiltonian cycle + chords to "
                   "out-degree 4 inside, 3 forward inter-component edges per vertex"),
    "lattice256": ("lattice", dict(side=256, p_bidir=0.1), 0x6B5350AA00000003,
                   "3-D lattice 256^3, each edge one random direction, 10% bidirectional"),
    "batch4096": ("batch", dict(count=4096, n_min=1000, n_max=1_000_000, deg_min=2.0, deg_max=16.0, p_rmat=0.5),
                  0x6B5350AA00000004,
                  "4096 independent graphs in one batch: G(n,m) or R-MAT(.57/.19/.19/.05) with p=.5, n log-uniform "
                  "in [1e3, 1e6], average out-degree log-uniform in [2, 16], deduplicated, concatenated CSR with "
                  "per-graph vertex/edge offsets"),
}
METHOD_NAMES = ["kaspan", "trim_ex", "hpc_like", "ispan_like", "hpc_trim_ex", "trim_ex_residual",
                "trim_ex_light_residual"]
LATENCY_SAMPLE = 128  # graphs of the batch solved one at a time for p50/p99
# FW-BW rounds when --fwbw-rounds is not given: the planted workload's many large SCCs want the
# partitioned multi-pivot rounds (32 is the plain rule's best in a sweep of 1..64)
DEFAULT_ROUNDS = {"planted50m": 32}
# bounded CPU-baseline samples for workloads whose full size takes the serial restatement too long
CPU_SAMPLE = {
    "planted50m": ("planted", dict(n=500_000, max_size=100_000, expo=2.1, intra_deg=4.0, inter_deg=3.0)),
}
CPU_MIN_S = 10.0  # the CPU baseline repeats a solve until this much time has passed (at most 1000 times)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", default="rmat24", choices=sorted(CONFIGS))
    ap.add_argument("--manifest", default=None, metavar="FILE",
                    help="solve the graph of a manifest file (convert_graph's format) instead of --config")
    ap.add_argument("--method", type=int, default=None, help="0 kaspan, 1 trim_ex (default: fastest known)")
    ap.add_argument("--variant", default=None,
                    help="one of the reference driver's variant names (sets --method): async, hpc_like, "
                         "hpc_trim_ex, ispan_like, trim_ex, trim_ex_light_residual, trim_ex_residual")
    ap.add_argument("--pivot", choices=["plain", "learned"], default="plain")
    ap.add_argument("--fwbw-rounds", type=int, default=None)
    ap.add_argument("--coop-edges", type=int, default=0, help="persistent-kernel frontier edge bound (0: library default)")
    ap.add_argument("--coop-workgroups", type=int, default=0, help="persistent-kernel workgroups (0: library default)")
    ap.add_argument("--alpha", type=float, default=14.0, help="push->pull switch (Beamer alpha)")
    ap.add_argument("--beta", type=float, default=24.0, help="pull->push switch (Beamer beta)")
    ap.add_argument("--autotune", action="store_true",
                    help="tune alpha/beta/persistent-kernel/round/tail/method options on this workload first")
    ap.add_argument("--ab", action="store_true", help="also time the other pivot rule; report both and the speedup")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-profile", action="store_true")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="only run a plain device copy of BYTES bytes 3 times (scales rocprofv3 FETCH/WRITE_SIZE)")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch

        src = torch.ones(args.calibrate_copy, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes": args.calibrate_copy}))
        return
    if args.fwbw_rounds is None:
        args.fwbw_rounds = DEFAULT_ROUNDS.get(args.config, 1)
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1 (multi-GPU sharding is a design, not built)"}))
        sys.exit(2)

    import numpy as np
    import torch

    import kaspan_amd as ka

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    if args.variant is not None:
        if args.variant not in ka.VARIANTS:
            raise SystemExit(f"unknown variant {args.variant!r}")
        args.method = ka.VARIANTS[args.variant]
    if args.manifest:
        from kaspan_amd import manifest as M

        mf = M.parse_manifest(args.manifest)
        args.config = f"manifest:{mf.name}"
        kind, kw, seed = "manifest", {}, None
        desc = (f"manifest graph {mf.name} (n={mf.node_count}, m={mf.edge_count}, {mf.head_bytes}/{mf.csr_bytes}-byte "
                f"{mf.endian}-endian arrays, self-loops: {mf.contains_self_loops}, duplicate edges: "
                f"{mf.contains_duplicate_edges})")
    else:
        kind, kw, seed, desc = CONFIGS[args.config]
    print(f"[bench] generating {args.config}", file=sys.stderr, flush=True)
    t0 = time.time()
    batch = None
    if kind == "manifest":
        g = M.load_manifest(mf)
    elif kind == "batch":
        bkw = dict(kw)
        count = bkw.pop("count")
        batch = ka.generate_batch(seed, count, **bkw)
        g = batch.graph
    else:
        g = ka.generate(kind, seed, **kw)
    torch.cuda.synchronize()
    gen_s = time.time() - t0
    print(f"[bench] n={g.n} m={g.m} generated in {gen_s:.1f}s", file=sys.stderr, flush=True)

    def solve(lab, **o):
        if batch is not None:
            return ka.scc_batch(batch, lab, **o)
        return ka.scc(g, lab, **o)
    method = args.method if args.method is not None else 0
    labels = torch.empty(g.n, dtype=torch.int32, device="cuda")
    stream = torch.cuda.current_stream()
    mlp = None
    if args.pivot == "learned" or args.ab:
        from kaspan_amd import pivot as P

        mlp = P.load_params()
        if mlp is None:
            raise SystemExit("no trained scorer: run python -m kaspan_amd.train_pivot first")
    opts = dict(method=method, fwbw_rounds=args.fwbw_rounds, coop_edges=args.coop_edges,
                coop_workgroups=args.coop_workgroups, alpha=args.alpha, beta=args.beta)
    if args.pivot == "learned":
        opts.update(pivot=1, mlp_params=mlp)

    tuned = None
    if args.autotune:
        from kaspan_amd import autotune

        space = dict(autotune.SPACE)
        if args.pivot == "learned":
            space.pop("method")
        t0 = time.time()
        best, best_ms, trials = autotune.tune(batch if batch is not None else g, space=space,
                                              start={**autotune.DEFAULTS, **{k: v for k, v in opts.items()
                                                                             if k in autotune.DEFAULTS}},
                                              log=lambda m: print(m, file=sys.stderr, flush=True))
        opts.update(best)
        tuned = {"opts": best, "ms": best_ms, "trials": len(trials), "seconds": time.time() - t0}

    def timed(o):
        for _ in range(args.warmup):
            solve(labels, **o)
        torch.cuda.synchronize()
        start = torch.cuda.Event(enable_timing=True)
        end = torch.cuda.Event(enable_timing=True)
        start.record(stream)
        for _ in range(args.steps):
            solve(labels, stats=False, **o)
        end.record(stream)
        torch.cuda.synchronize()
        ms = start.elapsed_time(end) / args.steps
        # statistics (per-scope HIP events) from one further, untimed solve
        _, st = solve(labels, **o)
        return ms, st

    ms_per_step, stats = timed(opts)
    value = g.m / (ms_per_step / 1e3)
    print(f"[bench] {ms_per_step:.3f} ms/step", file=sys.stderr, flush=True)
    ab = None
    if args.ab:
        import numpy as np

        first = labels.clone()
        other = {k: v for k, v in opts.items() if k not in ("pivot", "mlp_params")}
        if args.pivot == "plain":
            other.update(pivot=1, mlp_params=mlp)
        ms_other, st_other = timed(other)
        same = bool(torch.equal(first, labels))
        plain_ms, learned_ms = (ms_per_step, ms_other) if args.pivot == "plain" else (ms_other, ms_per_step)
        ab = {"plain_ms": plain_ms, "learned_ms": learned_ms, "speedup_learned": plain_ms / learned_ms,
              "labels_identical": same, "fwbw_rounds": args.fwbw_rounds,
              "fwbw_decided_plain": (stats if args.pivot == "plain" else st_other).decided_fwbw,
              "fwbw_decided_learned": (st_other if args.pivot == "plain" else stats).decided_fwbw}
        labels.copy_(first)

    # per-kernel profile over further steps (HIP events around every launch)
    roofline = None
    kernels = []
    if not args.no_profile:
        agg = {}
        for _ in range(args.steps):
            _, pst = solve(labels, profile=True, **opts)
            for k in pst.as_dict()["kernels"]:
                a = agg.setdefault(k["name"], {"launches": 0, "ms": 0.0, "bytes": 0.0})
                a["launches"] += k["launches"]
                a["ms"] += k["ms"]
                a["bytes"] += k["bytes"]
        for name, a in sorted(agg.items(), key=lambda kv: -kv[1]["ms"]):
            kernels.append({"name": name, "launches_per_step": a["launches"] / args.steps,
                            "ms_per_step": a["ms"] / args.steps,
                            "avg_launch_us": 1e3 * a["ms"] / max(a["launches"], 1),
                            "gbs": a["bytes"] / max(a["ms"], 1e-12) / 1e6})
        dom_name, dom = max(agg.items(), key=lambda kv: kv[1]["ms"])
        per_launch_bytes = dom["bytes"] / dom["launches"]
        per_launch_s = dom["ms"] / dom["launches"] / 1e3
        achieved = per_launch_bytes / per_launch_s / 1e9
        roofline = {"bound": "hbm", "kernel": dom_name, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                    "traffic": pmc_traffic(args.config, dom_name),
                    "bytes_per_launch": per_launch_bytes, "avg_launch_us": per_launch_s * 1e6}

    batch_info = None
    if batch is not None:
        # serving one graph at a time: latency of each sampled graph alone (same options)
        idx = sorted(set(np.linspace(0, batch.count - 1, LATENCY_SAMPLE).astype(int).tolist()))
        lat = []
        for i in idx:
            sg = batch.subgraph(i)
            slab = torch.empty(sg.n, dtype=torch.int32, device="cuda")
            ka.scc(sg, slab, **opts)  # warm-up of this size
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            ka.scc(sg, slab, **opts)
            e1.record(stream)
            torch.cuda.synchronize()
            lat.append(e0.elapsed_time(e1))
        lat = np.array(lat)
        batch_info = {"graphs": batch.count, "graphs_per_s": batch.count / (ms_per_step / 1e3),
                      "batch_latency_ms": ms_per_step,
                      "solo_latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                                          "mean": float(lat.mean()), "sample_graphs": len(idx)},
                      "solo_graphs_per_s_est": 1e3 / float(lat.mean()), "generate_s": gen_s}

    cpu = None
    if not args.no_cpu_baseline and batch is not None:
        from oracle import oracle as O  # checker / CPU baseline only

        # graphs of the batch in a fixed order until about 15 s of CPU time
        order = np.random.default_rng(0).permutation(batch.count)
        cpu_s, cpu_m, done, same = 0.0, 0, 0, True
        lab_h = labels.cpu().numpy().view(np.uint32)
        vo = batch.vtx_off.cpu().numpy().view(np.uint32)
        for i in order:
            fo, fa, bo, ba = batch.subgraph(int(i)).to_host()
            hg = O.HostGraph(batch.n_of(int(i)), fo, fa, bo, ba)
            t = time.perf_counter()
            ref, _ = O.kaspan_scc(hg)
            cpu_s += time.perf_counter() - t
            cpu_m += hg.m
            done += 1
            same &= bool(np.array_equal(O.canonicalize(ref), lab_h[vo[i]: vo[i + 1]]))
            if cpu_s > 15.0:
                break
        cpu = {"value": cpu_m / cpu_s, "unit": "edges/s", "cores": 1, "kind": "port",
               "sample": f"{done} random graphs of the batch ({cpu_m} edges) solved one by one; single-thread C "
                         f"restatement of kaspan::scc (scc/scc.hpp:21-87), lean form; {cpu_s:.2f}s; canonical "
                         f"labels equal to the GPU batch's: {same}"}
    elif not args.no_cpu_baseline:
        from oracle import oracle as O  # checker / CPU baseline only

        if args.config in CPU_SAMPLE:
            skind, skw = CPU_SAMPLE[args.config]
            sg = ka.generate(skind, seed, **skw)
            slab, _ = ka.scc(sg)
            what = f"{skind} sample {skw} (n={sg.n}, m={sg.m}) of the {args.config} workload"
        else:
            sg, slab = g, labels
            what = f"whole {args.config} graph (n={g.n}, m={g.m})"
        fo, fa, bo, ba = sg.to_host()
        hg = O.HostGraph(sg.n, fo, fa, bo, ba)
        reps, cpu_s = 0, 0.0
        while (cpu_s < CPU_MIN_S and reps < 1000) or reps == 0:
            t = time.perf_counter()
            ref, _ = O.kaspan_scc(hg)
            cpu_s += time.perf_counter() - t
            reps += 1
        same = bool(np.array_equal(O.canonicalize(ref), slab.cpu().numpy().view(np.uint32)))
        cpu = {"value": sg.m * reps / cpu_s, "unit": "edges/s", "cores": 1, "kind": "port",
               "sample": f"{what}, solved {reps} times; single-thread C restatement of kaspan::scc "
                         f"(scc/scc.hpp:21-87), lean form; {cpu_s:.2f}s in total; canonical labels equal to "
                         f"the GPU's: {same}"}

    out = {
        "metric": "edges/sec to full SCC labels on 1 MI355X",
        "value": value,
        "unit": "edges/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "uint32",
        "data": "synthetic",
        "config": {"workload": desc, "name": args.config, "n": g.n, "m": g.m,
                   "method": METHOD_NAMES[opts.get("method", method)], "pivot": args.pivot,
                   "fwbw_rounds": opts.get("fwbw_rounds", args.fwbw_rounds), "seed": seed,
                   "options": {k: v for k, v in opts.items() if k != "mlp_params"}},
        "ab": ab,
        "autotune": tuned,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "phases_ms": {k: getattr(stats, k) for k in ("ms_trim", "ms_fwbw", "ms_color", "ms_residual", "ms_canon")},
        "decided": {k: getattr(stats, k) for k in ("decided_trim", "decided_trim2", "decided_fwbw", "decided_color",
                                                   "decided_residual")},
        "solve": {k: getattr(stats, k) for k in ("pivot", "fwbw_levels_push", "fwbw_levels_pull", "color_rounds",
                                                 "color_iters", "edges_scanned", "residual_n")},
        "kernels": kernels[:8],
        "scopes": stats.phases(),
    }
    if batch_info is not None:
        out["batch"] = batch_info
    print(json.dumps(out))


if __name__ == "__main__":
    main()
