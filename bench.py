"""Benchmark: edges/sec to full SCC labels on one MI355X, learned vs plain pivots.

Default workload (the metric's own comparison): planted SCCs with n = 5e7,
component sizes from a power law (exponent 2.1, 1..1e7), each component a
Hamiltonian cycle plus chords to out-degree 4, three forward inter-component
edges per vertex; CSR + reverse CSR (uint32 ids, uint64 offsets) resident in
HBM. A step is one full solve (CSR in HBM -> canonical labels in HBM). Both
pivot rules are timed on the same graph, their labels must be identical, and
value = m / seconds per step with the learned pivot; `ab` holds both times.

Other workloads (--config): gnm64k (G(n,m), n=65536), rmat24 (R-MAT scale
24), lattice256 (3-D lattice 256^3), batch4096 (4096 graphs in one batch;
also graphs/sec and the p50/p99 latency of sampled graphs solved alone).
The default run also measures R-MAT scale 24 (plain pivot) under "extra".

Prints one JSON line. `roofline` times every kernel launch of further solves
with HIP events on the solve's stream (the library's profile mode); its
`traffic` is the HBM bytes per launch of that kernel from the committed
rocprofv3 FETCH_SIZE/WRITE_SIZE summary of the workload. `phases` gives each
statistics scope's algorithmic bytes (profile mode) over its time (HIP
events per scope, no profiling). `cpu_baseline` is the oracle's
single-thread restatement of kaspan::scc (lean form, and faithful form with
the ghost-cache work of scc/scc.hpp:53,77,82) on a bounded sample.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from kaspan_amd.workloads import CONFIGS, Workload  # noqa: E402  (no GPU work at import)

METRIC = "edges/sec to full SCC labels on 1 MI355X; learned vs plain pivots, same labels"
HBM_PEAK_GBS = 8000.0
# library profile names -> prefixes of the kernel symbols rocprofv3 reports for them (every instantiation)
ROCPROF_NAMES = {"pull_fw": ["kspn::k_pull<0,"], "pull_bw": ["kspn::k_pull<1,"], "trim_first": ["kspn::k_trim_first"],
                 "expand_fw": ["kspn::k_expand<0,"], "expand_bw": ["kspn::k_expand<1,", "kspn::k_expand<5,"],
                 "expand_chunk_fw": ["kspn::k_expand_chunk<0>"],
                 "expand_chunk_bw": ["kspn::k_expand_chunk<1>", "kspn::k_expand_chunk<5>"],
                 "expand_color": ["kspn::k_expand<3,", "kspn::k_expand_chunk<3>"],
                 "expand_color_bw": ["kspn::k_expand<4,", "kspn::k_expand_chunk<4>"],
                 "label_scc": ["kspn::k_label_bits", "kspn::k_label_prep", "kspn::k_part_apply", "kspn::k_part_label"],
                 "tail_solve": ["kspn::k_tail_solve"], "trim_sweep": ["kspn::k_trim_sweep<"],
                 "reach_coop": ["kspn::k_reach_coop<", "kspn::k_chain_level<", "kspn::k_narrow_level<"],
                 "bitmap_ops": ["kspn::k_part_split", "kspn::k_part_classify"],
                 "pivot_reduce": ["kspn::k_part_key", "kspn::k_part_prio"],
                 "frontier_deg": ["kspn::k_deg_scan", "kspn::k_frontier_deg"], "color_pull": ["kspn::k_color_pull"],
                 "pull_list": ["kspn::k_pull_list<"]}

METHOD_NAMES = ["kaspan", "trim_ex", "hpc_like", "ispan_like", "hpc_trim_ex", "trim_ex_residual",
                "trim_ex_light_residual"]
LATENCY_SAMPLE = 128  # graphs of the batch solved one at a time for p50/p99
# FW-BW rounds when neither --fwbw-rounds nor a tuned file gives them: the planted workload's many
# large SCCs want the partitioned multi-pivot rounds
DEFAULT_ROUNDS = {"planted50m": 48}
# bounded CPU-baseline samples for workloads whose full size takes the serial restatement too long
CPU_SAMPLE = {
    "planted50m": ("planted", dict(n=500_000, max_size=100_000, expo=2.1, intra_deg=4.0, inter_deg=3.0)),
}
CPU_MIN_S = 6.0  # each CPU form repeats its solve until this much time has passed (at most 1000 times)


def pmc_traffic(config, kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE
    summary of this workload (copy-calibrated, launch-weighted over the kernel's
    instantiations), or None."""
    path = os.path.join(ROOT, "profiles", f"{config}_pmc_traffic.json")
    if not os.path.exists(path):
        return None
    data = json.load(open(path)).get("kernels", {})
    prefixes = ROCPROF_NAMES.get(kernel, [])
    found = [v for k, v in data.items() if any(k.startswith(p) for p in prefixes)]
    launches = sum(k["launches"] for k in found)
    if not launches:
        return None
    return sum(k["bytes_per_launch"] * k["launches"] for k in found) / launches


def host_cpu():
    """cores and model of the host the benchmark runs on"""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count()
    return {"nproc": avail, "cpu_count": os.cpu_count(), "model": model}


def timed(wl, labels, opts, steps, warmup):
    """ms per solve over `steps` solves after `warmup` (HIP events on the stream), and
    the statistics of one further solve"""
    import torch

    stream = torch.cuda.current_stream()
    for _ in range(warmup):
        wl.solve(labels, **opts)
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for _ in range(steps):
        wl.solve(labels, stats=False, **opts)
    end.record(stream)
    torch.cuda.synchronize()
    ms = start.elapsed_time(end) / steps
    _, st = wl.solve(labels, **opts)
    return ms, st


def profile(wl, labels, opts, steps):
    """per-kernel and per-scope profile over `steps` solves in profile mode"""
    agg, scopes = {}, {}
    for _ in range(steps):
        _, pst = wl.solve(labels, profile=True, **opts)
        d = pst.as_dict()
        for k in d["kernels"]:
            a = agg.setdefault(k["name"], {"launches": 0, "ms": 0.0, "bytes": 0.0})
            a["launches"] += k["launches"]
            a["ms"] += k["ms"]
            a["bytes"] += k["bytes"]
        for ph in d["phases"]:
            s = scopes.setdefault(ph["name"], {"bytes": 0.0, "kernel_ms": 0.0})
            s["bytes"] += ph["bytes"]
            s["kernel_ms"] += ph["kernel_ms"]
    kernels = [{"name": name, "launches_per_step": a["launches"] / steps, "ms_per_step": a["ms"] / steps,
                "avg_launch_us": 1e3 * a["ms"] / max(a["launches"], 1),
                "gbs": a["bytes"] / max(a["ms"], 1e-12) / 1e6}
               for name, a in sorted(agg.items(), key=lambda kv: -kv[1]["ms"])]
    dom_name, dom = max(agg.items(), key=lambda kv: kv[1]["ms"])
    per_launch_bytes = dom["bytes"] / dom["launches"]
    per_launch_s = dom["ms"] / dom["launches"] / 1e3
    achieved = per_launch_bytes / per_launch_s / 1e9
    roofline = {"bound": "hbm", "kernel": dom_name, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": pmc_traffic(wl.name, dom_name), "bytes_per_launch": per_launch_bytes,
                "avg_launch_us": per_launch_s * 1e6, "launches_per_step": dom["launches"] / steps}
    per_scope = {k: {"bytes": v["bytes"] / steps, "kernel_ms": v["kernel_ms"] / steps} for k, v in scopes.items()}
    return kernels, roofline, per_scope


def phase_table(stats, per_scope):
    """each statistics scope: time (no profiling), algorithmic bytes (profile mode), GB/s and its
    fraction of the HBM peak, and the same over the scope's summed kernel time"""
    out = []
    for ph in stats.phases():
        b = per_scope.get(ph["name"], {}).get("bytes") if per_scope else None
        kms = per_scope.get(ph["name"], {}).get("kernel_ms") if per_scope else None
        row = {"scope": ph["name"], "ms": ph["ms"], "decided": ph["decided"]}
        if b is not None:
            gbs = b / max(ph["ms"], 1e-12) / 1e6
            row.update(bytes=b, gbs=round(gbs, 2), frac=round(gbs / HBM_PEAK_GBS, 4), kernel_ms=kms,
                       gbs_kernels=round(b / max(kms, 1e-12) / 1e6, 2))
        out.append(row)
    return out


def tuned_options(name):
    """the autotuner's stored options for this workload (kaspan_amd/tuned/<name>.json), or None"""
    from kaspan_amd import autotune

    return autotune.load(name)


def cpu_baseline(wl, labels_gpu, what_pivot):
    """the oracle's single-thread restatement of kaspan::scc, lean and faithful forms, on a bounded
    sample of the workload; canonical labels checked against the GPU's"""
    import numpy as np

    from oracle import oracle as O  # checker / CPU baseline only

    ka = wl.ka
    forms = {}
    if wl.batch is not None:
        order = np.random.default_rng(0).permutation(wl.batch.count)
        lab_h = labels_gpu.cpu().numpy().view(np.uint32)
        vo = wl.batch.vtx_off.cpu().numpy().view(np.uint32)
        graphs, m_total = [], 0
        for i in order:
            fo, fa, bo, ba = wl.batch.subgraph(int(i)).to_host()
            graphs.append((int(i), O.HostGraph(wl.batch.n_of(int(i)), fo, fa, bo, ba)))
            m_total += graphs[-1][1].m
            if m_total > 30_000_000:
                break
        for form in ("lean", "faithful"):
            cpu_s, m_done, same, done = 0.0, 0, True, 0
            for i, hg in graphs:
                t = time.perf_counter()
                ref, _ = O.kaspan_scc(hg, faithful=form == "faithful")
                cpu_s += time.perf_counter() - t
                m_done += hg.m
                done += 1
                same &= bool(np.array_equal(O.canonicalize(ref), lab_h[vo[i]: vo[i + 1]]))
                if cpu_s > CPU_MIN_S:
                    break
            forms[form] = {"value": m_done / cpu_s, "seconds": cpu_s, "graphs": done, "edges": m_done,
                           "labels_equal_gpu": same}
        sample = (f"random graphs of the batch solved one by one ({forms['lean']['graphs']} lean, "
                  f"{forms['faithful']['graphs']} faithful)")
    else:
        if wl.name in CPU_SAMPLE:
            skind, skw = CPU_SAMPLE[wl.name]
            sg = ka.generate(skind, wl.seed, **skw)
            slab, _ = ka.scc(sg)
            sample = f"{skind} sample {skw} (n={sg.n}, m={sg.m}) of the {wl.name} workload, same seed"
        else:
            sg, slab = wl.g, labels_gpu
            sample = f"whole {wl.name} graph (n={sg.n}, m={sg.m})"
        fo, fa, bo, ba = sg.to_host()
        hg = O.HostGraph(sg.n, fo, fa, bo, ba)
        want = slab.cpu().numpy().view(np.uint32)
        for form in ("lean", "faithful"):
            reps, cpu_s = 0, 0.0
            while (cpu_s < CPU_MIN_S and reps < 1000) or reps == 0:
                t = time.perf_counter()
                ref, _ = O.kaspan_scc(hg, faithful=form == "faithful")
                cpu_s += time.perf_counter() - t
                reps += 1
            forms[form] = {"value": sg.m * reps / cpu_s, "seconds": cpu_s, "solves": reps,
                           "labels_equal_gpu": bool(np.array_equal(O.canonicalize(ref), want))}
    cpu = host_cpu()
    return {"value": forms["lean"]["value"], "unit": "edges/s", "cores": 1, "kind": "port",
            "sample": (f"{sample}; single-thread C restatement of kaspan::scc (scc/scc.hpp:21-87), plain pivot "
                       f"(the reference has no learned pivot); value = lean form (no ghost caches); faithful form "
                       f"(ghost-cache hash set/map allocated and cleared as scc.hpp:53,77,82 do) in forms; host "
                       f"{cpu['model']}, {cpu['nproc']} cores available, 1 used (KaSpan has no intra-rank "
                       f"threading)"),
            "forms": forms, "host": cpu}


def latency_sample(wl, opts):
    """p50/p99 of sampled graphs of the batch solved one at a time"""
    import numpy as np
    import torch

    ka, batch = wl.ka, wl.batch
    stream = torch.cuda.current_stream()
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
    return {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)), "mean": float(lat.mean()),
            "sample_graphs": len(idx)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", default="planted50m", choices=sorted(CONFIGS))
    ap.add_argument("--manifest", default=None, metavar="FILE",
                    help="solve the graph of a manifest file (convert_graph's format) instead of --config")
    ap.add_argument("--method", type=int, default=None, help="0 kaspan, 1 trim_ex (default: tuned or 0)")
    ap.add_argument("--variant", default=None,
                    help="one of the reference driver's variant names (sets --method): async, hpc_like, "
                         "hpc_trim_ex, ispan_like, trim_ex, trim_ex_light_residual, trim_ex_residual")
    ap.add_argument("--pivot", choices=["plain", "learned"], default=None,
                    help="pivot rule of the measured solve (default: learned on planted50m with --ab, else plain)")
    ap.add_argument("--fwbw-rounds", type=int, default=None)
    ap.add_argument("--coop-edges", type=int, default=None, help="persistent-kernel frontier edge bound")
    ap.add_argument("--coop-workgroups", type=int, default=None, help="persistent-kernel workgroups")
    ap.add_argument("--alpha", type=float, default=None, help="push->pull switch (Beamer alpha)")
    ap.add_argument("--beta", type=float, default=None, help="pull->push switch (Beamer beta)")
    ap.add_argument("--flags", type=int, default=None, help="KSPN_FLAG_* bits for the solves")
    ap.add_argument("--graph", action="store_true", help="the device-driven solve, captured into a HIP graph")
    ap.add_argument("--no-tuned", action="store_true", help="ignore kaspan_amd/tuned/<config>.json")
    ap.add_argument("--autotune", action="store_true", help="tune the options on this workload first")
    ap.add_argument("--ab", action="store_true", help="time both pivot rules (on by default for planted50m)")
    ap.add_argument("--no-ab", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-profile", action="store_true")
    ap.add_argument("--no-latency", action="store_true", help="batch: skip the one-graph-at-a-time latency sample")
    ap.add_argument("--no-extra", action="store_true", help="skip the R-MAT 24 line of the default run")
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
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1 (multi-GPU sharding is a "
                                   "design, not built)"}))
        sys.exit(2)
    ab = (args.ab or (args.config == "planted50m" and not args.manifest)) and not args.no_ab
    pivot = args.pivot or ("learned" if ab else "plain")

    import numpy as np
    import torch

    import kaspan_amd as ka
    from kaspan_amd import autotune

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    print(f"[bench] generating {args.manifest or args.config}", file=sys.stderr, flush=True)
    t0 = time.time()
    wl = Workload(ka, args.config, args.manifest)
    torch.cuda.synchronize()
    gen_s = time.time() - t0
    g = wl.g
    desc, kind = wl.desc, wl.kind
    print(f"[bench] n={g.n} m={g.m} generated in {gen_s:.1f}s", file=sys.stderr, flush=True)

    # options: library defaults < tuned file < command line
    opts = dict(autotune.DEFAULTS)
    opts["fwbw_rounds"] = DEFAULT_ROUNDS.get(args.config, 1)
    tuned = None if (args.no_tuned or args.manifest) else tuned_options(wl.name)
    if tuned:
        opts.update(tuned["opts"])
    for k, v in (("method", args.method), ("fwbw_rounds", args.fwbw_rounds), ("coop_edges", args.coop_edges),
                 ("coop_workgroups", args.coop_workgroups), ("alpha", args.alpha), ("beta", args.beta),
                 ("flags", args.flags)):
        if v is not None:
            opts[k] = v
    if args.variant is not None:
        if args.variant not in ka.VARIANTS:
            raise SystemExit(f"unknown variant {args.variant!r}")
        opts["method"] = ka.VARIANTS[args.variant]
    mlp = None
    if pivot == "learned" or ab:
        from kaspan_amd import pivot as P

        mlp = P.load_params()
        if mlp is None:
            raise SystemExit("no trained scorer: run python -m kaspan_amd.train_pivot first")
    plain_opts = {k: v for k, v in opts.items() if k not in ("pivot", "mlp_params")}
    learned_opts = dict(plain_opts, pivot=1, mlp_params=mlp)
    if tuned and tuned.get("learned_opts") and mlp is not None:
        learned_opts.update(tuned["learned_opts"])
    main_opts = learned_opts if pivot == "learned" else plain_opts
    if args.graph:
        main_opts = dict(main_opts, flags=main_opts.get("flags", 0) | ka._lib.FLAG_GRAPH)

    tune_info = {"source": tuned["path"], "opts": tuned["opts"]} if tuned else None
    if args.autotune:
        t0 = time.time()
        space = dict(autotune.SPACE)
        if pivot == "learned":
            space.pop("method")
        best, best_ms, trials = autotune.tune(wl.batch if wl.batch is not None else g, space=space,
                                              start={k: v for k, v in main_opts.items() if k in autotune.DEFAULTS},
                                              fixed={k: v for k, v in main_opts.items() if k not in autotune.DEFAULTS},
                                              log=lambda m: print(m, file=sys.stderr, flush=True))
        main_opts.update(best)
        tune_info = {"opts": best, "ms": best_ms, "trials": len(trials), "seconds": time.time() - t0}

    labels = torch.empty(g.n, dtype=torch.int32, device="cuda")
    ms_per_step, stats = timed(wl, labels, main_opts, args.steps, args.warmup)
    print(f"[bench] {pivot}: {ms_per_step:.3f} ms/step", file=sys.stderr, flush=True)
    value = g.m / (ms_per_step / 1e3)
    ab_out = None
    if ab:
        first = labels.clone()
        other_opts = plain_opts if pivot == "learned" else learned_opts
        ms_other, st_other = timed(wl, labels, other_opts, args.steps, args.warmup)
        same = bool(torch.equal(first, labels))
        print(f"[bench] other rule: {ms_other:.3f} ms/step, labels identical: {same}", file=sys.stderr, flush=True)
        plain_ms, learned_ms = (ms_other, ms_per_step) if pivot == "learned" else (ms_per_step, ms_other)
        st_plain, st_learned = (st_other, stats) if pivot == "learned" else (stats, st_other)
        ab_out = {"plain_ms": plain_ms, "learned_ms": learned_ms, "speedup_learned": plain_ms / learned_ms,
                  "labels_identical": same, "fwbw_rounds_plain": plain_opts.get("fwbw_rounds"),
                  "fwbw_rounds_learned": learned_opts.get("fwbw_rounds"),
                  "fwbw_decided_plain": st_plain.decided_fwbw, "fwbw_decided_learned": st_learned.decided_fwbw,
                  "plain_edges_per_s": g.m / (plain_ms / 1e3), "learned_edges_per_s": g.m / (learned_ms / 1e3)}
        # the two rules at the same options (each rule's tuned options in turn): the pivot rule is then
        # the only difference
        knobs_l = {k: v for k, v in learned_opts.items() if k not in ("pivot", "mlp_params")}
        if knobs_l != plain_opts:
            same = {}
            for name, knobs in (("plain_options", plain_opts), ("learned_options", knobs_l)):
                pl = plain_ms if knobs is plain_opts else None
                ll = learned_ms if knobs is knobs_l else None
                if pl is None:
                    pl, _ = timed(wl, labels, knobs, args.steps, args.warmup)
                if ll is None:
                    ll, _ = timed(wl, labels, dict(knobs, pivot=1, mlp_params=mlp), args.steps, args.warmup)
                same[name] = {"fwbw_rounds": knobs.get("fwbw_rounds"), "plain_ms": pl, "learned_ms": ll,
                              "speedup_learned": pl / ll, "labels_identical": bool(torch.equal(first, labels))}
            ab_out["same_options"] = same
        if kind == "planted":
            planted = ka.planted_labels(wl.seed, wl.kw["n"], max_size=wl.kw["max_size"], expo=wl.kw["expo"])
            ab_out["labels_equal_planted"] = bool(torch.equal(first, planted))
        labels.copy_(first)

    roofline, kernels, phases = None, [], phase_table(stats, None)
    if not args.no_profile:
        psteps = max(1, min(args.steps, 5 if g.m > 1e8 else args.steps))
        kernels, roofline, per_scope = profile(wl, labels, main_opts, psteps)
        phases = phase_table(stats, per_scope)

    batch_info = None
    if wl.batch is not None:
        batch_info = {"graphs": wl.batch.count, "graphs_per_s": wl.batch.count / (ms_per_step / 1e3),
                      "batch_latency_ms": ms_per_step, "generate_s": gen_s}
        if not args.no_latency:
            lat = latency_sample(wl, main_opts)
            batch_info["solo_latency_ms"] = lat
            batch_info["solo_graphs_per_s_est"] = 1e3 / lat["mean"]
        if pivot == "plain" and not (main_opts.get("flags", 0) & ka._lib.FLAG_GRAPH):
            # the device-driven solve, captured once into a HIP graph and replayed
            gopts = dict(main_opts, flags=main_opts.get("flags", 0) | ka._lib.FLAG_GRAPH)
            glab = torch.empty_like(labels)
            gms, gst = timed(wl, glab, gopts, args.steps, args.warmup)
            batch_info["captured"] = {"ms": gms, "graphs_per_s": wl.batch.count / (gms / 1e3),
                                      "edges_per_s": g.m / (gms / 1e3), "labels_identical": bool(torch.equal(glab, labels)),
                                      "colour_rounds": gst.color_rounds, "decided_fwbw": gst.decided_fwbw,
                                      "levels": gst.fwbw_levels_push + gst.fwbw_levels_pull}
            if not args.no_latency:
                batch_info["captured"]["solo_latency_ms"] = latency_sample(wl, gopts)
            del glab

    cpu = None if args.no_cpu_baseline else cpu_baseline(wl, labels, pivot)

    extra = None
    if args.config == "planted50m" and not args.no_extra and not args.manifest:
        # R-MAT scale 24 with the plain rule: per-phase GB/s on the giant-SCC workload
        del wl, labels
        torch.cuda.empty_cache()
        wx = Workload(ka, "rmat24")
        xopts = dict(autotune.DEFAULTS)
        xt = None if args.no_tuned else tuned_options("rmat24")
        if xt:
            xopts.update(xt["opts"])
        lx = torch.empty(wx.g.n, dtype=torch.int32, device="cuda")
        xms, xst = timed(wx, lx, xopts, args.steps, args.warmup)
        extra = {"rmat24": {"workload": CONFIGS["rmat24"][3], "n": wx.g.n, "m": wx.g.m, "ms_per_step": xms,
                            "value": wx.g.m / (xms / 1e3), "unit": "edges/s", "pivot": "plain",
                            "options": xopts}}
        if not args.no_profile:
            _, xroof, xscope = profile(wx, lx, xopts, args.steps)
            extra["rmat24"]["roofline"] = xroof
            extra["rmat24"]["phases"] = phase_table(xst, xscope)

    out = {
        "metric": METRIC,
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
        "config": {"workload": desc, "name": args.manifest or
                   args.config, "n": g.n, "m": g.m, "method": METHOD_NAMES[main_opts.get("method", 0)],
                   "pivot": pivot, "fwbw_rounds": main_opts.get("fwbw_rounds"), "seed": CONFIGS[args.config][2]
                   if not args.manifest else None,
                   "options": {k: v for k, v in main_opts.items() if k != "mlp_params"}},
        "ab": ab_out,
        "autotune": tune_info,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "phases": phases,
        "phases_ms": {k: getattr(stats, k) for k in ("ms_trim", "ms_fwbw", "ms_color", "ms_residual", "ms_canon")},
        "decided": {k: getattr(stats, k) for k in ("decided_trim", "decided_trim2", "decided_fwbw", "decided_color",
                                                   "decided_residual")},
        "solve": {k: getattr(stats, k) for k in ("pivot", "fwbw_levels_push", "fwbw_levels_pull", "fwbw_levels_coop",
                                                 "fwbw_rounds_done", "color_rounds", "color_iters", "edges_scanned",
                                                 "residual_n")},
        "kernels": kernels[:10],
    }
    if batch_info is not None:
        out["batch"] = batch_info
    if extra is not None:
        out["extra"] = extra
    print(json.dumps(out))


if __name__ == "__main__":
    main()
