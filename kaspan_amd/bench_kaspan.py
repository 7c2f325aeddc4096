"""Command-line driver mirroring the reference's ``bench_kaspan``
(tool/bench_kaspan/src/main.cpp:32-126): load a graph, run one SCC variant
on the GPU, write the statistics tree as JSON.

    python -m kaspan_amd.bench_kaspan (--manifest_file F | --graph SPEC) --output_file OUT
                                      [--variant V] [--warmup W] [--trim2]

``--variant`` takes the reference's names (async, hpc_like, hpc_trim_ex,
ispan_like, trim_ex, trim_ex_light_residual, trim_ex_residual; none for
kaspan::scc). ``--graph`` replaces ``--kagen_option_string`` (KaGen is not
part of this build): ``gnm:n=65536,draws=524288,seed=1``,
``rmat:scale=24,draws=268435456``, ``lattice:side=256``, ``planted:n=50000000``.
``--warmup`` untimed solves come first (the first solve on a context
allocates its scratch); the reference runs the solve once.
"""
from __future__ import annotations

import argparse
import sys
import time

import torch


def parse_graph_spec(spec: str) -> tuple[str, int, dict]:
    kind, _, rest = spec.partition(":")
    kw = {}
    seed = 1
    for item in filter(None, rest.split(",")):
        k, _, v = item.partition("=")
        if k == "seed":
            seed = int(v)
        elif k in ("a", "b", "c", "p_bidir", "expo", "intra_deg", "inter_deg"):
            kw[k] = float(v)
        else:
            kw[k] = int(v)
    return kind, seed, kw


def main(argv=None) -> int:
    p = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    src = p.add_mutually_exclusive_group(required=True)
    src.add_argument("--manifest_file")
    src.add_argument("--graph", help="generated graph, e.g. rmat:scale=20,draws=16777216")
    src.add_argument("--kagen_option_string", help=argparse.SUPPRESS)
    p.add_argument("--output_file", required=True)
    p.add_argument("--variant", default=None)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--trim2", action="store_true", help="trim_ex with the trim-2 pass (its own \"trim_2\" scope)")
    args = p.parse_args(argv)

    import kaspan_amd as ka
    from kaspan_amd import _lib
    from kaspan_amd import manifest as M
    from kaspan_amd import statistics as S

    if args.kagen_option_string is not None:
        print("KaGen is not part of this build; use --graph (gnm/rmat/lattice/planted) or --manifest_file",
              file=sys.stderr)
        return 1
    if args.variant not in ka.VARIANTS:
        p.print_usage(sys.stderr)
        print(f"unknown variant {args.variant!r}", file=sys.stderr)
        return 1
    t_start = time.perf_counter_ns()
    rss = S.resident_set_bytes()
    t0 = time.perf_counter_ns()
    if args.manifest_file:
        source = "load"
        g = M.load_manifest(args.manifest_file)
    else:
        source = "kagen"
        kind, seed, kw = parse_graph_spec(args.graph)
        g = ka.generate(kind, seed, **kw)
    torch.cuda.synchronize()
    source_ns = time.perf_counter_ns() - t0
    labels = torch.empty(g.n, dtype=torch.int32, device=g.device)
    # the reference's stages only, unless --trim2 adds the trim-2 pass to trim_ex
    opts = dict(flags=_lib.FLAG_STATS_MEMORY | (0 if args.trim2 else _lib.FLAG_NO_TRIM2))
    for _ in range(args.warmup):
        ka.scc_variant(args.variant, g, labels, **opts)
    torch.cuda.synchronize()
    t0 = time.perf_counter_ns()
    _, st = ka.scc_variant(args.variant, g, labels, **opts)
    torch.cuda.synchronize()
    scc_ns = time.perf_counter_ns() - t0
    tree = S.statistics_json(st, g, scc_ns=scc_ns, labels=labels, variant=args.variant,
                             total_ns=time.perf_counter_ns() - t_start, source=source, source_ns=source_ns,
                             memory=rss)
    S.write_json(tree, args.output_file)
    b = tree["0"]["benchmark"]
    print(f"n={b['n']} m={b['m']} variant={args.variant or 'scc'} scc={scc_ns / 1e6:.3f} ms "
          f"components={b.get('global_component_count')}", file=sys.stderr)
    return 0


if __name__ == "__main__":
    sys.exit(main())
