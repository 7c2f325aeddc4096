"""Solve statistics in the reference's JSON tree.

The reference records nested scopes with a ``duration`` in nanoseconds and
named unsigned entries (``KASPAN_STATISTIC_SCOPE/PUSH/ADD/POP``,
debug/statistic.hpp:89-121) and writes them as ``{"<rank>": {scope...}}``
(debug/statistic.hpp:124-205). The driver (tool/bench_kaspan/src/main.cpp:
44-126) opens a ``benchmark`` scope with process entries, a ``kagen`` or
``load`` scope, the graph sizes, the ``scc`` scope of the variant (one child
scope per stage with ``decided_count`` and, for some, ``memory``) and the
component counts. ``statistics_json`` builds that tree for one device from a
solve's ``Stats``; the reference's plotting (experiment/horeka/plot_utils.py:
151-245) reads it unchanged.

Differences: stage durations are GPU time of the stage's kernels (HIP events
on the solve's stream); the ``scc`` duration is the host's wall time around
the solve call, as in the reference; ``memory`` entries of stages are the
device's bytes in use (``hipMemGetInfo``) at the end of the stage where the
reference records the process's resident set, and the benchmark's own
``memory`` entry is the host process's resident set as in the reference.
"""
from __future__ import annotations

import json
import os


def resident_set_bytes() -> int:
    """the process's resident set (debug/process.hpp's get_resident_set_bytes)"""
    try:
        with open("/proc/self/statm") as f:
            return int(f.read().split()[1]) * os.sysconf("SC_PAGE_SIZE")
    except (OSError, ValueError, IndexError):
        return 0


def scc_scope(stats, duration_ns: int, variant: str | None = None) -> dict:
    """the ``scc`` scope: duration and one child per stage, in the order the stages ran"""
    node = {"duration": int(duration_ns)}
    for ph in stats.phases():
        child = {"duration": int(round(ph["ms"] * 1e6))}
        if ph["name"] == "trim_1_first" and variant == "ispan_like":
            child["pivot"] = int(stats.pivot)  # scc_ispan_like.hpp:38
        if ph["name"] != "select_pivot":  # the reference's select_pivot scope holds a duration only
            child["decided_count"] = int(ph["decided"])
        if ph["has_memory"]:
            child["memory"] = int(ph["memory"])
        node[ph["name"]] = child
    return node


def statistics_json(stats, graph, *, scc_ns: int, labels=None, variant: str | None = None, total_ns: int = 0,
                    source: str = "load", source_ns: int = 0, memory: int | None = None) -> dict:
    """``{"0": {"benchmark": {...}}}`` for one solve of ``graph`` (n, m) on one device.

    ``source`` is "load" (a manifest) or "kagen" (a generated graph) with its duration;
    ``labels`` (canonical ids, device or host) gives the component counts."""
    bench = {"duration": int(total_ns), "memory": int(resident_set_bytes() if memory is None else memory),
             "world_rank": 0, "world_size": 1, "memcheck": 0, "callgrind": 0,
             source: {"duration": int(source_ns)},
             "n": int(graph.n), "local_n": int(graph.n), "m": int(graph.m), "local_fw_m": int(graph.m),
             "local_bw_m": int(graph.m), "scc": scc_scope(stats, scc_ns, variant)}
    if labels is not None:
        import torch

        lab = labels if isinstance(labels, torch.Tensor) else torch.as_tensor(labels)
        ids = torch.arange(lab.numel(), device=lab.device, dtype=torch.int64)
        count = int((lab.to(torch.int64) & 0xFFFFFFFF).eq(ids).sum().item())
        bench["local_component_count"] = count
        bench["global_component_count"] = count
    return {"0": {"benchmark": bench}}


def write_json(tree: dict, path: str):
    with open(path, "w") as f:
        json.dump(tree, f, separators=(",", ":"))
