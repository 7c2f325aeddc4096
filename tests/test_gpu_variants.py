"""The driver's variants (tool/bench_kaspan/src/main.cpp:63-71) on the GPU:
canonical labels equal Tarjan's for every variant, and the statistics scopes
each variant records (names, order, decided counts) agree with the oracle's
restatement of that variant's stages:

* the first trim (trim_1_first, or trim_1_exhaustive_first -- whose two
  passes reach the trim-1 fixpoint), the pivot and the FW-BW decisions are
  equal to the oracle's;
* trim_1_normal (scc, async, ispan_like) is equal;
* scc_trim_ex's "trim_ex" scope is the trim-1 fixpoint after FW-BW, which
  holds at least what the reference's incremental trim decides there (that
  trim leaves the out-degree counters of its in-pass decisions stale,
  trim_1_exhaustive.hpp:48-69, so it can stop short of the fixpoint);
* colour and residual scopes decide the rest: every scope's count sums to n.
"""
import numpy as np
import pytest

from tests import helpers as H

pytestmark = pytest.mark.gpu

NO_TRIM2 = 2


def host_graph(O, g):
    fo, fa, bo, ba = g.to_host()
    return O.HostGraph(g.n, fo, fa, bo, ba)


def graphs(ka, O):
    yield "ispan_tiny", ka.BidiGraph.from_host(7, H.ISPAN_TINY_FW_OFF, H.ISPAN_TINY_FW_ADJ, H.ISPAN_TINY_BW_OFF,
                                               H.ISPAN_TINY_BW_ADJ)
    yield "gnm20k", ka.generate("gnm", 71, n=20000, draws=50000)
    yield "rmat13", ka.generate("rmat", 72, scale=13, draws=8 << 13)
    yield "lattice30", ka.generate("lattice", 73, side=30, p_bidir=0.1)
    hg, _ = O.gen_planted(74, 20000)
    yield "planted20k", ka.BidiGraph.from_host(hg.n, hg.fw_off, hg.fw_adj, hg.bw_off, hg.bw_adj)


def expected_reference(O, hg, variant):
    """(scope names with the oracle's decided counts where they are comparable, oracle stats)"""
    n = hg.n
    if variant in (None, "async"):
        _, st = O.kaspan_scc(hg)
        plan = [("trim_1_first", st.decided_trim_first), ("forward_backward_search", st.decided_fwbw),
                ("trim_1_normal", st.decided_trim_normal), ("color", None)]
    elif variant == "hpc_like":
        _, st = O.scc_variant(hg, 2)
        plan = [("trim_1_first", st.decided_trim_first), ("forward_backward_search", st.decided_fwbw),
                ("color", None)]
    elif variant == "ispan_like":
        _, st = O.scc_variant(hg, 3)
        plan = [("trim_1_first", st.decided_trim_first), ("forward_backward_search", st.decided_fwbw),
                ("trim_1_normal", st.decided_trim_normal), ("residual", None)]
    elif variant == "hpc_trim_ex":
        _, st = O.scc_variant(hg, 4)
        plan = [("trim_ex_first", st.decided_trim_first), ("select_pivot", 0),
                ("forward_backward_search", st.decided_fwbw), ("color", None)]
    elif variant == "trim_ex":
        _, st = O.scc_trim_ex(hg)
        plan = [("trim_ex_first", st.decided_trim_first), ("select_pivot", 0),
                ("forward_backward_search", st.decided_fwbw), ("trim_ex", "fixpoint"), ("color_trim_ex", None)]
    elif variant == "trim_ex_residual":
        _, st = O.scc_trim_ex_residual(hg)
        plan = [("trim_ex_first", st.decided_trim_first), ("residual", None)]
    else:  # trim_ex_light_residual
        _, st = O.scc_variant(hg, 6)
        plan = [("trim_ex_first", st.decided_trim_first), ("trim_ex", 0), ("residual", None)]
    # a stage runs only while vertices are undecided (each variant returns early)
    out, left = [], n
    for name, d in plan:
        if left == 0:
            break
        out.append((name, d))
        if isinstance(d, int):
            left -= d
    return out, st


@pytest.mark.parametrize("variant", [None, "async", "hpc_like", "ispan_like", "hpc_trim_ex", "trim_ex",
                                     "trim_ex_residual", "trim_ex_light_residual"])
def test_variant_labels_and_scopes(ka, O, variant):
    for gname, g in graphs(ka, O):
        print(f"variant {variant} graph {gname}", flush=True)
        hg = host_graph(O, g)
        tar = O.tarjan(hg)
        lab, st = ka.scc_variant(variant, g, flags=NO_TRIM2)
        assert np.array_equal(lab.cpu().numpy().view(np.uint32), tar), (variant, gname)
        phases = st.phases()
        got = [p["name"] for p in phases]
        plan, ref_st = expected_reference(O, hg, variant)
        names = [name for name, _ in plan]
        if variant == "trim_ex":
            # the stronger trim may leave nothing for the colour rounds
            assert got in (names, [x for x in names if x != "color_trim_ex"]), (gname, got, names)
        else:
            assert got == names, (gname, got, names)
        by = {p["name"]: p for p in phases}
        for name, d in plan:
            if name not in by:
                continue
            if isinstance(d, int) and name != "select_pivot":
                assert by[name]["decided"] == d, (variant, gname, name, by[name]["decided"], d)
            assert by[name]["entries"] == 1
        assert sum(p["decided"] for p in phases) == g.n
        if "forward_backward_search" in by:
            assert st.pivot == ref_st.pivot, (variant, gname)
        if variant == "trim_ex" and "trim_ex" in by:
            # the trim-1 fixpoint of what FW-BW left
            _, live = O.trim_fixpoint(hg)
            _, live2, _ = O.fwbw(hg, live, ref_st.pivot)
            fix, _ = O.trim_fixpoint(hg, live2)
            assert by["trim_ex"]["decided"] == fix, (gname, by["trim_ex"]["decided"], fix)
        for p in phases:
            assert p["ms"] >= 0.0


def test_scope_memory_is_opt_in(ka):
    g = ka.generate("rmat", 75, scale=12, draws=8 << 12)
    _, st = ka.scc(g)
    assert all(p["memory"] == 0 for p in st.phases())
    _, st = ka.scc(g, flags=32)
    mem = [p for p in st.phases() if p["has_memory"]]
    assert mem and all(p["memory"] > 0 for p in mem)


def test_multi_round_scopes_accumulate(ka, O):
    """trims between single-pivot FW-BW rounds add to one scope per name"""
    g = ka.generate("rmat", 76, scale=14, draws=8 << 14)
    lab, st = ka.scc(g, method=ka.METHOD_KASPAN, fwbw_rounds=1)
    names1 = [p["name"] for p in st.phases()]
    lab2, st2 = ka.scc(g, fwbw_rounds=8)
    assert np.array_equal(lab.cpu().numpy(), lab2.cpu().numpy())
    names2 = [p["name"] for p in st2.phases()]
    assert len(names2) == len(set(names2))
    assert names1[0] == names2[0] == "trim_1_first"
    assert sum(p["decided"] for p in st2.phases()) == g.n


def test_statistics_tree(ka, O):
    """the JSON tree of one solve: the reference's nesting and keys"""
    from kaspan_amd import statistics as S

    g = ka.generate("gnm", 77, n=30000, draws=90000)
    lab, st = ka.scc_variant("hpc_like", g, flags=32)
    tree = S.statistics_json(st, g, scc_ns=1234, labels=lab, variant="hpc_like", total_ns=5678, source="kagen",
                             source_ns=99)
    b = tree["0"]["benchmark"]
    assert list(b)[:6] == ["duration", "memory", "world_rank", "world_size", "memcheck", "callgrind"]
    assert b["kagen"] == {"duration": 99}
    assert b["scc"]["duration"] == 1234
    stages = {k: v for k, v in b["scc"].items() if isinstance(v, dict)}
    assert list(stages) == [p["name"] for p in st.phases()]
    assert "memory" in stages["forward_backward_search"] and "memory" not in stages["trim_1_first"]
    assert sum(v["decided_count"] for v in stages.values()) == g.n
    tar = O.tarjan(host_graph(O, g))
    assert b["global_component_count"] == b["local_component_count"] == int(np.sum(tar == np.arange(g.n)))


@pytest.mark.parametrize("variant", [None, "hpc_like", "ispan_like", "trim_ex"])
def test_scopes_when_an_early_stage_decides_everything(ka, O, variant):
    """the reference returns after the stage that decides the last vertex
    (scc.hpp:44,60,70): a bipartite DAG ends after the first trim, a cycle
    after FW-BW"""
    import numpy as np

    n = 1000
    # sources -> sinks: every vertex has in- or out-degree 0
    rng = np.random.default_rng(3)
    src = rng.integers(0, n // 2, 4000)
    dst = rng.integers(n // 2, n, 4000)
    g = ka.BidiGraph.from_edges(n, src, dst)
    lab, st = ka.scc_variant(variant, g, flags=NO_TRIM2)
    assert lab.cpu().tolist() == list(range(n))
    first = "trim_ex_first" if variant == "trim_ex" else "trim_1_first"
    assert [p["name"] for p in st.phases()] == [first]
    assert st.phases()[0]["decided"] == n
    # one directed cycle with chords: FW-BW decides everything
    src = np.concatenate([np.arange(n), np.arange(0, n, 3)])
    dst = np.concatenate([(np.arange(n) + 1) % n, (np.arange(0, n, 3) + 5) % n])
    g = ka.BidiGraph.from_edges(n, src, dst)
    lab, st = ka.scc_variant(variant, g, flags=NO_TRIM2)
    assert lab.cpu().tolist() == [0] * n
    names = [p["name"] for p in st.phases()]
    expect = [first, "select_pivot", "forward_backward_search"] if variant == "trim_ex" else \
        [first, "forward_backward_search"]
    assert names == expect, names
    assert st.phases()[-1]["decided"] == n
    assert st.pivot < n
