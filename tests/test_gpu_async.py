"""The asynchronous reach (KSPN_FLAG_ASYNC): forward and backward searches
run as one persistent kernel over a device work queue instead of one launch
per BFS level. Labels must equal the oracle's Tarjan labels (and the planted
labels where the graph has them) for every method and round count, single
graphs and batches."""
import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.gpu

ASYNC = 256  # KSPN_FLAG_ASYNC
FULL_BW = 64  # KSPN_FLAG_FULL_BW


def host(O, g):
    fo, fa, bo, ba = g.to_host()
    return O.HostGraph(g.n, fo, fa, bo, ba)


def lab_np(t):
    return t.cpu().numpy().view(np.uint32)


def test_async_fuzzy_grid(ka, O):
    """the reference's fuzzy grid (test_scc_fuzzy.cpp:31-80): planted labels"""
    for n in (1, 6, 8, 16, 33, 64, 300):
        for run in range(6):
            hg, planted = O.gen_fuzzy(2000 * n + run, n)
            g = ka.BidiGraph.from_host(hg.n, hg.fw_off, hg.fw_adj, hg.bw_off, hg.bw_adj)
            for rounds in (1, 4):
                lab, st = ka.scc(g, flags=ASYNC, fwbw_rounds=rounds)
                assert np.array_equal(lab_np(lab), planted), (n, run, rounds)
                assert st.decided_trim + st.decided_trim2 + st.decided_fwbw + st.decided_color + st.decided_residual == n


@pytest.mark.parametrize("kind", ["gnm", "rmat", "lattice", "planted", "cycle", "dag", "hub"])
@pytest.mark.parametrize("opts", [dict(), dict(fwbw_rounds=8), dict(method=1), dict(fwbw_rounds=4, flags=FULL_BW),
                                  dict(fwbw_rounds=16, coop_edges=1)])
def test_async_single_graphs(ka, O, kind, opts):
    if kind == "gnm":
        g = ka.generate("gnm", 7, n=20000, draws=60000)
    elif kind == "rmat":
        g = ka.generate("rmat", 8, scale=15, draws=16 << 15)
    elif kind == "lattice":
        g = ka.generate("lattice", 9, side=24, p_bidir=0.1)
    elif kind == "planted":
        hg0, _ = O.gen_planted(16, 40000, max_size=8000)
        g = ka.BidiGraph.from_host(hg0.n, hg0.fw_off, hg0.fw_adj, hg0.bw_off, hg0.bw_adj)
    elif kind == "cycle":
        off, adj = H.cycle(5000)
        g = ka.BidiGraph.from_host(5000, off, adj)
    elif kind == "dag":
        off, adj = H.forward_dag(3000, 9000, np.random.default_rng(3))
        g = ka.BidiGraph.from_host(3000, off, adj)
    else:
        # a hub whose rows exceed the staging path's direct limit, inside a cycle
        n = 3000
        src = np.concatenate([np.zeros(n - 1, np.int64), np.arange(1, n), np.arange(1, n - 1)])
        dst = np.concatenate([np.arange(1, n), np.zeros(n - 1, np.int64), np.arange(2, n)])
        hg0 = O.csr_from_edges(n, src, dst)
        g = ka.BidiGraph.from_host(n, hg0.fw_off, hg0.fw_adj)
    ref = O.tarjan(host(O, g))
    o = dict(opts)
    flags = ASYNC | o.pop("flags", 0)
    lab, st = ka.scc(g, flags=flags, **o)
    assert np.array_equal(lab_np(lab), ref)
    assert st.decided_trim + st.decided_trim2 + st.decided_fwbw + st.decided_color + st.decided_residual == g.n


def test_async_repeat_and_batches(ka, O):
    """repeated solves reuse the queue (it must be left empty), and batches of
    graphs, with labels local per graph"""
    g = ka.generate("rmat", 11, scale=14, draws=8 << 14)
    ref = O.tarjan(host(O, g))
    for rep in range(3):
        lab, _ = ka.scc(g, flags=ASYNC, fwbw_rounds=4)
        assert np.array_equal(lab_np(lab), ref), rep
    kw = dict(n_min=1000, n_max=60_000, deg_min=2.0, deg_max=16.0)
    gb = ka.generate_batch(0x6B5350AA00000004, 128, **kw)
    u, vo, eo, graphs = O.gen_batch(0x6B5350AA00000004, 128, **kw)
    ref = np.concatenate([O.tarjan(x) for x in graphs])
    lab, st = ka.scc_batch(gb, flags=ASYNC)
    assert np.array_equal(lab_np(lab), ref)
    assert st.decided_trim + st.decided_trim2 + st.decided_fwbw + st.decided_color + st.decided_residual == gb.graph.n
