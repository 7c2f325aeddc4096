"""Edge pruning (KSPN_FLAG_PRUNE): an edge that a forward push level of the
multi-pivot rounds, a colour pull sweep or a colour push finds joining two
partitions, or reaching a decided vertex, is marked and skipped by later
levels. Partitions only split, so labels must equal the oracle's Tarjan
labels (the planted labels on the reference's fuzzy grid), and every round
must decide exactly what it decides without pruning: equal level counts,
equal colour rounds and equal decided counts per stage. (The number of
colour sweeps is not compared: a sweep reads words that other threads of
the same sweep may already have raised, so it can vary from run to run.)"""
import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.gpu

PRUNE = 4096  # KSPN_FLAG_PRUNE
NARROW = 2048
NO_COOP = 16
FULL_BW = 64


def host(O, g):
    fo, fa, bo, ba = g.to_host()
    return O.HostGraph(g.n, fo, fa, bo, ba)


def lab_np(t):
    return t.cpu().numpy().view(np.uint32)


def test_prune_fuzzy_grid(ka, O):
    """the reference's fuzzy grid (test_scc_fuzzy.cpp:31-80): planted labels"""
    for n in (1, 6, 8, 16, 33, 64, 300):
        for run in range(6):
            hg, planted = O.gen_fuzzy(9000 * n + run, n)
            g = ka.BidiGraph.from_host(hg.n, hg.fw_off, hg.fw_adj, hg.bw_off, hg.bw_adj)
            for rounds, flags in ((2, NO_COOP), (8, NO_COOP), (8, NARROW)):
                lab, st = ka.scc(g, flags=PRUNE | flags, fwbw_rounds=rounds, alpha=1e-9)
                assert np.array_equal(lab_np(lab), planted), (n, run, rounds, flags)


@pytest.mark.parametrize("kind", ["gnm", "rmat", "lattice", "planted", "cycle", "dag"])
@pytest.mark.parametrize("opts", [dict(fwbw_rounds=1), dict(method=1, fwbw_rounds=4),
                                  dict(fwbw_rounds=8), dict(fwbw_rounds=16, alpha=1e-9),
                                  dict(fwbw_rounds=8, alpha=1e-9, flags=NARROW),
                                  dict(fwbw_rounds=4, alpha=1e-9, flags=FULL_BW),
                                  dict(fwbw_rounds=8, pivot=2, seed=5)])
def test_prune_single_graphs(ka, O, kind, opts):
    if kind == "gnm":
        g = ka.generate("gnm", 37, n=20000, draws=60000)
    elif kind == "rmat":
        g = ka.generate("rmat", 38, scale=15, draws=16 << 15)
    elif kind == "lattice":
        g = ka.generate("lattice", 39, side=24, p_bidir=0.1)
    elif kind == "planted":
        hg0, _ = O.gen_planted(46, 40000, max_size=8000)
        g = ka.BidiGraph.from_host(hg0.n, hg0.fw_off, hg0.fw_adj, hg0.bw_off, hg0.bw_adj)
    elif kind == "cycle":
        off, adj = H.cycle(5000)
        g = ka.BidiGraph.from_host(5000, off, adj)
    else:
        off, adj = H.forward_dag(3000, 9000, np.random.default_rng(43))
        g = ka.BidiGraph.from_host(3000, off, adj)
    ref = O.tarjan(host(O, g))
    o = dict(opts)
    flags = PRUNE | o.pop("flags", 0)
    lab, st = ka.scc(g, flags=flags, **o)
    assert np.array_equal(lab_np(lab), ref)


@pytest.mark.parametrize("flags", [NO_COOP, NARROW])
def test_prune_rounds_claim_the_same(ka, O, flags):
    """pruned and unpruned rounds take the same levels and decide the same
    vertices; a second solve on the same context starts from a clear
    bitmap"""
    hg0, planted = O.gen_planted(47, 300000, max_size=60000)
    g = ka.BidiGraph.from_host(hg0.n, hg0.fw_off, hg0.fw_adj, hg0.bw_off, hg0.bw_adj)
    kw = dict(fwbw_rounds=16, alpha=1e-9)
    lab_e, st_e = ka.scc(g, flags=flags, **kw)
    assert np.array_equal(lab_np(lab_e), planted)
    for _ in range(2):
        lab_p, st_p = ka.scc(g, flags=flags | PRUNE, **kw)
        assert np.array_equal(lab_np(lab_p), planted)
        assert st_p.fwbw_levels_push == st_e.fwbw_levels_push
        assert st_p.fwbw_levels_coop == st_e.fwbw_levels_coop
        assert st_p.decided_fwbw == st_e.decided_fwbw
        assert st_p.fwbw_rounds_done == st_e.fwbw_rounds_done
        assert st_p.decided_color == st_e.decided_color
        assert st_p.color_rounds == st_e.color_rounds


def test_prune_graph_sizes_and_batch(ka, O):
    """the bitmap grows with the graph; a batch keeps labels local per graph"""
    for n, draws in ((1 << 12, 1 << 14), (1 << 18, 1 << 21)):
        g = ka.generate("gnm", 51, n=n, draws=draws)
        ref = O.tarjan(host(O, g))
        lab, st = ka.scc(g, flags=PRUNE | NARROW, fwbw_rounds=12)
        assert np.array_equal(lab_np(lab), ref), n
    kw = dict(n_min=1000, n_max=60_000, deg_min=2.0, deg_max=16.0)
    gb = ka.generate_batch(0x6B5350AA00000004, 128, **kw)
    u, vo, eo, graphs = O.gen_batch(0x6B5350AA00000004, 128, **kw)
    ref = np.concatenate([O.tarjan(x) for x in graphs])
    for flags in (PRUNE, PRUNE | NARROW):
        lab, st = ka.scc_batch(gb, flags=flags, fwbw_rounds=4)
        assert np.array_equal(lab_np(lab), ref), flags
