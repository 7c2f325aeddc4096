"""Chained push levels (KSPN_FLAG_CHAIN): the push levels the persistent
kernel would take run as chains of single-level launches with one host
round trip per chain. Labels must equal the oracle's Tarjan labels (planted
labels on the fuzzy grid) for every chain length, including chains that end
early and chains of one level, single graphs and batches."""
import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.gpu

CHAIN = 512  # KSPN_FLAG_CHAIN


def host(O, g):
    fo, fa, bo, ba = g.to_host()
    return O.HostGraph(g.n, fo, fa, bo, ba)


def lab_np(t):
    return t.cpu().numpy().view(np.uint32)


def test_chain_fuzzy_grid(ka, O):
    """the reference's fuzzy grid (test_scc_fuzzy.cpp:31-80): planted labels"""
    for n in (1, 6, 8, 16, 33, 64, 300):
        for run in range(6):
            hg, planted = O.gen_fuzzy(3000 * n + run, n)
            g = ka.BidiGraph.from_host(hg.n, hg.fw_off, hg.fw_adj, hg.bw_off, hg.bw_adj)
            for rounds, chain in ((1, 0), (4, 1), (4, 3)):
                lab, st = ka.scc(g, flags=CHAIN, fwbw_rounds=rounds, chain_levels=chain)
                assert np.array_equal(lab_np(lab), planted), (n, run, rounds, chain)


@pytest.mark.parametrize("kind", ["gnm", "rmat", "lattice", "planted", "cycle", "dag"])
@pytest.mark.parametrize("opts", [dict(), dict(fwbw_rounds=8), dict(method=1), dict(chain_levels=1),
                                  dict(chain_levels=64, coop_edges=1 << 12), dict(fwbw_rounds=4, flags=64),
                                  dict(coop_workgroups=7)])
def test_chain_single_graphs(ka, O, kind, opts):
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
    else:
        off, adj = H.forward_dag(3000, 9000, np.random.default_rng(3))
        g = ka.BidiGraph.from_host(3000, off, adj)
    ref = O.tarjan(host(O, g))
    o = dict(opts)
    flags = CHAIN | o.pop("flags", 0)
    lab, st = ka.scc(g, flags=flags, **o)
    assert np.array_equal(lab_np(lab), ref)
    assert st.decided_trim + st.decided_trim2 + st.decided_fwbw + st.decided_color + st.decided_residual == g.n
    plain, pst = ka.scc(g, **{k: v for k, v in opts.items() if k not in ("chain_levels",)})
    assert np.array_equal(lab_np(plain), ref)


def test_chain_levels_counted_and_batches(ka, O):
    """a long-diameter graph runs its levels in chains (counted as device
    levels); batches keep labels local per graph"""
    off, adj = H.cycle(20000)
    g = ka.BidiGraph.from_host(20000, off, adj)
    lab, st = ka.scc(g, flags=CHAIN, chain_levels=8)
    assert np.array_equal(lab_np(lab), np.zeros(20000, np.uint32))
    assert st.fwbw_levels_coop > 0
    kw = dict(n_min=1000, n_max=60_000, deg_min=2.0, deg_max=16.0)
    gb = ka.generate_batch(0x6B5350AA00000004, 128, **kw)
    u, vo, eo, graphs = O.gen_batch(0x6B5350AA00000004, 128, **kw)
    ref = np.concatenate([O.tarjan(x) for x in graphs])
    lab, st = ka.scc_batch(gb, flags=CHAIN)
    assert np.array_equal(lab_np(lab), ref)


def test_chain_option_bounds(ka):
    g = ka.generate("gnm", 5, n=4000, draws=12000)
    with pytest.raises(ka.KspnError) as ei:
        ka.scc(g, flags=CHAIN, chain_levels=257)
    assert ei.value.code == -1


NARROW = 2048  # KSPN_FLAG_NARROW


def test_narrow_fuzzy_grid(ka, O):
    """narrow chained levels on the reference's fuzzy grid: planted labels"""
    for n in (1, 6, 8, 16, 33, 64, 300):
        for run in range(6):
            hg, planted = O.gen_fuzzy(5000 * n + run, n)
            g = ka.BidiGraph.from_host(hg.n, hg.fw_off, hg.fw_adj, hg.bw_off, hg.bw_adj)
            for rounds, chain in ((1, 0), (4, 1), (4, 3)):
                lab, st = ka.scc(g, flags=NARROW, fwbw_rounds=rounds, chain_levels=chain)
                assert np.array_equal(lab_np(lab), planted), (n, run, rounds, chain)


@pytest.mark.parametrize("kind", ["gnm", "rmat", "lattice", "planted", "cycle", "dag", "star"])
@pytest.mark.parametrize("opts", [dict(), dict(fwbw_rounds=8), dict(method=1), dict(chain_levels=1),
                                  dict(chain_levels=2, fwbw_rounds=16), dict(fwbw_rounds=4, flags=64),
                                  dict(flags=CHAIN, coop_edges=1 << 12)])
def test_narrow_single_graphs(ka, O, kind, opts):
    """labels equal Tarjan's; R-MAT and the star hand levels whose frontier
    holds a vertex of degree above 64 back to the edge-balanced expansion"""
    if kind == "gnm":
        g = ka.generate("gnm", 17, n=20000, draws=60000)
    elif kind == "rmat":
        g = ka.generate("rmat", 18, scale=15, draws=16 << 15)
    elif kind == "lattice":
        g = ka.generate("lattice", 19, side=24, p_bidir=0.1)
    elif kind == "planted":
        hg0, _ = O.gen_planted(26, 40000, max_size=8000)
        g = ka.BidiGraph.from_host(hg0.n, hg0.fw_off, hg0.fw_adj, hg0.bw_off, hg0.bw_adj)
    elif kind == "cycle":
        off, adj = H.cycle(5000)
        g = ka.BidiGraph.from_host(5000, off, adj)
    elif kind == "star":
        # a cycle of 4000 with a hub of out-degree 200 at 2000 and the pivot (in and out degree 21) at 10: the
        # forward reach claims the hub in mid-reach, whose level is not narrow
        n = 4000
        cyc = np.arange(n)
        src = np.concatenate([cyc, np.full(200, 2000), np.arange(3000, 3020), np.full(20, 10)])
        dst = np.concatenate([(cyc + 1) % n, np.arange(100, 300), np.full(20, 10), np.arange(3100, 3120)])
        g = ka.BidiGraph.from_edges(n, src, dst)
    else:
        off, adj = H.forward_dag(3000, 9000, np.random.default_rng(13))
        g = ka.BidiGraph.from_host(3000, off, adj)
    ref = O.tarjan(host(O, g))
    o = dict(opts)
    flags = NARROW | o.pop("flags", 0)
    lab, st = ka.scc(g, flags=flags, **o)
    assert np.array_equal(lab_np(lab), ref)
    assert st.decided_trim + st.decided_trim2 + st.decided_fwbw + st.decided_color + st.decided_residual == g.n
    # the same run again: context state left by narrow chains does not leak into the next solve
    lab2, _ = ka.scc(g, flags=flags, **o)
    assert np.array_equal(lab_np(lab2), ref)


def test_narrow_levels_counted_and_batches(ka, O):
    """a cycle's levels run narrow on the device, as many as host-driven
    levels; a planted graph's multi-pivot rounds decide the same vertices
    narrow and host-driven; batches keep labels local per graph"""
    off, adj = H.cycle(5000)
    g = ka.BidiGraph.from_host(5000, off, adj)
    lab, st_n = ka.scc(g, flags=NARROW, chain_levels=8)
    assert np.array_equal(lab_np(lab), np.zeros(5000, np.uint32))
    lab, st_h = ka.scc(g, flags=16)
    assert np.array_equal(lab_np(lab), np.zeros(5000, np.uint32))
    assert st_n.fwbw_levels_coop > 0
    assert st_n.fwbw_levels_push + st_n.fwbw_levels_coop == st_h.fwbw_levels_push
    hg0, planted = O.gen_planted(27, 200000, max_size=50000)
    g = ka.BidiGraph.from_host(hg0.n, hg0.fw_off, hg0.fw_adj, hg0.bw_off, hg0.bw_adj)
    lab, st_n = ka.scc(g, flags=NARROW, fwbw_rounds=16, alpha=1e-9)
    assert np.array_equal(lab_np(lab), planted)
    lab, st_h = ka.scc(g, flags=16, fwbw_rounds=16, alpha=1e-9)
    assert np.array_equal(lab_np(lab), planted)
    assert st_n.fwbw_levels_coop > 0
    assert st_n.decided_fwbw == st_h.decided_fwbw
    kw = dict(n_min=1000, n_max=60_000, deg_min=2.0, deg_max=16.0)
    gb = ka.generate_batch(0x6B5350AA00000004, 128, **kw)
    u, vo, eo, graphs = O.gen_batch(0x6B5350AA00000004, 128, **kw)
    ref = np.concatenate([O.tarjan(x) for x in graphs])
    lab, st = ka.scc_batch(gb, flags=NARROW)
    assert np.array_equal(lab_np(lab), ref)

