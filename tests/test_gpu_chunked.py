"""Chunked push levels (KSPN_FLAG_CHUNKED): a reach frontier whose vertices
have at most 64 out-edges each, and every colour push level, is expanded by
chunks of 256 frontier entries with their degrees scanned in LDS, without
the device-wide degree scan.
Labels must equal the oracle's Tarjan labels (the planted labels on the
reference's fuzzy grid), and a reach must claim the same vertices level by
level as the edge-balanced expansion: equal level counts and decided
counts. Frontiers with a vertex of higher degree (R-MAT, the hub of the
star) go to the edge-balanced expansion."""
import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.gpu

CHUNKED = 8192  # KSPN_FLAG_CHUNKED
PRUNE = 4096
NARROW = 2048
NO_COOP = 16
FULL_BW = 64


def host(O, g):
    fo, fa, bo, ba = g.to_host()
    return O.HostGraph(g.n, fo, fa, bo, ba)


def lab_np(t):
    return t.cpu().numpy().view(np.uint32)


def test_chunked_fuzzy_grid(ka, O):
    """the reference's fuzzy grid (test_scc_fuzzy.cpp:31-80): planted labels"""
    for n in (1, 6, 8, 16, 33, 64, 300):
        for run in range(6):
            hg, planted = O.gen_fuzzy(11000 * n + run, n)
            g = ka.BidiGraph.from_host(hg.n, hg.fw_off, hg.fw_adj, hg.bw_off, hg.bw_adj)
            for rounds, flags in ((1, NO_COOP), (8, NO_COOP), (8, NARROW | PRUNE)):
                lab, st = ka.scc(g, flags=CHUNKED | flags, fwbw_rounds=rounds, alpha=1e-9)
                assert np.array_equal(lab_np(lab), planted), (n, run, rounds, flags)


@pytest.mark.parametrize("kind", ["gnm", "rmat", "lattice", "planted", "cycle", "dag", "star"])
@pytest.mark.parametrize("opts", [dict(), dict(alpha=1e-9), dict(fwbw_rounds=8, alpha=1e-9),
                                  dict(method=1, fwbw_rounds=4, alpha=1e-9),
                                  dict(fwbw_rounds=8, alpha=1e-9, flags=NARROW | PRUNE),
                                  dict(fwbw_rounds=4, alpha=1e-9, flags=FULL_BW)])
def test_chunked_single_graphs(ka, O, kind, opts):
    if kind == "gnm":
        g = ka.generate("gnm", 57, n=20000, draws=60000)
    elif kind == "rmat":
        g = ka.generate("rmat", 58, scale=15, draws=16 << 15)
    elif kind == "lattice":
        g = ka.generate("lattice", 59, side=24, p_bidir=0.1)
    elif kind == "planted":
        hg0, _ = O.gen_planted(66, 40000, max_size=8000)
        g = ka.BidiGraph.from_host(hg0.n, hg0.fw_off, hg0.fw_adj, hg0.bw_off, hg0.bw_adj)
    elif kind == "cycle":
        off, adj = H.cycle(5000)
        g = ka.BidiGraph.from_host(5000, off, adj)
    elif kind == "star":
        # a cycle of 4000 with a hub of out-degree 200 at 2000 and the pivot (in and out degree 21) at 10: the
        # level that holds the hub goes to the edge-balanced expansion
        n = 4000
        cyc = np.arange(n)
        src = np.concatenate([cyc, np.full(200, 2000), np.arange(3000, 3020), np.full(20, 10)])
        dst = np.concatenate([(cyc + 1) % n, np.arange(100, 300), np.full(20, 10), np.arange(3100, 3120)])
        g = ka.BidiGraph.from_edges(n, src, dst)
    else:
        off, adj = H.forward_dag(3000, 9000, np.random.default_rng(63))
        g = ka.BidiGraph.from_host(3000, off, adj)
    ref = O.tarjan(host(O, g))
    o = dict(opts)
    flags = CHUNKED | NO_COOP | o.pop("flags", 0)
    lab, st = ka.scc(g, flags=flags, **o)
    assert np.array_equal(lab_np(lab), ref)


def test_chunked_levels_equal_edge_balanced(ka, O):
    """same levels and decided vertices as the edge-balanced expansion:
    a single pivot (push only) and the multi-pivot rounds of a planted graph"""
    g = ka.generate("gnm", 61, n=200000, draws=1000000)
    ref = O.tarjan(host(O, g))
    for extra in (dict(alpha=1e-9), dict(alpha=1e-9, fwbw_rounds=6), dict(fwbw_rounds=6)):
        lab_c, st_c = ka.scc(g, flags=CHUNKED | NO_COOP, **extra)
        lab_e, st_e = ka.scc(g, flags=NO_COOP, **extra)
        assert np.array_equal(lab_np(lab_c), ref)
        assert np.array_equal(lab_np(lab_e), ref)
        assert st_c.fwbw_levels_push == st_e.fwbw_levels_push, extra
        assert st_c.fwbw_levels_pull == st_e.fwbw_levels_pull, extra
        assert st_c.decided_fwbw == st_e.decided_fwbw, extra
        if "alpha" in extra:
            # push levels only: the edges they expand are fixed by the levels (a pull sweep's scan
            # of a candidate ends at the first reached in-neighbour, which may be claimed meanwhile)
            assert st_c.edges_scanned == st_e.edges_scanned, extra
    hg0, planted = O.gen_planted(67, 300000, max_size=60000)
    g = ka.BidiGraph.from_host(hg0.n, hg0.fw_off, hg0.fw_adj, hg0.bw_off, hg0.bw_adj)
    for flags in (NO_COOP, NARROW | PRUNE):
        lab_c, st_c = ka.scc(g, flags=CHUNKED | flags, fwbw_rounds=16, alpha=1e-9)
        lab_e, st_e = ka.scc(g, flags=flags, fwbw_rounds=16, alpha=1e-9)
        assert np.array_equal(lab_np(lab_c), planted)
        assert np.array_equal(lab_np(lab_e), planted)
        assert st_c.fwbw_levels_push == st_e.fwbw_levels_push
        assert st_c.fwbw_levels_pull == st_e.fwbw_levels_pull
        assert st_c.fwbw_levels_coop == st_e.fwbw_levels_coop
        assert st_c.decided_fwbw == st_e.decided_fwbw
        assert st_c.decided_color == st_e.decided_color
        assert st_c.color_rounds == st_e.color_rounds


def test_chunked_batch(ka, O):
    """a batch keeps labels local per graph"""
    kw = dict(n_min=1000, n_max=60_000, deg_min=2.0, deg_max=16.0)
    gb = ka.generate_batch(0x6B5350AA00000004, 128, **kw)
    u, vo, eo, graphs = O.gen_batch(0x6B5350AA00000004, 128, **kw)
    ref = np.concatenate([O.tarjan(x) for x in graphs])
    for flags in (CHUNKED, CHUNKED | NARROW | PRUNE):
        lab, st = ka.scc_batch(gb, flags=flags, alpha=1e-9)
        assert np.array_equal(lab_np(lab), ref), flags


def test_chunked_profile_names(ka, O):
    """chunked levels are timed under their own profile names, so a level
    count and an average launch time name one kernel each"""
    g = ka.generate("gnm", 61, n=200000, draws=1000000)
    lab, st = ka.scc(g, flags=CHUNKED | NO_COOP, alpha=1e-9, profile=True)
    k = {x["name"]: x["launches"] for x in st.as_dict()["kernels"]}
    assert k.get("expand_chunk_fw", 0) > 0 and k.get("expand_chunk_bw", 0) > 0, k
    lab, st = ka.scc(g, flags=NO_COOP, alpha=1e-9, profile=True)
    k = {x["name"]: x["launches"] for x in st.as_dict()["kernels"]}
    assert k.get("expand_fw", 0) > 0 and "expand_chunk_fw" not in k and "expand_chunk_bw" not in k, k
