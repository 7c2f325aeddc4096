"""The device-driven solve (KSPN_FLAG_GRAPH): one fixed sequence of launches,
captured into a HIP graph on the first call and replayed afterwards. Labels
must equal the oracle's Tarjan labels and the host-driven solve's, on every
replay, with other GPU work and synchronisations between the replays."""
import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.gpu

GRAPH = 128  # KSPN_FLAG_GRAPH


def host(O, g):
    fo, fa, bo, ba = g.to_host()
    return O.HostGraph(g.n, fo, fa, bo, ba)


def lab_np(t):
    return t.cpu().numpy().view(np.uint32)


def test_graph_mode_fuzzy_grid(ka, O):
    """the reference's fuzzy grid (test_scc_fuzzy.cpp:31-80): planted labels"""
    for n in (1, 6, 8, 16, 33, 64):
        for run in range(8):
            hg, planted = O.gen_fuzzy(1000 * n + run, n)
            g = ka.BidiGraph.from_host(hg.n, hg.fw_off, hg.fw_adj, hg.bw_off, hg.bw_adj)
            lab, st = ka.scc(g, flags=GRAPH)
            assert np.array_equal(lab_np(lab), planted), (n, run)
            assert st.decided_trim + st.decided_fwbw + st.decided_color == n


@pytest.mark.parametrize("kind", ["gnm", "rmat", "lattice", "planted", "cycle", "dag"])
def test_graph_mode_single_graphs(ka, O, kind):
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
    lab, st = ka.scc(g, flags=GRAPH)
    assert np.array_equal(lab_np(lab), ref)
    eager, _ = ka.scc(g)
    assert np.array_equal(lab_np(eager), ref)


def test_graph_mode_replays(ka, O):
    """one capture, several replays into the same labels buffer, with eager
    solves of another graph and synchronisations in between"""
    import torch

    g = ka.generate("rmat", 11, scale=14, draws=8 << 14)
    other = ka.generate("gnm", 12, n=5000, draws=20000)
    ref = O.tarjan(host(O, g))
    lab = torch.full((g.n,), -1, dtype=torch.int32, device="cuda")
    for rep in range(4):
        lab.fill_(-1)
        ka.scc(g, lab, flags=GRAPH)
        assert np.array_equal(lab_np(lab), ref), rep
        ka.scc(other)  # eager work between replays
        torch.cuda.synchronize()


def test_graph_mode_batches(ka, O):
    """batches: a hand-made one (golden graph, a vertex, a cycle, a DAG, an
    edgeless graph, G(n,m)) and a generated one; labels local per graph,
    equal to the host-driven batch solve and to per-graph Tarjan"""
    parts = [
        (7, H.ISPAN_TINY_FW_OFF, H.ISPAN_TINY_FW_ADJ),
        (1, [0, 0], []),
        (9,) + H.cycle(9),
        (12,) + H.forward_dag(12, 20, np.random.default_rng(1)),
        (5, [0] * 6, []),
    ]
    s, d = O.gen_gnm(3, 500, 1500)
    gg = O.csr_from_edges(500, s, d)
    parts.append((500, gg.fw_off, gg.fw_adj))
    hostg = [O.HostGraph(n, np.array(fo, np.uint64), np.array(fa, np.uint32)) for n, fo, fa in parts]
    b = ka.Batch.from_graphs([ka.BidiGraph.from_host(x.n, x.fw_off, x.fw_adj, x.bw_off, x.bw_adj) for x in hostg])
    ref = np.concatenate([O.tarjan(x) for x in hostg])
    for _ in range(3):
        lab, st = ka.scc_batch(b, flags=GRAPH)
        assert np.array_equal(lab_np(lab), ref)
    assert lab_np(lab)[:7].tolist() == H.ISPAN_TINY_LABELS
    kw = dict(n_min=1000, n_max=60_000, deg_min=2.0, deg_max=16.0)
    gb = ka.generate_batch(0x6B5350AA00000004, 128, **kw)
    u, vo, eo, graphs = O.gen_batch(0x6B5350AA00000004, 128, **kw)
    ref = np.concatenate([O.tarjan(x) for x in graphs])
    eager, _ = ka.scc_batch(gb)
    assert np.array_equal(lab_np(eager), ref)
    for _ in range(3):
        lab, st = ka.scc_batch(gb, flags=GRAPH)
        assert np.array_equal(lab_np(lab), ref)
        assert st.decided_trim + st.decided_fwbw + st.decided_color == gb.graph.n


def test_graph_mode_options(ka, O):
    """unsupported combinations are refused; a round cap reports KSPN_E_ITER_CAP"""
    g = ka.generate("gnm", 5, n=4000, draws=12000)
    with pytest.raises(ka.KspnError) as ei:
        ka.scc(g, flags=GRAPH, method=1)
    assert ei.value.code == -5
    hg0, _ = O.gen_planted(77, 5000)
    pg = ka.BidiGraph.from_host(hg0.n, hg0.fw_off, hg0.fw_adj, hg0.bw_off, hg0.bw_adj)
    with pytest.raises(ka.KspnError) as ei:
        ka.scc(pg, flags=GRAPH, max_rounds=1)
    assert ei.value.code == -4
    lab, _ = ka.scc(pg, flags=GRAPH)
    assert np.array_equal(lab_np(lab), O.tarjan(hg0))
