"""GPU parity of batched solves: many independent graphs packed into one
union graph (batch-global ids, per-graph vertex and edge offsets) and
solved in one kspn_scc_batch call. Labels are local to each graph and must
equal the oracle's Tarjan ids of that graph alone, bit for bit."""
import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.gpu

BATCH_METHODS = [
    dict(method=0),
    dict(method=1),
    dict(method=0, fwbw_rounds=4),
    dict(method=0, flags=4),          # colour rounds only, no tail solver
    dict(method=0, flags=4 | 8),      # colour rounds over the union, no refinement
    dict(method=1, residual_threshold=500),
    dict(method=0, alpha=1e9, beta=1.0),  # pull levels
    dict(method=0, fwbw_rounds=3, pivot=2, seed=5),  # random pivots
]
BIDS = ["kaspan", "trim_ex", "fwbw4", "no_tail", "plain_colour", "residual500", "pull", "random_pivots"]


def local_tarjan(O, graphs):
    return np.concatenate([O.tarjan(g) for g in graphs]) if graphs else np.zeros(0, np.uint32)


def labels_of(ka, batch, **opts):
    lab, st = ka.scc_batch(batch, **opts)
    return lab.cpu().numpy().view(np.uint32), st


def test_generated_batch_matches_oracle(ka, O):
    kw = dict(n_min=40, n_max=3000, deg_min=1.0, deg_max=8.0)
    b = ka.generate_batch(21, 24, **kw)
    u, vo, eo, graphs = O.gen_batch(21, 24, **kw)
    assert b.count == 24 and b.graph.n == u.n and b.graph.m == u.m
    assert np.array_equal(b.vtx_off.cpu().numpy().view(np.uint32), vo)
    assert np.array_equal(b.edge_off.cpu().numpy().view(np.uint64), eo)
    fo, fa, bo, ba = b.graph.to_host()
    assert H.csr_digest(fo, fa) == H.csr_digest(u.fw_off, u.fw_adj)
    assert H.csr_digest(bo, ba) == H.csr_digest(u.bw_off, u.bw_adj)
    ref = local_tarjan(O, graphs)
    for opts in BATCH_METHODS:
        lab, st = labels_of(ka, b, **opts)
        assert np.array_equal(lab, ref), opts


@pytest.mark.parametrize("opts", BATCH_METHODS, ids=BIDS)
def test_hand_made_batch(ka, O, opts):
    """golden graph, cycle, DAG (all trimmed), single vertex, edgeless graph, G(n,m)"""
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
    host = [O.HostGraph(n, np.array(fo, np.uint64), np.array(fa, np.uint32)) for n, fo, fa in parts]
    dev = [ka.BidiGraph.from_host(g.n, g.fw_off, g.fw_adj, g.bw_off, g.bw_adj) for g in host]
    b = ka.Batch.from_graphs(dev)
    lab, st = labels_of(ka, b, **opts)
    assert np.array_equal(lab, local_tarjan(O, host))
    assert lab[:7].tolist() == H.ISPAN_TINY_LABELS
    # each graph alone gives the same labels
    vo = b.vtx_off.cpu().numpy().view(np.uint32)
    for i in range(b.count):
        solo, _ = ka.scc(b.subgraph(i), **opts)
        assert np.array_equal(solo.cpu().numpy().view(np.uint32), lab[vo[i]: vo[i + 1]]), i


def test_batch_where_every_graph_is_trimmed(ka, O):
    """edges only from the lower to the upper half: every vertex has in- or
    out-degree 0, so the first trim pass decides all and no graph keeps a
    live vertex for the FW-BW rounds"""
    def halves(k):
        return H.csr_from_pairs(k, [(u, v) for u in range(k // 2) for v in range(k // 2, k) if (u + v) % 3])

    host = [O.HostGraph(k, *halves(k)) for k in (3, 4, 8, 1, 11)]
    b = ka.Batch.from_graphs([ka.BidiGraph.from_host(g.n, g.fw_off, g.fw_adj) for g in host])
    lab, st = labels_of(ka, b, fwbw_rounds=2)
    assert np.array_equal(lab, local_tarjan(O, host))
    assert st.decided_trim == b.graph.n


def test_larger_generated_batch(ka, O):
    """256 graphs of 1e3..1e5 vertices, G(n,m) and R-MAT, against Tarjan on the union"""
    kw = dict(n_min=1000, n_max=100_000, deg_min=2.0, deg_max=16.0)
    b = ka.generate_batch(0x6B5350AA00000004, 256, **kw)
    u, vo, eo, graphs = O.gen_batch(0x6B5350AA00000004, 256, **kw)
    assert b.graph.m == u.m
    lab_u = O.tarjan(u)
    ref = np.empty_like(lab_u)
    for i in range(256):
        ref[vo[i]: vo[i + 1]] = lab_u[vo[i]: vo[i + 1]] - vo[i]
    for opts in (dict(method=0), dict(method=1, fwbw_rounds=2), dict(method=0, fwbw_rounds=8, flags=4)):
        lab, st = labels_of(ka, b, **opts)
        assert np.array_equal(lab, ref), opts
        assert st.fwbw_rounds_done >= 1 and st.partitions_max >= 1


def test_batch_errors(ka):
    g = ka.BidiGraph.from_host(2, [0, 1, 2], [1, 0])
    b = ka.Batch.from_graphs([g])
    b.count = 3  # more graphs than vertices
    b.vtx_off = ka._as_u32_tensor(np.array([0, 1, 1, 2], np.uint32), "cuda")
    with pytest.raises(ka.KspnError):
        ka.scc_batch(b)
