"""CPU tests of the oracle: golden vectors from the reference's own tests,
closed forms, brute force and planted labels pin every oracle function."""
import json
import os

import numpy as np
import pytest

import helpers as H

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def tiny(O):
    return O.HostGraph(7, H.ISPAN_TINY_FW_OFF, H.ISPAN_TINY_FW_ADJ)


# ---------------------------------------------------------------- RNG
def test_splitmix64_known_answers(O):
    # published SplitMix64 outputs (Steele, Lea, Flood; Vigna's splitmix64.c)
    assert [O.lib().orc_splitmix64_at(0, i) for i in range(3)] == [
        0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    assert [O.lib().orc_splitmix64_at(1234567, i) for i in range(2)] == [6457827717110365317, 3203168211198807973]


@pytest.mark.parametrize("bits", [1, 2, 5, 12, 16])
def test_perm_pow2_is_bijection(O, bits):
    img = {O.lib().orc_perm_pow2(99, bits, x) for x in range(1 << bits)}
    assert img == set(range(1 << bits))


# ---------------------------------------------------------------- golden vectors
def test_ispan_tiny_golden(O):
    g = tiny(O)
    # the transpose of the fw CSR is the bw CSR written in the test
    assert g.bw_off.tolist() == H.ISPAN_TINY_BW_OFF
    assert g.bw_adj.tolist() == H.ISPAN_TINY_BW_ADJ
    assert O.tarjan(g).tolist() == H.ISPAN_TINY_LABELS
    assert O.brute_force(g).tolist() == H.ISPAN_TINY_LABELS
    for fn in (O.kaspan_scc, O.scc_trim_ex, O.scc_trim_ex_residual):
        lab, _ = fn(g)
        assert O.canonicalize(lab).tolist() == H.ISPAN_TINY_LABELS


def test_transpose_golden(O):
    bo, ba = O.transpose(6, np.array(H.SUB_FW_OFF, np.uint64), np.array(H.SUB_FW_ADJ, np.uint32))
    assert bo.tolist() == H.SUB_BW_OFF
    assert ba.tolist() == H.SUB_BW_ADJ


def test_subgraph_golden(O):
    g = O.HostGraph(6, H.SUB_FW_OFF, H.SUB_FW_ADJ, H.SUB_BW_OFF, H.SUB_BW_ADJ)
    sid, so, sa = O.subgraph(g, H.SUB_MASK)
    assert sid.tolist() == H.SUB_SUPER_IDS
    assert so.tolist() == H.SUB_SUB_FW_OFF and sa.tolist() == H.SUB_SUB_FW_ADJ
    gb = O.HostGraph(6, H.SUB_BW_OFF, H.SUB_BW_ADJ, H.SUB_FW_OFF, H.SUB_FW_ADJ)
    sid, so, sa = O.subgraph(gb, H.SUB_MASK)
    assert so.tolist() == H.SUB_SUB_BW_OFF and sa.tolist() == H.SUB_SUB_BW_ADJ


def test_transpose_invariants(O):
    s, d = O.gen_gnm(5, 300, 2000)
    g = O.csr_from_edges(300, s, d)
    fw = {(u, int(v)) for u in range(300) for v in g.fw_adj[g.fw_off[u]:g.fw_off[u + 1]]}
    bw = {(int(u), v) for v in range(300) for u in g.bw_adj[g.bw_off[v]:g.bw_off[v + 1]]}
    assert fw == bw
    for v in range(300):
        row = g.bw_adj[g.bw_off[v]:g.bw_off[v + 1]]
        assert np.all(row[1:] > row[:-1])  # ascending sources
    # a planted mistake (rows swapped) must be caught by the multiset check
    bad = g.bw_adj.copy()
    if len(bad) > 1:
        bad[[0, -1]] = bad[[-1, 0]]
        bw_bad = {(int(u), v) for v in range(300) for u in bad[g.bw_off[v]:g.bw_off[v + 1]]}
        assert bw_bad != fw


# ---------------------------------------------------------------- closed forms
def test_closed_forms(O):
    rng = np.random.default_rng(3)
    # directed cycle: one SCC
    off, adj = H.cycle(50)
    g = O.HostGraph(50, off, adj)
    assert set(O.tarjan(g).tolist()) == {0}
    assert set(O.brute_force(g).tolist()) == {0}
    # forward DAG: singletons
    off, adj = H.forward_dag(60, 300, rng)
    g = O.HostGraph(60, off, adj)
    assert O.tarjan(g).tolist() == list(range(60))
    assert O.brute_force(g).tolist() == list(range(60))
    # empty graph and n = 1
    for n in (1, 5):
        g = O.HostGraph(n, np.zeros(n + 1, np.uint64), np.zeros(0, np.uint32))
        assert O.tarjan(g).tolist() == list(range(n))
        for fn in (O.kaspan_scc, O.scc_trim_ex, O.scc_trim_ex_residual):
            assert O.canonicalize(fn(g)[0]).tolist() == list(range(n))
    # fully bidirectional lattice: one SCC
    s, d = O.gen_lattice(1, 5, 1.0)
    g = O.csr_from_edges(125, s, d)
    assert set(O.tarjan(g).tolist()) == {0}


def test_canonicalize(O):
    assert O.canonicalize([3, 3, 1, 1, 3]).tolist() == [0, 0, 2, 2, 0]
    assert O.canonicalize([4, 4, 4, 4, 4]).tolist() == [0, 0, 0, 0, 0]
    assert O.canonicalize([0, 1, 2]).tolist() == [0, 1, 2]


@pytest.mark.parametrize("seed", range(12))
def test_tarjan_vs_brute_force(O, seed):
    n = 40 + 20 * seed
    s, d = O.gen_gnm(seed, n, int(n * (0.8 + 0.2 * seed)))
    g = O.csr_from_edges(n, s, d)
    assert np.array_equal(O.tarjan(g), O.brute_force(g))


# ---------------------------------------------------------------- planted labels
@pytest.mark.parametrize("n", [1, 6, 8, 16, 33, 64])
def test_fuzzy_grid_all_pipelines(O, n):
    """test_mpi/test_scc_fuzzy.cpp:31-80 grid with fixed seeds"""
    for run in range(40):
        g, planted = O.gen_fuzzy(1000 * n + run, n)
        assert np.array_equal(O.tarjan(g), planted)
        for fn in (O.kaspan_scc, O.scc_trim_ex, O.scc_trim_ex_residual):
            lab, _ = fn(g)
            assert np.array_equal(O.canonicalize(lab), planted), (fn.__name__, run)


@pytest.mark.parametrize("d", [0.5, 1.0, 4.0])
def test_fuzzy_generator_vs_tarjan(O, d):
    """test/test_tarjan_fuzzy.cpp:17-40 and test_fuzzy_generator.cpp:14-76"""
    for n in range(0, 300, 23):
        for i in range(3):
            g, planted = O.gen_fuzzy(n * 31 + i, n, d)
            t = O.tarjan(g)
            assert np.array_equal(t, planted)
            assert int(np.sum(planted == np.arange(n))) == len(np.unique(t))
            # the bw view has the same components
            gb = O.HostGraph(n, g.bw_off, g.bw_adj, g.fw_off, g.fw_adj)
            assert np.array_equal(O.tarjan(gb), planted)
            if n > 1:
                assert g.m / n >= min(d, (n - 1) / 2.0) - 1e-9 or g.m >= n * (n - 1) // 2


def test_planted_power_law(O):
    n = 20000
    g, planted = O.gen_planted(42, n)
    assert np.array_equal(O.tarjan(g), planted)
    sizes = O.planted_sizes(42, n)
    assert int(sizes.sum()) == n
    assert len(np.unique(planted)) == len(sizes)
    assert 0.4 < np.mean(sizes == 1) < 0.8


# ---------------------------------------------------------------- stages
def test_trim_1_first_closed_form(O):
    s, d = O.gen_rmat(9, 10, 3000)
    g = O.csr_from_edges(1024, s, d)
    pivot, live, scc_id = O.trim_1_first(g)
    outd = np.diff(g.fw_off.astype(np.int64))
    ind = np.diff(g.bw_off.astype(np.int64))
    dead = (outd == 0) | (ind == 0)
    assert np.array_equal(live == 0, dead)
    assert np.array_equal(scc_id[dead], np.arange(1024)[dead])
    prod = np.where(dead, -1, ind * outd)
    assert pivot == int(np.argmax(prod))  # first maximum = smallest id


@pytest.mark.parametrize("seed", range(6))
def test_fwbw_decides_pivot_component(O, seed):
    s, d = O.gen_gnm(seed, 500, 900 + 100 * seed)
    g = O.csr_from_edges(500, s, d)
    pivot, live, _ = O.trim_1_first(g)
    if pivot >= 500:
        return
    n_dec, live2, scc_id = O.fwbw(g, live, pivot)
    t = O.tarjan(g)
    comp = np.flatnonzero(t == t[pivot])
    decided = np.flatnonzero((live == 1) & (live2 == 0))
    assert np.array_equal(np.sort(decided), comp)
    assert n_dec == len(comp) and np.all(scc_id[comp] == pivot)


@pytest.mark.parametrize("seed", range(6))
def test_color_round_invariants(O, seed):
    n = 120
    s, d = O.gen_gnm(100 + seed, n, 170)
    g = O.csr_from_edges(n, s, d)
    live = np.ones(n, np.uint8)
    r = 1 + seed
    n_dec, live2, color, scc_id = O.color_round(g, live, r)
    t = O.tarjan(g)
    # colour = min rotated id over ancestors (all vertices live)
    reach = np.zeros((n, n), bool)  # reach[u, v]: u reaches v
    for u in range(n):
        stack, seen = [u], {u}
        while stack:
            x = stack.pop()
            for y in g.fw_adj[g.fw_off[x]:g.fw_off[x + 1]]:
                if int(y) not in seen:
                    seen.add(int(y))
                    stack.append(int(y))
        reach[u, list(seen)] = True
    rot = np.array([((v >> r) | (v << (32 - r))) & 0xFFFFFFFF for v in range(n)], np.uint64)
    for v in range(n):
        assert color[v] == rot[reach[:, v]].min()
    # decided vertices form whole components labelled by a member
    dec = np.flatnonzero(live2 == 0)
    assert n_dec == len(dec) > 0
    for v in dec:
        assert t[scc_id[v]] == t[v]
        assert np.all(live2[t == t[v]] == 0)


def test_pipeline_stats_consistent(O):
    s, d = O.gen_gnm(0x6B5350AA00000000, 65536, 524288)
    g = O.csr_from_edges(65536, s, d)
    lab, st = O.kaspan_scc(g)
    assert np.array_equal(O.canonicalize(lab), O.tarjan(g))
    total = st.decided_trim_first + st.decided_fwbw + st.decided_trim_normal + st.decided_color
    assert total == 65536
    lab2, _ = O.kaspan_scc(g, faithful=True)
    assert np.array_equal(lab, lab2)


# ---------------------------------------------------------------- generators
def test_generator_fixtures(O):
    from golden.make_generator_fixtures import build
    for case in json.load(open(os.path.join(GOLDEN, "generators.json"))):
        n, s, d, g = build(case)
        assert n == case["n_vertices"] and len(s) == case["raw_edges"] and g.m == case["m"]
        assert [[int(a), int(b)] for a, b in zip(s[:8], d[:8])] == case["first_edges"]
        assert H.csr_digest(g.fw_off, g.fw_adj) == case["fw_digest"]
        assert H.csr_digest(g.bw_off, g.bw_adj) == case["bw_digest"]
        assert len(np.unique(O.tarjan(g))) == case["components"]


def test_generator_shapes(O):
    s, d = O.gen_lattice(3, 6, 0.1)
    assert len(s) >= 3 * 36 * 5 and np.all(s != d)
    diff = np.abs(s.astype(np.int64) - d.astype(np.int64))
    assert set(np.unique(diff).tolist()) <= {1, 6, 36}
    s, d = O.gen_rmat(4, 12, 1 << 16)
    deg = np.bincount(s, minlength=4096)
    assert deg.max() > 20 * deg.mean()  # heavy-tailed
    s, d = O.gen_gnm(4, 1000, 5000)
    assert s.max() < 1000 and d.max() < 1000 and np.all(s != d)


def test_rmat_n_reduces_to_rmat_on_powers_of_two(O):
    # for n = 2^s nothing folds and the cycle-walking bijection is perm_pow2:
    # the batch's R-MAT stream is then exactly the pinned scale-s stream
    for seed, s in [(3, 6), (9, 10), (21, 13)]:
        a, b = O.gen_rmat_n(seed, 1 << s, 5000), O.gen_rmat(seed, s, 5000, permute=True)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    # any n: ids stay in range, no self-loops, the degree stays heavy-tailed
    s, d = O.gen_rmat_n(5, 3000, 60000)
    assert s.max() < 3000 and d.max() < 3000 and np.all(s != d)
    deg = np.bincount(s, minlength=3000)
    assert deg.max() > 10 * deg.mean()
    assert O.gen_rmat_n(5, 1, 10)[0].size == 0  # one vertex: every draw is a self-loop


def test_batch_plan_distribution(O):
    kind, n, draws, seeds = O.batch_plan(0x6B5350AA00000004, 4096)
    assert n.min() >= 1000 and n.max() <= 1_000_000
    deg = draws / n
    assert deg.min() >= 2.0 - 1e-3 and deg.max() <= 16.0
    # log-uniform: log n has mean (ln 1e3 + ln 1e6) / 2 and the kinds split evenly
    assert abs(np.log(n.astype(np.float64)).mean() - np.log(1e3 * 1e6) / 2) < 0.15
    assert 0.45 < kind.mean() < 0.55
    assert len(np.unique(seeds)) == 4096
    # the whole plan of the 4096-graph workload: about 5.9e8 vertices, 4.0e9 draws
    assert 5.0e8 < int(n.sum()) < 7.0e8 and 3.4e9 < int(draws.sum()) < 4.8e9


def test_batch_union_labels(O):
    u, vo, eo, graphs = O.gen_batch(11, 20, 40, 2000, 1.0, 6.0)
    assert u.n == int(vo[-1]) and u.m == int(eo[-1])
    lab = O.tarjan(u)
    for i, g in enumerate(graphs):
        local = O.tarjan(g)
        # the union's components never cross graphs, so its min ids are local ids + the graph's base
        assert np.array_equal(lab[vo[i]: vo[i + 1]] - vo[i], local)
        # rows point inside the graph's own ranges
        assert int(u.fw_off[vo[i]]) == int(eo[i])
        seg = u.fw_adj[eo[i]: eo[i + 1]]
        assert seg.size == 0 or (seg.min() >= vo[i] and seg.max() < vo[i + 1])


# ---------------------------------------------------------------- MLP
def test_mlp_gradients_finite_differences(O):
    rng = np.random.default_rng(0)
    x = rng.normal(size=(17, 6)).astype(np.float32)
    t = rng.normal(size=17).astype(np.float32)
    p = [rng.normal(scale=0.4, size=k) for k in (192, 32, 1024, 32, 32)]
    b3 = 0.1
    loss, grads = O.mlp_loss_grad(x, t, *p, b3)
    eps = 1e-6
    for pi, gi in zip(range(5), grads[:5]):
        for j in rng.choice(len(p[pi]), 6, replace=False):
            q = [a.copy() for a in p]
            q[pi][j] += eps
            lp, _ = O.mlp_loss_grad(x, t, *q, b3)
            q[pi][j] -= 2 * eps
            lm, _ = O.mlp_loss_grad(x, t, *q, b3)
            assert abs((lp - lm) / (2 * eps) - gi[j]) < 1e-5 * max(1.0, abs(gi[j]))
    lp, _ = O.mlp_loss_grad(x, t, *p, b3 + eps)
    lm, _ = O.mlp_loss_grad(x, t, *p, b3 - eps)
    assert abs((lp - lm) / (2 * eps) - grads[5]) < 1e-6
    out = O.mlp_forward(x, *p, b3)
    assert abs(0.5 * np.mean((out - t) ** 2) - loss) < 1e-9


def np_features(n, src, dst):
    """The six pivot features from an edge list (duplicates and self-loops
    kept, as rows of the CSR hold them), written as degree counts and
    sparse matrix-vector products instead of the oracle's per-vertex loops:
    log1p(in), log1p(out), log1p(in)*log1p(out)/8, log1p(A @ out),
    log1p(A^T @ in), log1p(min(in, out)) with A[u, v] = #edges u -> v."""
    src = np.asarray(src, np.int64)
    dst = np.asarray(dst, np.int64)
    outd = np.bincount(src, minlength=n).astype(np.float64)
    ind = np.bincount(dst, minlength=n).astype(np.float64)
    s2o = np.bincount(src, weights=outd[dst], minlength=n)  # (A @ out)[u] = sum over u -> w of out(w)
    s2i = np.bincount(dst, weights=ind[src], minlength=n)   # (A^T @ in)[v] = sum over w -> v of in(w)
    li, lo = np.log1p(ind), np.log1p(outd)
    return np.stack([li, lo, li * lo / 8, np.log1p(s2o), np.log1p(s2i), np.log1p(np.minimum(ind, outd))], axis=1)


def csr_with_multiedges(n, src, dst):
    """CSR + reverse CSR that keep every listed edge (duplicates, self-loops)"""
    fo, fa = H.csr_from_pairs(n, list(zip(src.tolist(), dst.tolist())))
    bo, ba = H.csr_from_pairs(n, list(zip(dst.tolist(), src.tolist())))
    return fo, fa, bo, ba


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_pivot_features_against_numpy(O, seed):
    """orc_pivot_features == a NumPy restatement on graphs with self-loops and
    duplicate edges, to float32 rounding (the oracle sums in integers and
    takes log1pf in float32)"""
    rng = np.random.default_rng(seed)
    n = 300
    src = rng.integers(0, n, 2000)
    dst = rng.integers(0, n, 2000)
    src = np.concatenate([src, src[:50], np.arange(20)])  # 50 duplicates, 20 self-loops
    dst = np.concatenate([dst, dst[:50], np.arange(20)])
    fo, fa, bo, ba = csr_with_multiedges(n, src, dst)
    got = O.pivot_features(O.HostGraph(n, fo, fa, bo, ba)).astype(np.float64)
    want = np_features(n, src, dst)
    assert np.allclose(got, want, rtol=2e-6, atol=1e-6)
    # the two 2-hop sums must not be interchangeable: swapping directions is caught
    swapped = np_features(n, dst, src)
    assert not np.allclose(got[:, 3:5], swapped[:, 3:5], rtol=1e-3)


def test_pivot_features_closed_forms(O):
    """out-star 0 -> 1..k and a 2-cycle 10 <-> 11: every feature by hand"""
    k = 5
    src = np.array([0] * k + [10, 11])
    dst = np.array(list(range(1, k + 1)) + [11, 10])
    fo, fa, bo, ba = csr_with_multiedges(12, src, dst)
    f = O.pivot_features(O.HostGraph(12, fo, fa, bo, ba)).astype(np.float64)
    l = np.log1p
    assert np.allclose(f[0], [0, l(k), 0, 0, 0, 0], atol=1e-6)              # leaves have no out-edges
    assert np.allclose(f[1], [l(1), 0, 0, 0, l(0), 0], atol=1e-6)           # the centre has in-degree 0
    assert np.allclose(f[10], [l(1), l(1), l(1) ** 2 / 8, l(1), l(1), l(1)], atol=1e-6)
    assert np.allclose(f[6], 0, atol=0)                                       # isolated vertex


def test_mlp_forward_against_numpy(O):
    """orc_mlp_forward == relu(relu(x W1^T + b1) W2^T + b2) w3 + b3 in float64 matrix form"""
    rng = np.random.default_rng(3)
    x = rng.normal(size=(257, 6)).astype(np.float32)
    w1, b1, w2, b2, w3 = (rng.normal(scale=0.5, size=k) for k in (192, 32, 1024, 32, 32))
    b3 = -0.3
    got = O.mlp_forward(x, w1, b1, w2, b2, w3, b3)
    h1 = np.maximum(x.astype(np.float64) @ w1.reshape(32, 6).T + b1, 0)
    h2 = np.maximum(h1 @ w2.reshape(32, 32).T + b2, 0)
    want = h2 @ w3 + b3
    assert np.allclose(got, want, rtol=1e-12, atol=1e-12)
    # a transposed W2 is caught
    h2t = np.maximum(h1 @ w2.reshape(32, 32) + b2, 0)
    assert not np.allclose(got, h2t @ w3 + b3, rtol=1e-6)


# ---------------------------------------------------------------- full-size helpers
@pytest.mark.parametrize("seed,n,max_size", [(5, 1000, 100), (6, 20000, 5000), (0x6B5350AA00000002, 50000, 10000)])
def test_planted_labels_without_edges(O, seed, n, max_size):
    """planted_labels (no edges) == gen_planted's labels == Tarjan on its graph"""
    g, lab = O.gen_planted(seed, n, max_size=max_size)
    only = O.planted_labels(seed, n, max_size=max_size)
    assert np.array_equal(only, lab)
    assert np.array_equal(only, O.tarjan(g))


def test_tarjan_batch_equals_tarjan_per_graph(O):
    u, vo, eo, graphs = O.gen_batch(11, 40, n_min=20, n_max=3000, deg_min=1.0, deg_max=8.0)
    got = O.tarjan_batch(vo, u.fw_off, u.fw_adj, threads=4)
    want = np.concatenate([O.tarjan(g) for g in graphs])
    assert np.array_equal(got, want)
    lab_u = O.tarjan(u)
    for i in range(40):
        assert np.array_equal(got[vo[i]: vo[i + 1]], lab_u[vo[i]: vo[i + 1]] - vo[i])
