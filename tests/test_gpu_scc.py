"""GPU parity: libkspn through its C ABI against the CPU oracle and the
reference's golden vectors. Integer work, so every comparison is bit-exact."""
import ctypes as C
import json
import os

import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
C1_SEED = 0x6B5350AA00000000
C2_SEED = 0x6B5350AA00000001

METHODS = [
    dict(method=0),
    dict(method=1),
    dict(method=1, flags=2),
    dict(method=1, fwbw_rounds=4),
    dict(method=1, residual_threshold=300),
    dict(method=0, residual_threshold=0xFFFFFFFF),
    dict(method=0, alpha=1e9, beta=1.0),   # pull levels as soon as possible
    dict(method=0, alpha=1e-9),            # push levels only
    dict(method=0, flags=4),               # colour rounds only, no tail solver
    dict(method=1, flags=4),
    dict(method=0, tail_threshold=1 << 30),  # tail solver takes everything after FW-BW
    dict(method=0, fwbw_rounds=64, flags=4),   # partitioned multi-pivot FW-BW, then colour rounds
    dict(method=0, fwbw_rounds=3, alpha=1e9, beta=1.0),  # partitioned, pull levels
    dict(method=0, flags=4 | 8),               # the reference's colour rounds: no tail, no refinement
    dict(method=0, flags=16),                  # every FW/BW level as separate launches
    dict(method=0, fwbw_rounds=8, flags=16),   # partitioned rounds, separate launches
    dict(method=0, flags=1024),                # device push levels in one persistent launch
    dict(method=0, fwbw_rounds=8, flags=1024),
]
MIDS = ["kaspan", "trim_ex", "trim_ex_no_trim2", "fwbw4", "residual300", "kaspan_residual_all", "pull", "push",
        "kaspan_no_tail", "trim_ex_no_tail", "tail_all", "multi_pivot64", "multi_pivot3_pull", "kaspan_plain_colour",
        "no_coop", "multi_pivot8_no_coop", "persistent", "multi_pivot8_persistent"]


def host_graph(O, g):
    fo, fa, bo, ba = g.to_host()
    return O.HostGraph(g.n, fo, fa, bo, ba)


def gpu_labels(ka, g, **opts):
    labels, st = ka.scc(g, **opts)
    return labels.cpu().numpy().view(np.uint32), st


def test_ispan_tiny(ka, O):
    g = ka.BidiGraph.from_host(7, H.ISPAN_TINY_FW_OFF, H.ISPAN_TINY_FW_ADJ, H.ISPAN_TINY_BW_OFF, H.ISPAN_TINY_BW_ADJ)
    for opts in METHODS:
        lab, st = gpu_labels(ka, g, **opts)
        assert lab.tolist() == H.ISPAN_TINY_LABELS, opts


def test_transpose_golden(ka):
    bo, ba = ka.backward_complement_graph(6, np.array(H.SUB_FW_OFF, np.uint64), np.array(H.SUB_FW_ADJ, np.uint32))
    assert bo.cpu().tolist() == H.SUB_BW_OFF
    assert ba.cpu().tolist() == H.SUB_BW_ADJ


def test_subgraph_golden(ka):
    import torch

    mask = torch.tensor(H.SUB_MASK, dtype=torch.uint8)
    sid, so, sa = ka.allgather_fw_sub_graph(6, np.array(H.SUB_FW_OFF, np.uint64), np.array(H.SUB_FW_ADJ, np.uint32), mask)
    assert sid.cpu().tolist() == H.SUB_SUPER_IDS
    assert so.cpu().tolist() == H.SUB_SUB_FW_OFF and sa.cpu().tolist() == H.SUB_SUB_FW_ADJ
    sid, so, sa = ka.allgather_fw_sub_graph(6, np.array(H.SUB_BW_OFF, np.uint64), np.array(H.SUB_BW_ADJ, np.uint32), mask)
    assert so.cpu().tolist() == H.SUB_SUB_BW_OFF and sa.cpu().tolist() == H.SUB_SUB_BW_ADJ


def test_subgraph_random_vs_oracle(ka, O):
    import torch

    s, d = O.gen_gnm(21, 3000, 12000)
    hg = O.csr_from_edges(3000, s, d)
    mask = (np.random.default_rng(1).random(3000) < 0.3).astype(np.uint8)
    sid, so, sa = ka.allgather_fw_sub_graph(3000, hg.fw_off, hg.fw_adj, torch.from_numpy(mask))
    esid, eso, esa = O.subgraph(hg, mask)
    assert np.array_equal(sid.cpu().numpy().view(np.uint32), esid)
    assert np.array_equal(so.cpu().numpy().view(np.uint64), eso)
    assert np.array_equal(sa.cpu().numpy().view(np.uint32), esa)


def test_generators_match_oracle(ka, O):
    for case in json.load(open(os.path.join(GOLDEN, "generators.json"))):
        kw = {k: case[k] for k in ("n", "draws", "scale", "side", "p_bidir", "permute") if k in case}
        g = ka.generate(case["kind"], case["seed"], **kw)
        assert g.n == case["n_vertices"] and g.m == case["m"], case["kind"]
        fo, fa, bo, ba = g.to_host()
        assert H.csr_digest(fo, fa) == case["fw_digest"], case
        assert H.csr_digest(bo, ba) == case["bw_digest"], case


def test_build_csr_from_edges(ka, O):
    rng = np.random.default_rng(5)
    src = rng.integers(0, 500, 4000).astype(np.uint32)
    dst = rng.integers(0, 500, 4000).astype(np.uint32)
    g = ka.BidiGraph.from_edges(500, src, dst)
    hg = O.csr_from_edges(500, src, dst)
    fo, fa, bo, ba = g.to_host()
    assert np.array_equal(fo, hg.fw_off) and np.array_equal(fa, hg.fw_adj)
    assert np.array_equal(bo, hg.bw_off) and np.array_equal(ba, hg.bw_adj)


def test_tarjan_kernel(ka, O):
    for seed, n, draws in [(1, 7, 0), (2, 2000, 2500), (3, 5000, 9000)]:
        if draws == 0:
            hg = O.HostGraph(7, H.ISPAN_TINY_FW_OFF, H.ISPAN_TINY_FW_ADJ)
        else:
            s, d = O.gen_gnm(seed, n, draws)
            hg = O.csr_from_edges(n, s, d)
        lab = ka.tarjan(hg.n, hg.fw_off, hg.fw_adj)
        assert np.array_equal(lab.cpu().numpy().view(np.uint32), O.tarjan(hg))


def test_canonicalize_kernel(ka, O):
    import torch

    rng = np.random.default_rng(9)
    for n, k in [(1, 1), (100, 7), (100000, 300), (300000, 3)]:
        reps = rng.choice(n, size=k, replace=False).astype(np.uint32)
        lab = reps[rng.integers(0, k, n)]
        t = torch.from_numpy(lab.view(np.int32)).cuda()
        ka.normalise_scc_id(t)
        assert np.array_equal(t.cpu().numpy().view(np.uint32), O.canonicalize(lab))


@pytest.mark.parametrize("n", [1, 6, 8, 16, 33, 64, 300])
def test_fuzzy_grid(ka, O, n):
    """test_mpi/test_scc_fuzzy.cpp:31-80 grid (fixed seeds) for every method"""
    for run in range(40 if n < 300 else 8):
        hg, planted = O.gen_fuzzy(1000 * n + run, n)
        g = ka.BidiGraph.from_host(n, hg.fw_off, hg.fw_adj, hg.bw_off, hg.bw_adj)
        for opts in METHODS:
            lab, st = gpu_labels(ka, g, **opts)
            assert np.array_equal(lab, planted), (n, run, opts)


@pytest.mark.parametrize("opts", METHODS, ids=MIDS)
@pytest.mark.parametrize("kind", ["gnm_sparse", "gnm_dense", "rmat14", "rmat16", "lattice24", "planted"])
def test_random_graphs(ka, O, kind, opts):
    if kind == "gnm_sparse":
        g = ka.generate("gnm", 11, n=30000, draws=33000)
    elif kind == "gnm_dense":
        g = ka.generate("gnm", 12, n=20000, draws=160000)
    elif kind == "rmat14":
        g = ka.generate("rmat", 13, scale=14, draws=16 << 14)
    elif kind == "rmat16":
        g = ka.generate("rmat", 14, scale=16, draws=16 << 16)
    elif kind == "lattice24":
        g = ka.generate("lattice", 15, side=24, p_bidir=0.1)
    else:
        hg0, _ = O.gen_planted(16, 40000)
        g = ka.BidiGraph.from_host(hg0.n, hg0.fw_off, hg0.fw_adj, hg0.bw_off, hg0.bw_adj)
    hg = host_graph(O, g)
    lab, st = gpu_labels(ka, g, **opts)
    assert np.array_equal(lab, O.tarjan(hg))
    total = st.decided_trim + st.decided_trim2 + st.decided_fwbw + st.decided_color + st.decided_residual
    assert total == g.n


def test_persistent_levels(ka, O):
    """long-diameter lattice: the device runs the push levels, as chained
    launches (default) or in one persistent launch, and expands its own claims
    for a few rounds per level, so it needs fewer levels; the same labels and
    the same FW-BW decisions as separate launches per level"""
    g = ka.generate("lattice", 31, side=40, p_bidir=0.05)
    ref = O.tarjan(host_graph(O, g))
    for rounds in (1, 16):
        lab2, st2 = gpu_labels(ka, g, fwbw_rounds=rounds, flags=16)
        assert np.array_equal(lab2, ref) and st2.fwbw_levels_coop == 0
        for flags in (0, 1024):
            lab, st = gpu_labels(ka, g, fwbw_rounds=rounds, flags=flags)
            assert np.array_equal(lab, ref)
            assert st.fwbw_levels_coop > 0
            assert st.fwbw_levels_coop + st.fwbw_levels_push <= st2.fwbw_levels_push
            assert st.decided_fwbw == st2.decided_fwbw


def test_c1_gnm_full(ka, O):
    """config 1: G(n=65536, m=524288), degree-product pivot, oracle = Tarjan"""
    g = ka.generate("gnm", C1_SEED, n=65536, draws=524288)
    hg = host_graph(O, g)
    ref_raw, ref_st = O.kaspan_scc(hg)
    tar = O.tarjan(hg)
    assert np.array_equal(O.canonicalize(ref_raw), tar)
    lab, st = gpu_labels(ka, g, method=0)
    assert np.array_equal(lab, tar)
    # stage parity with the restated reference: same pivot, same first trim, same FW-BW set
    assert st.pivot == ref_st.pivot
    assert st.decided_fwbw == ref_st.decided_fwbw
    lab2, _ = gpu_labels(ka, g, method=1)
    assert np.array_equal(lab2, tar)


def test_stage_parity_kaspan(ka, O):
    for seed in range(4):
        g = ka.generate("rmat", 100 + seed, scale=13, draws=8 << 13)
        hg = host_graph(O, g)
        _, ref_st = O.kaspan_scc(hg)
        lab, st = gpu_labels(ka, g, method=0)
        assert st.pivot == ref_st.pivot
        assert st.decided_fwbw == ref_st.decided_fwbw
        assert np.array_equal(lab, O.tarjan(hg))


def test_edge_cases(ka, O):
    import torch

    # n = 0
    g = ka.BidiGraph.from_host(0, [0], [], [0], [])
    lab, st = ka.scc(g)
    assert lab.numel() == 0
    # isolated vertices
    g = ka.BidiGraph.from_host(5, [0] * 6, [], [0] * 6, [])
    assert gpu_labels(ka, g)[0].tolist() == [0, 1, 2, 3, 4]
    # self-loops only: singletons
    off = np.arange(6, dtype=np.uint64)
    g = ka.BidiGraph.from_host(5, off, np.arange(5, dtype=np.uint32), off, np.arange(5, dtype=np.uint32))
    for opts in METHODS:
        assert gpu_labels(ka, g, **opts)[0].tolist() == [0, 1, 2, 3, 4]
    # a long directed cycle and a long path
    for n in (2, 1000):
        off, adj = H.cycle(n)
        g = ka.BidiGraph.from_host(n, off, adj)
        for opts in METHODS:
            assert set(gpu_labels(ka, g, **opts)[0].tolist()) == {0}
    off, adj = H.csr_from_pairs(500, [(i, i + 1) for i in range(499)])
    g = ka.BidiGraph.from_host(500, off, adj)
    for opts in METHODS:
        assert gpu_labels(ka, g, **opts)[0].tolist() == list(range(500))
    # bad arguments
    with pytest.raises(ka.KspnError):
        ka.scc(g, torch.empty(3, dtype=torch.int32, device="cuda"))


def test_iteration_cap(ka, O):
    # a forward DAG with every vertex live after trim-1's one pass needs many colour rounds
    hg0, _ = O.gen_planted(77, 5000)
    g = ka.BidiGraph.from_host(hg0.n, hg0.fw_off, hg0.fw_adj, hg0.bw_off, hg0.bw_adj)
    for flags in (0, 4, 12):  # cap inside the tail solver and in the host colour loop
        with pytest.raises(ka.KspnError) as ei:
            ka.scc(g, method=0, max_rounds=1, flags=flags)
        assert ei.value.code == -4
    lab, _ = gpu_labels(ka, g, method=0, max_rounds=0)
    assert np.array_equal(lab, O.tarjan(hg0))


def test_profile_stats(ka, O):
    g = ka.generate("rmat", 3, scale=15, draws=16 << 15)
    lab, st = ka.scc(g, profile=True)
    d = st.as_dict()
    names = {k["name"] for k in d["kernels"]}
    assert "trim_first" in names and "label_scc" in names
    assert all(k["ms"] >= 0 and k["launches"] > 0 for k in d["kernels"])
    assert st.ms_total > 0


@pytest.mark.parametrize("offset", [0, 1, 2, 3])
def test_labels_any_alignment(ka, O, offset):
    """The dense label pass after the seeded first FW-BW stores four labels
    per 16-byte write when the labels start on a 16-byte boundary, one label
    per write otherwise: a caller's buffer at every 4-byte offset, and n not
    a multiple of 4, give Tarjan's labels; the words around it stay as they
    were."""
    s, d = O.gen_rmat(61, 14, 16 << 14)
    n = (1 << 14) - 3
    keep = (s < n) & (d < n)
    hg = O.csr_from_edges(n, s[keep], d[keep])
    g = ka.BidiGraph.from_host(hg.n, hg.fw_off, hg.fw_adj, hg.bw_off, hg.bw_adj)
    import torch

    buf = torch.full((n + 8,), -7, dtype=torch.int32, device="cuda")
    assert buf.data_ptr() % 16 == 0
    view = buf[offset:offset + n]
    lab, _ = ka.scc(g, view)
    assert np.array_equal(lab.cpu().numpy().view(np.uint32), O.tarjan(hg))
    rest = buf.cpu().numpy()
    assert (rest[:offset] == -7).all() and (rest[offset + n:] == -7).all()


@pytest.mark.parametrize("kind", ["rmat20", "gnm20", "rmat20_cycle"])
def test_speculative_sweep(ka, O, kind):
    """On graphs of at least 2^20 vertices the sweep after a direction's
    first push level is launched before the host has decided (R-MAT: the
    device's rule chooses pull; sparse G(n,m): it does not, and the gated
    launches and their publication must leave no trace; R-MAT with a
    24-vertex cycle through its largest hub: a list batch that does not end
    the reach, one cycle vertex per level, then push levels again). Labels
    equal Tarjan's, and the level counts equal a profiled solve's, which
    never speculates; three solves in a row on one context."""
    if kind == "gnm20":
        g = ka.generate("gnm", 92, n=1 << 20, draws=3 << 20)
    else:
        g = ka.generate("rmat", 91, scale=20, draws=16 << 20)
    if kind == "rmat20_cycle":
        fo, fa, _, _ = g.to_host()
        deg = np.diff(fo.astype(np.int64))
        hub = int(np.argmax(deg))
        n0 = g.n
        src = np.repeat(np.arange(n0, dtype=np.int64), deg)
        dst = fa.astype(np.int64)
        cyc = np.arange(n0, n0 + 24, dtype=np.int64)
        src = np.concatenate([src, [hub], cyc[:-1], [cyc[-1]]])
        dst = np.concatenate([dst, [cyc[0]], cyc[1:], [hub]])
        g = ka.BidiGraph.from_edges(n0 + 24, src, dst)
    tar = O.tarjan(host_graph(O, g))
    ref, st_ref = ka.scc(g, profile=True)
    assert np.array_equal(ref.cpu().numpy().view(np.uint32), tar)
    for _ in range(3):
        lab, st = ka.scc(g)
        assert np.array_equal(lab.cpu().numpy().view(np.uint32), tar)
        assert (st.fwbw_levels_push, st.fwbw_levels_pull, st.fwbw_levels_coop) == \
            (st_ref.fwbw_levels_push, st_ref.fwbw_levels_pull, st_ref.fwbw_levels_coop)
    if kind == "rmat20":
        assert st.fwbw_levels_pull > 0


@pytest.mark.parametrize("shape", [(16, 4096), (32, 32768)], ids=["l16x4096", "l32x32768"])
@pytest.mark.parametrize("orient", ["down", "up"])
def test_repeated_sweeps(ka, O, shape, orient):
    """Layered graphs whose reach advances about one layer per pull sweep, so
    a sweep leaves more than n/16 candidates, lists none, and the host sweeps
    again (at 2^20 vertices after a speculative first sweep). Labels equal
    Tarjan's in a profiled solve and three without statistics; the profiled solve
    sweeps at least three times per direction. Level counts are not compared:
    claims made earlier in the same sweep are seen or not by timing."""
    layers, width = shape
    n, src, dst = H.layered(7, layers, width)
    if orient == "up":
        src, dst = dst, src
    g = ka.BidiGraph.from_edges(n, src, dst)
    tar = O.tarjan(host_graph(O, g))
    lab, st = ka.scc(g, profile=True)
    assert np.array_equal(lab.cpu().numpy().view(np.uint32), tar)
    sweeps = {k["name"]: k["launches"] for k in st.as_dict()["kernels"]}
    assert sweeps.get("pull_fw", 0) >= 3 and sweeps.get("pull_bw", 0) >= 3, sweeps
    for _ in range(3):
        lab, _ = ka.scc(g, stats=False)  # the list levels then sum no edges
        assert np.array_equal(lab.cpu().numpy().view(np.uint32), tar)


@pytest.mark.parametrize("dag_edges", [500, 8000], ids=["lds", "global"])
def test_tail_solver_sizes(ka, O, dag_edges):
    """A dense 5000-vertex SCC takes the pivot; what FW-BW leaves is 1000
    three-cycles joined by forward DAG edges between cycles, all for the tail
    solver. With 500 DAG edges its sub-CSR (3500 edges per direction) runs
    from LDS, with 8000 (11 000) from the global buffers. Labels equal
    Tarjan's and the tail took all 3000 cycle vertices."""
    rng = np.random.default_rng(17)
    big, cyc = 5000, 1000
    src = [np.arange(big), rng.integers(0, big, 20 * big)]
    dst = [(np.arange(big) + 1) % big, rng.integers(0, big, 20 * big)]
    base = big + 3 * np.arange(cyc)
    for k in range(3):
        src.append(base + k)
        dst.append(base + (k + 1) % 3)
    a = rng.integers(0, cyc, 2 * dag_edges)
    b = rng.integers(0, cyc, 2 * dag_edges)
    keep = a < b
    a, b = a[keep][:dag_edges], b[keep][:dag_edges]
    src.append(big + 3 * a + rng.integers(0, 3, a.size))
    dst.append(big + 3 * b + rng.integers(0, 3, b.size))
    n = big + 3 * cyc
    g = ka.BidiGraph.from_edges(n, np.concatenate(src).astype(np.int64), np.concatenate(dst).astype(np.int64))
    tar = O.tarjan(host_graph(O, g))
    lab, st = gpu_labels(ka, g)
    assert np.array_equal(lab, tar)
    assert st.residual_n == 3 * cyc


def test_c2_rmat24_full(ka, O):
    """config 2 at full size: R-MAT scale 24, edge factor 16, permuted ids;
    labels equal the oracle's Tarjan labels bit for bit."""
    g = ka.generate("rmat", C2_SEED, scale=24, draws=16 << 24)
    lab, st = gpu_labels(ka, g, method=0)
    hg = host_graph(O, g)
    tar = O.tarjan(hg)
    assert np.array_equal(lab, tar)
    lab2, _ = gpu_labels(ka, g, method=1)
    assert np.array_equal(lab2, tar)


def test_cpp_mirror(ka, tmp_path):
    """include/kspn/kaspan.hpp from C++: golden labels under every variant, a
    manifest loaded through kaspan::manifest::load, error -> exception"""
    import subprocess

    from kaspan_amd import manifest as M

    mf = M.write_manifest(str(tmp_path / "tiny"), 7, H.ISPAN_TINY_FW_OFF, H.ISPAN_TINY_FW_ADJ, H.ISPAN_TINY_BW_OFF,
                          H.ISPAN_TINY_BW_ADJ, endian="big", head_bytes=3, csr_bytes=2)
    exe = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "test_scc_cpp")
    out = subprocess.run([exe, mf], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "cpp ok" in out.stdout


def test_tail_guard_reports_bad_ids(ka, O):
    """The tail kernels check every listed id against n: an id >= n is
    reported and skipped, never used as an index. The graph's arrays hold all
    ids below 4096 and the kernels are told n = 2048, so the ids 2048..4095
    in the list stay inside the arrays whatever the kernels do."""
    import ctypes as C

    import torch

    from kaspan_amd import _lib

    g = ka.generate("gnm", 5, n=4096, draws=16384)
    ctx = ka.Context.default()

    def flagged(ids):
        lst = torch.tensor(np.array(ids, np.uint32).view(np.int32), device="cuda")
        out = C.c_uint32(0)
        gv = g.view()
        _lib.check(_lib.lib().kspn_test_tail_guard(ctx.ptr, C.byref(gv), 2048, C.c_void_p(lst.data_ptr()), len(ids),
                                                   C.byref(out), None), "kspn_test_tail_guard")
        return out.value

    assert flagged(list(range(0, 2048, 7))) == 0
    assert flagged([3, 2048, 9, 4095, 11]) == 7
    # the flag is cleared: the next solve works and its labels are right
    lab, _ = gpu_labels(ka, g)
    assert np.array_equal(lab, O.tarjan(host_graph(O, g)))
