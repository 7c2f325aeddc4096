"""Manifest loading on the GPU (kspn_manifest_load / kspn_unpack_dense) against
the oracle's host reading of the same files, the solver on loaded graphs
(self-loops and duplicate edges included) against Tarjan, and the loader's
rejections: wrong file sizes, out-of-range values, malformed offsets."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def M():
    from kaspan_amd import manifest

    return manifest


def same_graph(g, h):
    fo, fa, bo, ba = g.to_host()
    return (np.array_equal(fo, h.fw_off) and np.array_equal(fa, h.fw_adj) and np.array_equal(bo, h.bw_off)
            and np.array_equal(ba, h.bw_adj))


@pytest.mark.parametrize("head_bytes,csr_bytes,endian", [(None, None, "little"), (3, 2, "big"), (8, 8, "little"),
                                                         (5, 3, "big"), (2, 7, "little"), (4, 4, "big")])
def test_load_matches_oracle_and_solves(ka, O, M, tmp_path, head_bytes, csr_bytes, endian):
    print(f"manifest widths {head_bytes}/{csr_bytes} {endian}", flush=True)
    s, d = O.gen_rmat(21, 12, 12 << 12)
    hg = O.csr_from_edges(1 << 12, s, d)
    path = M.write_manifest(str(tmp_path / "r12"), hg.n, hg.fw_off, hg.fw_adj, hg.bw_off, hg.bw_adj,
                            endian=endian, head_bytes=head_bytes, csr_bytes=csr_bytes)
    ref = O.load_manifest(path)
    g = M.load_manifest(path)
    assert (g.n, g.m) == (hg.n, hg.m)
    assert same_graph(g, ref) and same_graph(g, hg)
    lab, _ = ka.scc(g)
    assert np.array_equal(lab.cpu().numpy().view(np.uint32), O.tarjan(hg))


def test_multi_chunk_load(ka, O, M, tmp_path):
    """more than one 64 MiB staging chunk per array (3-byte ids, big endian)"""
    print("manifest multi-chunk rmat21", flush=True)
    g0 = ka.generate("rmat", 41, scale=21, draws=32 << 20)
    fo, fa, bo, ba = g0.to_host()
    assert g0.m * 3 > 64 << 20
    path = M.write_manifest(str(tmp_path / "big"), g0.n, fo, fa, bo, ba, endian="big", head_bytes=5, csr_bytes=3)
    g = M.load_manifest(path)
    assert all(np.array_equal(a, b) for a, b in zip(g.to_host(), (fo, fa, bo, ba)))


def test_self_loops_and_duplicates(ka, O, M, tmp_path):
    """convert_edgelist keeps self-loops and duplicate edges, as the reference's
    converter does; every variant still gives Tarjan's labels"""
    rng = np.random.default_rng(5)
    n = 3000
    u = rng.integers(0, n, 12000)
    v = (u + rng.integers(-3, 4, 12000)) % n  # local edges: many small cycles
    u = np.concatenate([u, u[:500], np.arange(0, n, 7)])
    v = np.concatenate([v, v[:500], np.arange(0, n, 7)])  # duplicates, self-loops
    src = tmp_path / "loops.txt"
    src.write_text("% test\n" + "".join(f"{a} {b}\n" for a, b in zip(u, v)))
    path = M.convert_edgelist(str(src))
    mf = M.parse_manifest(path)
    assert mf.contains_self_loops and mf.contains_duplicate_edges
    hg = O.load_manifest(path)
    ref = O.tarjan(hg)
    g = M.load_manifest(path)
    assert same_graph(g, hg)
    for variant in ka.VARIANTS:
        lab, _ = ka.scc_variant(variant, g)
        assert np.array_equal(lab.cpu().numpy().view(np.uint32), ref), variant


def test_empty_and_edgeless(ka, M, tmp_path):
    path = M.write_manifest(str(tmp_path / "iso"), 5, [0] * 6, [])
    g = M.load_manifest(path)
    assert (g.n, g.m) == (5, 0)
    lab, _ = ka.scc(g)
    assert lab.cpu().tolist() == [0, 1, 2, 3, 4]
    path = M.write_manifest(str(tmp_path / "nil"), 0, [0], [])
    g = M.load_manifest(path)
    assert (g.n, g.m) == (0, 0)


def _poke(path, index, value, width=2):
    """overwrite element `index` of a little-endian dense array"""
    with open(path, "r+b") as f:
        f.seek(index * width)
        f.write(int(value).to_bytes(width, "little"))


@pytest.mark.parametrize("case", ["short_file", "long_file", "id_out_of_range", "offset_out_of_range",
                                  "decreasing", "last_offset", "first_offset"])
def test_loader_rejects(ka, M, tmp_path, case):
    from kaspan_amd import KspnError
    from tests.helpers import ISPAN_TINY_BW_ADJ, ISPAN_TINY_BW_OFF, ISPAN_TINY_FW_ADJ, ISPAN_TINY_FW_OFF

    print(f"manifest rejection {case}", flush=True)
    prefix = str(tmp_path / "t")
    path = M.write_manifest(prefix, 7, ISPAN_TINY_FW_OFF, ISPAN_TINY_FW_ADJ, ISPAN_TINY_BW_OFF, ISPAN_TINY_BW_ADJ,
                            head_bytes=2, csr_bytes=2)
    code = -1
    if case == "short_file":
        with open(prefix + ".bw.csr.bin", "r+b") as f:
            f.truncate(21)
        code = -6
    elif case == "long_file":
        with open(prefix + ".fw.head.bin", "ab") as f:
            f.write(b"\0\0")
        code = -6
    elif case == "id_out_of_range":
        _poke(prefix + ".fw.csr.bin", 3, 7)  # n = 7
    elif case == "offset_out_of_range":
        _poke(prefix + ".bw.head.bin", 4, 12)  # m = 11
    elif case == "decreasing":
        _poke(prefix + ".fw.head.bin", 3, 7)
        _poke(prefix + ".fw.head.bin", 4, 5)
    elif case == "last_offset":
        _poke(prefix + ".bw.head.bin", 7, 10)
    elif case == "first_offset":
        _poke(prefix + ".fw.head.bin", 0, 1)
    with pytest.raises(KspnError) as e:
        M.load_manifest(path)
    assert e.value.code == code, str(e.value)
    # the same files, repaired, load again (the context is still usable)
    if case not in ("short_file", "long_file"):
        M.write_manifest(prefix, 7, ISPAN_TINY_FW_OFF, ISPAN_TINY_FW_ADJ, ISPAN_TINY_BW_OFF, ISPAN_TINY_BW_ADJ,
                         head_bytes=2, csr_bytes=2)
        g = M.load_manifest(path)
        assert ka.scc(g)[0].cpu().tolist() == [0, 1, 0, 1, 1, 5, 5]


@pytest.mark.parametrize("width", range(1, 9))
@pytest.mark.parametrize("endian", ["little", "big"])
def test_unpack_dense_matches_oracle(ka, O, width, endian):
    import torch

    rng = np.random.default_rng(100 + width)
    count = 100_003
    raw = rng.integers(0, 256, count * width, dtype=np.uint8)
    ref = O.decode_dense_unsigned(raw.tobytes(), count, width, endian)
    from kaspan_amd import manifest as M

    got = M.unpack_dense(torch.from_numpy(raw).cuda(), count, width, endian=endian)
    assert np.array_equal(got.cpu().numpy().view(np.uint64), ref)
    if width <= 4:
        got32 = M.unpack_dense(torch.from_numpy(raw).cuda(), count, width, endian=endian, out_bytes=4)
        assert np.array_equal(got32.cpu().numpy().view(np.uint32), ref.astype(np.uint32))
    # a limit below the largest value is reported
    from kaspan_amd import KspnError

    with pytest.raises(KspnError):
        M.unpack_dense(torch.from_numpy(raw).cuda(), count, width, endian=endian, limit=int(ref.max()) - 1)


def test_bench_driver_reads_manifest(ka, O, M, tmp_path):
    """python -m kaspan_amd.bench_kaspan --manifest_file: the statistics tree the
    reference's plotting reads (experiment/horeka/plot_utils.py:151-245)"""
    import json
    import subprocess
    import sys

    s, d = O.gen_gnm(3, 5000, 20000)
    hg = O.csr_from_edges(5000, s, d)
    path = M.write_manifest(str(tmp_path / "g"), hg.n, hg.fw_off, hg.fw_adj, hg.bw_off, hg.bw_adj)
    out = tmp_path / "stats.json"
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, "-m", "kaspan_amd.bench_kaspan", "--manifest_file", path, "--output_file",
                        str(out), "--variant", "ispan_like"], cwd=root, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    tree = json.loads(out.read_text())
    b = tree["0"]["benchmark"]
    assert (b["n"], b["m"], b["world_size"]) == (5000, hg.m, 1)
    assert b["load"]["duration"] > 0
    scc = b["scc"]
    assert scc["duration"] > 0
    stages = [k for k, v in scc.items() if isinstance(v, dict)]
    assert stages[0] == "trim_1_first" and "pivot" in scc["trim_1_first"]
    assert sum(scc[k]["decided_count"] for k in stages) == 5000
    tar = O.tarjan(hg)
    assert b["global_component_count"] == int(np.sum(tar == np.arange(5000)))
