"""Manifest graphs on the CPU: the oracle's decoder against Python's own
integer decoding, the writer and edge-list converter against hand-computed
files, and libkspn's manifest parser (host code of the C ABI, no GPU calls)
against the oracle's reading of the same files, including every malformed
case the reference rejects (scc/adapter/manifest.hpp:20-197)."""
import os
import struct

import numpy as np
import pytest

from tests.helpers import ISPAN_TINY_BW_ADJ, ISPAN_TINY_BW_OFF, ISPAN_TINY_FW_ADJ, ISPAN_TINY_FW_OFF


@pytest.fixture(scope="module")
def M():
    from kaspan_amd import manifest

    return manifest


def test_representing_bytes_boundaries(M):
    """memory/buffer.hpp:14-25: the smallest byte count whose range holds the value"""
    assert M.representing_bytes(0) == 1
    for b in range(1, 8):
        assert M.representing_bytes((1 << (8 * b)) - 1) == b
        assert M.representing_bytes(1 << (8 * b)) == b + 1
    assert M.representing_bytes((1 << 64) - 1) == 8


@pytest.mark.parametrize("width", range(1, 9))
@pytest.mark.parametrize("endian", ["little", "big"])
def test_dense_unsigned_decode_matches_int_from_bytes(O, M, width, endian):
    """the oracle's dense_unsigned decoding equals int.from_bytes element by element,
    and the writer's encoding equals int.to_bytes (and struct for 1/2/4/8 bytes)"""
    rng = np.random.default_rng(width * 2 + (endian == "big"))
    vals = [int(x) for x in rng.integers(0, 1 << min(8 * width, 63), size=257, dtype=np.uint64)]
    vals += [0, (1 << (8 * width)) - 1]
    enc = M._encode(np.array(vals, dtype=np.uint64), width, endian)
    assert enc == b"".join(v.to_bytes(width, endian) for v in vals)
    if width in (1, 2, 4, 8):
        fmt = {1: "B", 2: "H", 4: "I", 8: "Q"}[width]
        assert enc == struct.pack(("<" if endian == "little" else ">") + fmt * len(vals), *vals)
    dec = O.decode_dense_unsigned(enc, len(vals), width, endian)
    assert [int(x) for x in dec] == vals
    # a transposed byte order is caught
    if width > 1:
        other = "big" if endian == "little" else "little"
        assert [int(x) for x in O.decode_dense_unsigned(enc, len(vals), width, other)] != vals


def _tiny(M, tmp_path, **kw):
    return M.write_manifest(str(tmp_path / "tiny"), 7, ISPAN_TINY_FW_OFF, ISPAN_TINY_FW_ADJ, ISPAN_TINY_BW_OFF,
                            ISPAN_TINY_BW_ADJ, **kw)


def test_writer_layout_and_oracle_round_trip(O, M, tmp_path):
    path = _tiny(M, tmp_path)
    text = open(path).read().splitlines()
    # the key order and values convert_graph writes (external_edgelist.hpp:353-366)
    assert text == ["schema.version 1", "graph.code tiny", "graph.name tiny", "graph.endian little",
                    "graph.node_count 7", "graph.edge_count 11", "graph.contains_self_loops false",
                    "graph.contains_duplicate_edges false", "graph.head.bytes 1", "graph.csr.bytes 1",
                    "fw.head.path tiny.fw.head.bin", "fw.csr.path tiny.fw.csr.bin", "bw.head.path tiny.bw.head.bin",
                    "bw.csr.path tiny.bw.csr.bin"]
    assert open(tmp_path / "tiny.fw.csr.bin", "rb").read() == bytes(ISPAN_TINY_FW_ADJ)
    g = O.load_manifest(path)
    assert g.n == 7 and g.m == 11
    assert list(g.fw_off) == ISPAN_TINY_FW_OFF and list(g.fw_adj) == ISPAN_TINY_FW_ADJ
    assert list(g.bw_off) == ISPAN_TINY_BW_OFF and list(g.bw_adj) == ISPAN_TINY_BW_ADJ


@pytest.mark.parametrize("head_bytes,csr_bytes,endian", [(2, 2, "little"), (3, 2, "big"), (8, 4, "little"),
                                                         (5, 7, "big"), (2, 8, "big")])
def test_writer_widths_round_trip(O, M, tmp_path, head_bytes, csr_bytes, endian):
    s, d = O.gen_gnm(11, 300, 2000)
    g = O.csr_from_edges(300, s, d)
    path = M.write_manifest(str(tmp_path / "g"), g.n, g.fw_off, g.fw_adj, g.bw_off, g.bw_adj, endian=endian,
                            head_bytes=head_bytes, csr_bytes=csr_bytes)
    h = O.load_manifest(path)
    for a, b in ((h.fw_off, g.fw_off), (h.fw_adj, g.fw_adj), (h.bw_off, g.bw_off), (h.bw_adj, g.bw_adj)):
        assert np.array_equal(a, b)
    assert os.path.getsize(tmp_path / "g.fw.head.bin") == 301 * head_bytes
    assert os.path.getsize(tmp_path / "g.bw.csr.bin") == g.m * csr_bytes


def test_writer_rejects_narrow_widths(O, M, tmp_path):
    from kaspan_amd import KspnError

    s, d = O.gen_gnm(11, 300, 2000)
    g = O.csr_from_edges(300, s, d)
    with pytest.raises(KspnError):
        M.write_manifest(str(tmp_path / "g"), g.n, g.fw_off, g.fw_adj, head_bytes=1)  # m > 255
    with pytest.raises(KspnError):
        M.write_manifest(str(tmp_path / "g"), g.n, g.fw_off, g.fw_adj, csr_bytes=1)  # ids > 255


def test_writer_transposes_when_bw_is_missing(O, M, tmp_path):
    s, d = O.gen_rmat(5, 8, 3000)
    g = O.csr_from_edges(256, s, d)
    path = M.write_manifest(str(tmp_path / "r"), g.n, g.fw_off, g.fw_adj)
    h = O.load_manifest(path)
    assert np.array_equal(h.bw_off, g.bw_off) and np.array_equal(h.bw_adj, g.bw_adj)


def test_convert_edgelist(O, M, tmp_path):
    """convert_graph (external_edgelist.hpp:196-367): comment lines skipped, extra
    columns ignored, self-loops and duplicates kept and flagged, ascending rows"""
    src = tmp_path / "web.edges.txt"
    src.write_text("% a comment\n# another\n// and another\n\n3 1 0.5\n0 2\n2 0\n2 0\n1 1\n3 0\n0 3\n")
    path = M.convert_edgelist(str(src), name="web graph")
    assert os.path.basename(path) == "web.manifest"
    mf = O.read_manifest(path)
    assert mf["code"] == "web" and mf["name"] == "web graph"
    assert mf["node_count"] == 4 and mf["edge_count"] == 7
    assert mf["contains_self_loops"] and mf["contains_duplicate_edges"]
    assert mf["head_bytes"] == 1 and mf["csr_bytes"] == 1
    g = O.load_manifest(path)
    # rows: 0 -> {2, 3}, 1 -> {1}, 2 -> {0, 0}, 3 -> {0, 1}
    assert list(g.fw_off) == [0, 2, 3, 5, 7] and list(g.fw_adj) == [2, 3, 1, 0, 0, 0, 1]
    # reverse rows: 0 <- {2, 2, 3}, 1 <- {1, 3}, 2 <- {0}, 3 <- {0}
    assert list(g.bw_off) == [0, 3, 5, 6, 7] and list(g.bw_adj) == [2, 2, 3, 1, 3, 0, 0]
    with pytest.raises(Exception):
        bad = tmp_path / "bad.txt"
        bad.write_text("1 x\n")
        M.convert_edgelist(str(bad))


def test_parser_agrees_with_oracle(O, M, tmp_path):
    path = _tiny(M, tmp_path, endian="big", head_bytes=3, csr_bytes=2, name="tiny graph")
    # comments and extra spaces after the key are allowed
    text = open(path).read().replace("graph.node_count 7", "graph.node_count    7")
    open(path, "w").write("% written by a test\n" + text + "% trailing comment\n")
    mf = M.parse_manifest(path)
    ref = O.read_manifest(path)
    assert mf.schema_version == ref["schema_version"] == 1
    assert (mf.code, mf.name, mf.endian) == (ref["code"], ref["name"], ref["endian"]) == ("tiny", "tiny graph", "big")
    assert (mf.node_count, mf.edge_count, mf.head_bytes, mf.csr_bytes) == (7, 11, 3, 2)
    assert (mf.contains_self_loops, mf.contains_duplicate_edges) == (False, False)
    for key in ("fw_head_path", "fw_csr_path", "bw_head_path", "bw_csr_path"):
        assert getattr(mf, key) == ref[key]


def _rewrite(path, old, new):
    text = open(path).read()
    assert old in text
    open(path, "w").write(text.replace(old, new))


@pytest.mark.parametrize("old,new,code,needle", [
    ("graph.edge_count 11\n", "", -1, "graph.edge_count"),               # missing key
    ("graph.node_count 7", "graph.node_count 7x", -1, "graph.node_count"),  # trailing garbage
    ("graph.node_count 7", "graph.node_count -7", -1, "graph.node_count"),  # sign
    ("graph.contains_self_loops false", "graph.contains_self_loops no", -1, "contains_self_loops"),
    ("graph.endian little", "graph.endian middle", -1, "endian"),
    ("schema.version 1\n", "schema.version 1\nschema.version 2\n", -1, "repeated key"),
    ("schema.version 1\n", "schema.version 1\nloose_line\n", -1, "without a value"),
    ("fw.csr.path tiny.fw.csr.bin", "fw.csr.path nowhere.bin", -6, "fw.csr.path"),
])
def test_parser_rejects_malformed(M, tmp_path, old, new, code, needle):
    from kaspan_amd import KspnError

    path = _tiny(M, tmp_path)
    _rewrite(path, old, new)
    with pytest.raises(KspnError) as e:
        M.parse_manifest(path)
    assert e.value.code == code
    assert needle in str(e.value)


def test_parser_missing_manifest(M, tmp_path):
    from kaspan_amd import KspnError

    with pytest.raises(KspnError) as e:
        M.parse_manifest(str(tmp_path / "absent.manifest"))
    assert e.value.code == -6
