"""Graphs stored as a manifest: a text file of ``key value`` lines naming four
binary arrays of dense unsigned integers (rows offsets and adjacency of the
CSR and of the reverse CSR), 1..8 bytes per element, little or big endian.

==================================  ==================================================================
kaspan_amd.manifest                 reference
==================================  ==================================================================
``parse_manifest(path)``            ``manifest::load`` (scc/adapter/manifest.hpp:136-197)
``load_manifest(path)``             ``load_graph_from_manifest`` (scc/adapter/manifest.hpp:200-250)
``unpack_dense(packed, ...)``       ``dense_unsigned_accessor::get`` (memory/dense_unsigned_ops.hpp:14-44)
``write_manifest(prefix, ...)``     the writing half of ``convert_graph`` (scc/adapter/external_edgelist.hpp:248-367)
``convert_edgelist(path)``          ``convert_graph`` (scc/adapter/external_edgelist.hpp:196-367)
``representing_bytes(v)``           ``representing_bytes`` (memory/buffer.hpp:14-25)
==================================  ==================================================================

Parsing and loading run in libkspn (C ABI ``kspn_manifest_*``): the packed
bytes stream through pinned host buffers into HBM and a kernel widens them
to the solver's layout with range checks. Writing and converting are host
utilities (numpy) for preparing inputs.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np
import torch

from . import _lib
from ._lib import KspnError

__all__ = ["Manifest", "parse_manifest", "load_manifest", "unpack_dense", "write_manifest", "convert_edgelist",
           "representing_bytes"]


def representing_bytes(max_val: int) -> int:
    """bytes needed to store max_val, 1..8 (memory/buffer.hpp:14-25)"""
    for b in range(1, 8):
        if max_val < (1 << (8 * b)):
            return b
    return 8


class Manifest:
    """the parsed key/value file (scc/adapter/manifest.hpp:117-134); paths are absolute"""

    def __init__(self, raw: _lib.Manifest):
        self.raw = raw
        self.schema_version = raw.schema_version
        self.code = raw.code.decode()
        self.name = raw.name.decode()
        self.endian = "big" if raw.big_endian else "little"
        self.node_count = raw.node_count
        self.edge_count = raw.edge_count
        self.contains_self_loops = bool(raw.contains_self_loops)
        self.contains_duplicate_edges = bool(raw.contains_duplicate_edges)
        self.head_bytes = raw.head_bytes
        self.csr_bytes = raw.csr_bytes
        self.fw_head_path = raw.fw_head_path.decode()
        self.fw_csr_path = raw.fw_csr_path.decode()
        self.bw_head_path = raw.bw_head_path.decode()
        self.bw_csr_path = raw.bw_csr_path.decode()

    def __repr__(self):
        return (f"Manifest(name={self.name!r}, n={self.node_count}, m={self.edge_count}, endian={self.endian}, "
                f"head_bytes={self.head_bytes}, csr_bytes={self.csr_bytes})")


def parse_manifest(path: str) -> Manifest:
    """Parse a manifest file (host only). Raises KspnError with KSPN_E_INVALID for a
    malformed file and KSPN_E_IO for a missing one; the message names the key."""
    raw = _lib.Manifest()
    _lib.check(_lib.lib().kspn_manifest_parse(os.fsencode(path), C.byref(raw)), "kspn_manifest_parse", detail=True)
    return Manifest(raw)


def load_manifest(mf, device="cuda", ctx=None, stream=None):
    """``load_graph_from_manifest``: the graph of a manifest (path or Manifest) as a
    BidiGraph in HBM (uint64 offsets, uint32 ids)."""
    from . import BidiGraph, Context, _require_gpu, _stream_ptr

    _require_gpu()
    if not isinstance(mf, Manifest):
        mf = parse_manifest(mf)
    if mf.node_count + 1 > 0xFFFFFFFF:
        raise KspnError(-1, "load_manifest: graph.node_count does not fit 32-bit vertex ids")
    dev = torch.device(device)
    ctx = ctx or Context.default(dev.index or 0)
    n, m = int(mf.node_count), int(mf.edge_count)
    fo = torch.empty(n + 1, dtype=torch.int64, device=dev)
    bo = torch.empty(n + 1, dtype=torch.int64, device=dev)
    fa = torch.empty(max(m, 1), dtype=torch.int32, device=dev)
    ba = torch.empty(max(m, 1), dtype=torch.int32, device=dev)
    rc = _lib.lib().kspn_manifest_load(ctx.ptr, C.byref(mf.raw), C.c_void_p(fo.data_ptr()), C.c_void_p(fa.data_ptr()),
                                       C.c_void_p(bo.data_ptr()), C.c_void_p(ba.data_ptr()), _stream_ptr(stream))
    _lib.check(rc, "kspn_manifest_load", detail=True)
    g = BidiGraph.__new__(BidiGraph)
    g.n, g.m, g.device = n, m, dev
    g.fw_off, g.fw_adj, g.bw_off, g.bw_adj = fo, fa[:m], bo, ba[:m]
    return g


def unpack_dense(packed: torch.Tensor, count: int, elem_bytes: int, endian="little", limit=None, out_bytes=8,
                 ctx=None, stream=None) -> torch.Tensor:
    """Widen ``count`` dense unsigned elements (``elem_bytes`` each) of a device uint8 tensor
    to int64 (``out_bytes`` 8) or int32 (4) bit patterns; KSPN_E_INVALID if one exceeds ``limit``."""
    from . import Context, _require_gpu, _stream_ptr

    _require_gpu()
    if packed.dtype != torch.uint8 or not packed.is_contiguous() or packed.numel() < count * elem_bytes:
        raise KspnError(-1, "unpack_dense: packed must be a contiguous uint8 tensor of count * elem_bytes bytes")
    if limit is None:
        limit = (1 << (8 * out_bytes)) - 1
    ctx = ctx or Context.default(packed.device.index or 0)
    out = torch.empty(max(count, 1), dtype=torch.int64 if out_bytes == 8 else torch.int32, device=packed.device)
    rc = _lib.lib().kspn_unpack_dense(ctx.ptr, C.c_void_p(packed.data_ptr()), count, elem_bytes,
                                      1 if endian == "big" else 0, limit, C.c_void_p(out.data_ptr()), out_bytes,
                                      _stream_ptr(stream))
    _lib.check(rc, "kspn_unpack_dense", detail=True)
    return out[:count]


def _encode(values: np.ndarray, width: int, endian: str) -> bytes:
    """dense unsigned encoding (dense_unsigned_ops::set, memory/dense_unsigned_ops.hpp:46-75)"""
    v = np.ascontiguousarray(values, dtype=np.uint64)
    if width in (1, 2, 4, 8):
        return v.astype(np.dtype(f"{'<' if endian == 'little' else '>'}u{width}")).tobytes()
    shifts = (8 * np.arange(width, dtype=np.uint64))
    b = ((v[:, None] >> shifts[None, :]) & np.uint64(0xFF)).astype(np.uint8)
    if endian == "big":
        b = b[:, ::-1]
    return np.ascontiguousarray(b).tobytes()


def _transpose(n, fw_off, fw_adj):
    """reverse CSR with ascending sources per row (the (v, u) sort of convert_graph)"""
    src = np.repeat(np.arange(n, dtype=np.uint64), np.diff(fw_off).astype(np.int64))
    dst = fw_adj.astype(np.uint64)
    order = np.lexsort((src, dst))
    bw_adj = src[order].astype(np.uint64)
    bw_off = np.zeros(n + 1, np.uint64)
    np.cumsum(np.bincount(dst.astype(np.int64), minlength=n), out=bw_off[1:])
    return bw_off, bw_adj


def write_manifest(prefix: str, n: int, fw_off, fw_adj, bw_off=None, bw_adj=None, *, code=None, name=None,
                   endian="little", head_bytes=None, csr_bytes=None) -> str:
    """Write ``prefix.manifest`` and ``prefix.{fw,bw}.{head,csr}.bin`` in the layout
    ``convert_graph`` writes (scc/adapter/external_edgelist.hpp:248-367): widths default to
    ``representing_bytes(m)`` and ``representing_bytes(largest id)``, paths are stored
    relative to the manifest's directory. Returns the manifest's path. Host arrays
    (numpy or CPU tensors) or a BidiGraph's ``to_host()`` tuple."""
    fw_off = np.asarray(fw_off, dtype=np.uint64)
    fw_adj = np.asarray(fw_adj, dtype=np.uint64)
    if fw_off.size != n + 1 or (n >= 0 and fw_off[-1] != fw_adj.size):
        raise KspnError(-1, "write_manifest: fw_off must have n+1 entries ending at len(fw_adj)")
    if bw_off is None:
        bw_off, bw_adj = _transpose(n, fw_off, fw_adj)
    bw_off = np.asarray(bw_off, dtype=np.uint64)
    bw_adj = np.asarray(bw_adj, dtype=np.uint64)
    m = int(fw_adj.size)
    max_node = int(fw_adj.max()) if m else 0
    head_bytes = head_bytes or representing_bytes(m)
    csr_bytes = csr_bytes or representing_bytes(max_node)
    if not (representing_bytes(m) <= head_bytes <= 8 and representing_bytes(max_node) <= csr_bytes <= 8):
        raise KspnError(-1, f"write_manifest: widths must hold the values and be at most 8 bytes "
                            f"(head needs {representing_bytes(m)}, csr {representing_bytes(max_node)})")
    src = np.repeat(np.arange(n, dtype=np.uint64), np.diff(fw_off).astype(np.int64))
    self_loops = bool(m and np.any(src == fw_adj))
    keys = src * np.uint64(n if n else 1) + fw_adj
    duplicates = bool(m and np.unique(keys).size != m)
    base = os.path.basename(prefix)
    code = code if code is not None else base
    name = name if name is not None else code
    files = {}
    for key, arr, w in (("fw.head", fw_off, head_bytes), ("fw.csr", fw_adj, csr_bytes), ("bw.head", bw_off, head_bytes),
                        ("bw.csr", bw_adj, csr_bytes)):
        fname = f"{base}.{key}.bin"
        with open(f"{prefix}.{key}.bin", "wb") as f:
            f.write(_encode(arr, w, endian))
        files[key] = fname
    lines = [
        "schema.version 1", f"graph.code {code}", f"graph.name {name}", f"graph.endian {endian}",
        f"graph.node_count {n}", f"graph.edge_count {m}",
        f"graph.contains_self_loops {'true' if self_loops else 'false'}",
        f"graph.contains_duplicate_edges {'true' if duplicates else 'false'}",
        f"graph.head.bytes {head_bytes}", f"graph.csr.bytes {csr_bytes}",
        f"fw.head.path {files['fw.head']}", f"fw.csr.path {files['fw.csr']}",
        f"bw.head.path {files['bw.head']}", f"bw.csr.path {files['bw.csr']}",
    ]
    path = f"{prefix}.manifest"
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    return path


def _parse_edge_lines(path):
    """(u, v) pairs of a whitespace-separated edge list; lines starting with '%', '#'
    or '//' and empty lines are skipped, further columns ignored
    (parse_uv and the line filter, scc/adapter/external_edgelist.hpp:71-100,233-236)"""
    us, vs = [], []
    with open(path, "r") as f:
        for line in f:
            if not line.strip() or line[0] in "%#" or line.startswith("//"):
                continue
            parts = line.split()
            if len(parts) < 2 or not parts[0].isdigit() or not parts[1].isdigit():
                raise KspnError(-1, f"convert_edgelist: cannot parse {line.rstrip()!r}")
            us.append(int(parts[0]))
            vs.append(int(parts[1]))
    return np.array(us, dtype=np.uint64), np.array(vs, dtype=np.uint64)


def convert_edgelist(input_file: str, name: str | None = None, endian="little") -> str:
    """``convert_graph`` (scc/adapter/external_edgelist.hpp:196-367) for edge lists that fit
    in host memory: ``n`` is the largest id + 1, self-loops and duplicate edges are kept
    (and flagged), rows list targets in ascending order. The files go next to the input,
    named after its stem up to the first '.'. Returns the manifest's path."""
    u, v = _parse_edge_lines(input_file)
    m = u.size
    max_node = int(max(u.max(), v.max())) if m else 0
    n = max_node + 1 if m else 0
    if n + 1 > 0xFFFFFFFF:
        raise KspnError(-1, "convert_edgelist: ids do not fit 32-bit vertex ids")
    order = np.lexsort((v, u))
    fw_adj = v[order]
    fw_off = np.zeros(n + 1, np.uint64)
    if n:
        np.cumsum(np.bincount(u.astype(np.int64), minlength=n), out=fw_off[1:])
    order = np.lexsort((u, v))
    bw_adj = u[order]
    bw_off = np.zeros(n + 1, np.uint64)
    if n:
        np.cumsum(np.bincount(v.astype(np.int64), minlength=n), out=bw_off[1:])
    d, fname = os.path.split(input_file)
    code = fname.split(".", 1)[0]
    return write_manifest(os.path.join(d, code), n, fw_off, fw_adj, bw_off, bw_adj, code=code,
                          name=name or code, endian=endian, head_bytes=representing_bytes(m),
                          csr_bytes=representing_bytes(max_node))
