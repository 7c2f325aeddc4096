"""ctypes view of the CPU oracle (libkspn_oracle.so). TEST INFRASTRUCTURE ONLY.

Loaded by tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` -- as the checker or the timed CPU baseline, never as the product.
Every function mirrors a C function in ``kspn_oracle.c`` whose header cites the
KaSpan source it restates.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libkspn_oracle.so")
UNDECIDED = 0xFFFFFFFF

_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


class SccStats(C.Structure):
    _fields_ = [
        ("pivot", C.c_uint32),
        ("decided_trim_first", C.c_uint32),
        ("decided_fwbw", C.c_uint32),
        ("decided_trim_normal", C.c_uint32),
        ("decided_color", C.c_uint32),
        ("color_rounds", C.c_uint32),
        ("decided_residual", C.c_uint32),
        ("residual_n", C.c_uint32),
        ("residual_m", C.c_uint64),
    ]


def build() -> str:
    """Compile the oracle with its Makefile (gcc, -O3)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        u32, u64, dbl, i32 = C.c_uint32, C.c_uint64, C.c_double, C.c_int
        sig = {
            "orc_splitmix64_at": (u64, [u64, u64]),
            "orc_perm_pow2": (u32, [u64, u32, u32]),
            "orc_gen_gnm": (u64, [u64, u32, u64, _u32p, _u32p]),
            "orc_gen_rmat": (u64, [u64, u32, u64, dbl, dbl, dbl, i32, _u32p, _u32p]),
            "orc_gen_lattice": (u64, [u64, u32, dbl, _u32p, _u32p]),
            "orc_gen_rmat_n": (u64, [u64, u32, u64, dbl, dbl, dbl, _u32p, _u32p]),
            "orc_batch_plan": (None, [u64, u32, u32, u32, dbl, dbl, dbl, _u32p, _u32p, _u64p, _u64p]),
            "orc_lattice_max_edges": (u64, [u32]),
            "orc_gen_fuzzy": (u64, [u64, u32, dbl, _u32p, _u32p, u64, _u32p]),
            "orc_gen_planted": (u64, [u64, u32, dbl, u32, dbl, dbl, _u32p, _u32p, u64, _u32p]),
            "orc_planted_sizes": (u64, [u64, u32, dbl, u32, _u32p, u64]),
            "orc_csr_from_edges": (u64, [u32, u64, _u32p, _u32p, _u64p, _u32p]),
            "orc_transpose": (None, [u32, _u64p, _u32p, _u64p, _u32p]),
            "orc_subgraph": (u64, [u32, _u64p, _u32p, _u8p, _u32p, C.POINTER(u32), _u64p, _u32p]),
            "orc_tarjan": (u32, [u32, _u64p, _u32p, _u32p]),
            "orc_tarjan_batch": (None, [u32, _u32p, _u64p, _u32p, _u32p, u32]),
            "orc_planted_labels": (None, [u64, u32, dbl, u32, _u32p]),
            "orc_canonicalize": (None, [u32, _u32p]),
            "orc_brute_force": (i32, [u32, _u64p, _u32p, _u32p]),
            "orc_trim_1_first": (u32, [u32, _u64p, _u64p, _u8p, _u32p]),
            "orc_fwbw": (u32, [u32, _u64p, _u32p, _u64p, _u32p, _u8p, u32, _u32p]),
            "orc_color_round": (u32, [u32, _u64p, _u32p, _u64p, _u32p, _u8p, u32, _u32p, _u32p]),
            "orc_kaspan_scc": (i32, [u32, _u64p, _u32p, _u64p, _u32p, _u32p, i32, C.POINTER(SccStats)]),
            "orc_scc_trim_ex_residual": (i32, [u32, _u64p, _u32p, _u64p, _u32p, _u32p, C.POINTER(SccStats)]),
            "orc_scc_trim_ex": (i32, [u32, _u64p, _u32p, _u64p, _u32p, _u32p, C.POINTER(SccStats)]),
            "orc_scc_variant": (i32, [u32, _u64p, _u32p, _u64p, _u32p, i32, _u32p, C.POINTER(SccStats)]),
            "orc_trim_fixpoint": (u32, [u32, _u64p, _u32p, _u64p, _u32p, _u8p]),
            "orc_pivot_features": (None, [u32, _u64p, _u32p, _u64p, _u32p, _f32p]),
            "orc_mlp_forward": (None, [u32, _f32p, _f64p, _f64p, _f64p, _f64p, _f64p, dbl, _f64p]),
            "orc_mlp_loss_grad": (dbl, [u32, _f32p, _f32p, _f64p, _f64p, _f64p, _f64p, _f64p, dbl,
                                        _f64p, _f64p, _f64p, _f64p, _f64p, C.POINTER(dbl)]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


# --------------------------------------------------------------------------- graphs
class HostGraph:
    """CSR + reverse CSR on the host: uint64 offsets, uint32 ids."""

    def __init__(self, n, fw_off, fw_adj, bw_off=None, bw_adj=None):
        self.n = int(n)
        self.fw_off = np.ascontiguousarray(fw_off, dtype=np.uint64)
        self.fw_adj = np.ascontiguousarray(fw_adj, dtype=np.uint32)
        if bw_off is None:
            bw_off, bw_adj = transpose(self.n, self.fw_off, self.fw_adj)
        self.bw_off = np.ascontiguousarray(bw_off, dtype=np.uint64)
        self.bw_adj = np.ascontiguousarray(bw_adj, dtype=np.uint32)

    @property
    def m(self):
        return int(self.fw_off[-1]) if self.n else 0


def csr_from_edges(n, src, dst) -> HostGraph:
    src = np.ascontiguousarray(src, dtype=np.uint32)
    dst = np.ascontiguousarray(dst, dtype=np.uint32)
    off = np.zeros(n + 1, np.uint64)
    adj = np.zeros(max(len(src), 1), np.uint32)
    m = lib().orc_csr_from_edges(n, len(src), src, dst, off, adj)
    return HostGraph(n, off, adj[:m].copy())


def transpose(n, fw_off, fw_adj):
    m = int(fw_off[-1]) if n else 0
    bw_off = np.zeros(n + 1, np.uint64)
    bw_adj = np.zeros(max(m, 1), np.uint32)
    lib().orc_transpose(n, np.ascontiguousarray(fw_off, np.uint64), np.ascontiguousarray(fw_adj if m else np.zeros(1, np.uint32), np.uint32), bw_off, bw_adj)
    return bw_off, bw_adj[:m].copy()


def _nonempty(a, dtype):
    a = np.ascontiguousarray(a, dtype=dtype)
    return a if a.size else np.zeros(1, dtype)


def gen_gnm(seed, n, draws):
    src = np.zeros(max(draws, 1), np.uint32)
    dst = np.zeros(max(draws, 1), np.uint32)
    k = lib().orc_gen_gnm(seed, n, draws, src, dst)
    return src[:k], dst[:k]


def gen_rmat(seed, scale, draws, a=0.57, b=0.19, c=0.19, permute=True):
    src = np.zeros(max(draws, 1), np.uint32)
    dst = np.zeros(max(draws, 1), np.uint32)
    k = lib().orc_gen_rmat(seed, scale, draws, a, b, c, int(permute), src, dst)
    return src[:k], dst[:k]


def gen_rmat_n(seed, n, draws, a=0.57, b=0.19, c=0.19):
    src = np.zeros(max(draws, 1), np.uint32)
    dst = np.zeros(max(draws, 1), np.uint32)
    k = lib().orc_gen_rmat_n(seed, n, draws, a, b, c, src, dst)
    return src[:k], dst[:k]


def batch_plan(seed, count, n_min=1000, n_max=1_000_000, deg_min=2.0, deg_max=16.0, p_rmat=0.5):
    """Per-graph (kind, n, draws, seed) arrays of a batch (kind 0 G(n,m), 1 R-MAT)."""
    kind = np.zeros(max(count, 1), np.uint32)
    n = np.zeros(max(count, 1), np.uint32)
    draws = np.zeros(max(count, 1), np.uint64)
    gseed = np.zeros(max(count, 1), np.uint64)
    lib().orc_batch_plan(seed, count, n_min, n_max, deg_min, deg_max, p_rmat, kind, n, draws, gseed)
    return kind[:count].copy(), n[:count].copy(), draws[:count].copy(), gseed[:count].copy()


def gen_batch(seed, count, n_min=1000, n_max=1_000_000, deg_min=2.0, deg_max=16.0, p_rmat=0.5, a=0.57, b=0.19, c=0.19):
    """The graphs of a batch plan, each deduplicated, and their union with
    batch-global ids: (union HostGraph, vtx_off[count+1], edge_off[count+1], graphs)."""
    kind, ns, draws, gseed = batch_plan(seed, count, n_min, n_max, deg_min, deg_max, p_rmat)
    graphs = []
    for i in range(count):
        if kind[i]:
            src, dst = gen_rmat_n(int(gseed[i]), int(ns[i]), int(draws[i]), a, b, c)
        else:
            src, dst = gen_gnm(int(gseed[i]), int(ns[i]), int(draws[i]))
        graphs.append(csr_from_edges(int(ns[i]), src, dst))
    vtx_off = np.zeros(count + 1, np.uint32)
    edge_off = np.zeros(count + 1, np.uint64)
    for i, g in enumerate(graphs):
        vtx_off[i + 1] = vtx_off[i] + g.n
        edge_off[i + 1] = edge_off[i] + g.m
    return pack_batch(graphs), vtx_off, edge_off, graphs


def pack_batch(graphs) -> "HostGraph":
    """Union of independent graphs: graph i's ids shift by the vertices before it."""
    fo, fa, bo, ba = [np.zeros(1, np.uint64)], [], [np.zeros(1, np.uint64)], []
    vb, eb = 0, 0
    for g in graphs:
        fo.append(g.fw_off[1:] + np.uint64(eb))
        bo.append(g.bw_off[1:] + np.uint64(eb))
        fa.append(g.fw_adj[: g.m] + np.uint32(vb))
        ba.append(g.bw_adj[: g.m] + np.uint32(vb))
        vb += g.n
        eb += g.m
    cat = lambda xs, dt: np.concatenate(xs).astype(dt) if xs else np.zeros(0, dt)
    return HostGraph(vb, cat(fo, np.uint64), cat(fa, np.uint32), cat(bo, np.uint64), cat(ba, np.uint32))


def gen_lattice(seed, side, p_bidir=0.1):
    cap = max(int(lib().orc_lattice_max_edges(side)), 1)
    src = np.zeros(cap, np.uint32)
    dst = np.zeros(cap, np.uint32)
    k = lib().orc_gen_lattice(seed, side, p_bidir, src, dst)
    return src[:k], dst[:k]


def gen_fuzzy(seed, n, d=-1.0):
    """Planted-SCC graph shaped like test/fuzzy.hpp; returns (HostGraph, labels)."""
    cap = max(int(n * max(n - 1, 1)), 1) if n < 3000 else int(n * 64)
    src = np.zeros(cap, np.uint32)
    dst = np.zeros(cap, np.uint32)
    labels = np.zeros(max(n, 1), np.uint32)
    k = lib().orc_gen_fuzzy(seed, n, float(d), src, dst, cap, labels)
    return csr_from_edges(n, src[:k], dst[:k]), labels[:n].copy()


def gen_planted(seed, n, expo=2.1, max_size=10_000_000, intra_deg=4.0, inter_deg=3.0):
    cap = int(n * (round(intra_deg) + round(inter_deg) + 1)) + 1
    src = np.zeros(cap, np.uint32)
    dst = np.zeros(cap, np.uint32)
    labels = np.zeros(max(n, 1), np.uint32)
    k = lib().orc_gen_planted(seed, n, expo, max_size, intra_deg, inter_deg, src, dst, cap, labels)
    return csr_from_edges(n, src[:k], dst[:k]), labels[:n].copy()


def planted_sizes(seed, n, expo=2.1, max_size=10_000_000):
    sizes = np.zeros(max(n, 1), np.uint32)
    k = lib().orc_planted_sizes(seed, n, expo, max_size, sizes, n)
    return sizes[:k].copy()


# --------------------------------------------------------------------------- SCC
def tarjan(g: HostGraph) -> np.ndarray:
    labels = np.zeros(max(g.n, 1), np.uint32)
    lib().orc_tarjan(g.n, g.fw_off, _nonempty(g.fw_adj, np.uint32), labels)
    return labels[: g.n].copy()


def tarjan_batch(vtx_off, fw_off, fw_adj, threads=16) -> np.ndarray:
    """Tarjan labels of every graph of a batch union, local to each graph
    (graph i owns vertices [vtx_off[i], vtx_off[i+1]))."""
    vtx_off = np.ascontiguousarray(vtx_off, np.uint32)
    n = int(vtx_off[-1])
    labels = np.zeros(max(n, 1), np.uint32)
    lib().orc_tarjan_batch(len(vtx_off) - 1, vtx_off, np.ascontiguousarray(fw_off, np.uint64),
                           _nonempty(np.ascontiguousarray(fw_adj, np.uint32), np.uint32), labels, threads)
    return labels[:n]


def planted_labels(seed, n, expo=2.1, max_size=10_000_000) -> np.ndarray:
    """the labels gen_planted returns, without generating the edges"""
    labels = np.zeros(max(n, 1), np.uint32)
    lib().orc_planted_labels(seed, n, expo, max_size, labels)
    return labels[:n]


def canonicalize(labels) -> np.ndarray:
    lab = np.array(labels, dtype=np.uint32, copy=True)
    if lab.size:
        lib().orc_canonicalize(lab.size, lab)
    return lab


def brute_force(g: HostGraph) -> np.ndarray:
    labels = np.zeros(max(g.n, 1), np.uint32)
    rc = lib().orc_brute_force(g.n, g.fw_off, _nonempty(g.fw_adj, np.uint32), labels)
    if rc:
        raise ValueError("brute force is limited to n <= 4096")
    return labels[: g.n].copy()


def _pipeline(fn, g: HostGraph, *extra):
    labels = np.full(max(g.n, 1), UNDECIDED, np.uint32)
    st = SccStats()
    rc = fn(g.n, g.fw_off, _nonempty(g.fw_adj, np.uint32), g.bw_off, _nonempty(g.bw_adj, np.uint32),
            labels, *extra, C.byref(st))
    if rc:
        raise RuntimeError(f"oracle pipeline failed with {rc}")
    return labels[: g.n].copy(), st


def kaspan_scc(g: HostGraph, faithful=False):
    """Restated kaspan::scc (scc/scc.hpp:21-87): raw ids, stats."""
    return _pipeline(lib().orc_kaspan_scc, g, int(faithful))


def scc_trim_ex_residual(g: HostGraph):
    return _pipeline(lib().orc_scc_trim_ex_residual, g)


def scc_trim_ex(g: HostGraph):
    return _pipeline(lib().orc_scc_trim_ex, g)


def scc_variant(g: HostGraph, method: int):
    """scc_hpc_like (2), scc_ispan_like (3), scc_hpc_trim_ex (4), scc_trim_ex_light_residual (6)
    at one rank (scc/variant/*.hpp): raw ids, stats"""
    labels = np.full(max(g.n, 1), UNDECIDED, np.uint32)
    st = SccStats()
    rc = lib().orc_scc_variant(g.n, g.fw_off, _nonempty(g.fw_adj, np.uint32), g.bw_off,
                               _nonempty(g.bw_adj, np.uint32), int(method), labels, C.byref(st))
    if rc:
        raise RuntimeError(f"oracle variant {method} failed with {rc}")
    return labels[: g.n].copy(), st


def trim_fixpoint(g: HostGraph, live=None):
    """(count, live) of the trim-1 fixpoint from `live` (all vertices if None)"""
    live = np.ones(max(g.n, 1), np.uint8) if live is None else np.array(live, np.uint8, copy=True)
    if live.size == 0:
        live = np.ones(1, np.uint8)
    d = lib().orc_trim_fixpoint(g.n, g.fw_off, _nonempty(g.fw_adj, np.uint32), g.bw_off,
                                _nonempty(g.bw_adj, np.uint32), live)
    return int(d), live[: g.n].copy()


# --------------------------------------------------------------------------- manifests
def decode_dense_unsigned(data: bytes, count: int, width: int, endian: str) -> np.ndarray:
    """dense_unsigned_ops::get for every element (memory/dense_unsigned_ops.hpp:14-44):
    element i is bytes [i*width, (i+1)*width) read in `endian` byte order"""
    b = np.frombuffer(data, dtype=np.uint8, count=count * width).reshape(count, width).astype(np.uint64)
    if endian == "big":
        b = b[:, ::-1]
    out = np.zeros(count, np.uint64)
    for j in range(width):
        out |= b[:, j] << np.uint64(8 * j)
    return out


def read_manifest(path: str) -> dict:
    """manifest::load (scc/adapter/manifest.hpp:20-51,136-197): key/value lines, '%' comments,
    key up to the first space, value after the spaces; paths relative to the manifest"""
    kv = {}
    with open(path, "r", newline="\n") as f:
        for line in f.read().split("\n"):
            if not line or line[0] == "%":
                continue
            if " " not in line:
                raise ValueError(f"line without a value: {line!r}")
            k, v = line.split(" ", 1)
            if k in kv:
                raise ValueError(f"repeated key {k}")
            kv[k] = v.lstrip(" ")
    base = os.path.dirname(os.path.abspath(path))
    out = {
        "schema_version": int(kv["schema.version"]), "code": kv["graph.code"], "name": kv["graph.name"],
        "endian": kv["graph.endian"], "node_count": int(kv["graph.node_count"]),
        "edge_count": int(kv["graph.edge_count"]),
        "contains_self_loops": {"true": True, "false": False}[kv["graph.contains_self_loops"]],
        "contains_duplicate_edges": {"true": True, "false": False}[kv["graph.contains_duplicate_edges"]],
        "head_bytes": int(kv["graph.head.bytes"]), "csr_bytes": int(kv["graph.csr.bytes"]),
    }
    if out["endian"] not in ("little", "big"):
        raise ValueError("graph.endian")
    for key in ("fw.head.path", "fw.csr.path", "bw.head.path", "bw.csr.path"):
        out[key.replace(".", "_")] = os.path.realpath(os.path.join(base, kv[key]))
    return out


def load_manifest(path: str) -> HostGraph:
    """load_graph_from_manifest (scc/adapter/manifest.hpp:200-250) on the host"""
    mf = read_manifest(path)
    n, m = mf["node_count"], mf["edge_count"]
    arrays = []
    for key, count, width in (("fw_head_path", n + 1, mf["head_bytes"]), ("fw_csr_path", m, mf["csr_bytes"]),
                              ("bw_head_path", n + 1, mf["head_bytes"]), ("bw_csr_path", m, mf["csr_bytes"])):
        with open(mf[key], "rb") as f:
            data = f.read()
        if len(data) != count * width:
            raise ValueError(f"{mf[key]}: {len(data)} bytes, expected {count * width}")
        arrays.append(decode_dense_unsigned(data, count, width, mf["endian"]))
    fo, fa, bo, ba = arrays
    if np.any(fo > m) or np.any(bo > m) or (m and (fa.max() >= n or ba.max() >= n)):
        raise ValueError("value out of range")
    return HostGraph(n, fo, fa.astype(np.uint32), bo, ba.astype(np.uint32))


def trim_1_first(g: HostGraph):
    live = np.ones(max(g.n, 1), np.uint8)
    scc_id = np.full(max(g.n, 1), UNDECIDED, np.uint32)
    pivot = lib().orc_trim_1_first(g.n, g.fw_off, g.bw_off, live, scc_id)
    return int(pivot), live[: g.n].copy(), scc_id[: g.n].copy()


def fwbw(g: HostGraph, live, pivot):
    live = np.array(live, np.uint8, copy=True)
    scc_id = np.full(max(g.n, 1), UNDECIDED, np.uint32)
    d = lib().orc_fwbw(g.n, g.fw_off, _nonempty(g.fw_adj, np.uint32), g.bw_off, _nonempty(g.bw_adj, np.uint32), live, pivot, scc_id)
    return int(d), live, scc_id[: g.n].copy()


def color_round(g: HostGraph, live, rotation):
    live = np.array(live, np.uint8, copy=True)
    color = np.zeros(max(g.n, 1), np.uint32)
    scc_id = np.full(max(g.n, 1), UNDECIDED, np.uint32)
    d = lib().orc_color_round(g.n, g.fw_off, _nonempty(g.fw_adj, np.uint32), g.bw_off, _nonempty(g.bw_adj, np.uint32), live, rotation, color, scc_id)
    return int(d), live, color[: g.n].copy(), scc_id[: g.n].copy()


def subgraph(g: HostGraph, mask):
    mask = np.ascontiguousarray(mask, dtype=np.uint8)
    super_ids = np.zeros(max(g.n, 1), np.uint32)
    sub_off = np.zeros(g.n + 1, np.uint64)
    sub_adj = np.zeros(max(g.m, 1), np.uint32)
    sn = C.c_uint32(0)
    sm = lib().orc_subgraph(g.n, g.fw_off, _nonempty(g.fw_adj, np.uint32), mask, super_ids, C.byref(sn), sub_off, sub_adj)
    return super_ids[: sn.value].copy(), sub_off[: sn.value + 1].copy(), sub_adj[:sm].copy()


def pivot_features(g: HostGraph) -> np.ndarray:
    feat = np.zeros((max(g.n, 1), 6), np.float32)
    lib().orc_pivot_features(g.n, g.fw_off, _nonempty(g.fw_adj, np.uint32), g.bw_off, _nonempty(g.bw_adj, np.uint32), feat)
    return feat[: g.n].copy()


def mlp_forward(x, w1, b1, w2, b2, w3, b3):
    x = np.ascontiguousarray(x, np.float32)
    out = np.zeros(max(len(x), 1), np.float64)
    lib().orc_mlp_forward(len(x), x, *(np.ascontiguousarray(a, np.float64) for a in (w1, b1, w2, b2, w3)), float(b3), out)
    return out[: len(x)]


def mlp_loss_grad(x, t, w1, b1, w2, b2, w3, b3):
    x = np.ascontiguousarray(x, np.float32)
    t = np.ascontiguousarray(t, np.float32)
    g = [np.zeros(192), np.zeros(32), np.zeros(1024), np.zeros(32), np.zeros(32)]
    gb3 = C.c_double(0)
    loss = lib().orc_mlp_loss_grad(len(x), x, t, *(np.ascontiguousarray(a, np.float64) for a in (w1, b1, w2, b2, w3)), float(b3), *g, C.byref(gb3))
    return loss, g + [gb3.value]
