"""kaspan_amd -- KaSpan's ``scc(graph) -> label per vertex`` on one MI355X.

Host-side mirror of the reference's interface for this path (paths are in
the rawbby/KaSpan tree):

=====================================  ===========================================================
kaspan_amd                             reference
=====================================  ===========================================================
``BidiGraph``                          ``bidi_graph`` / ``bidi_graph_view`` (graph/bidi_graph.hpp:17,250)
``scc(graph, scc_id)``                 ``scc(bidi_graph_part_view, vertex_t* scc_id)`` (scc/scc.hpp:21-87)
``scc_trim_ex(graph, scc_id)``         ``scc_trim_ex`` (scc/variant/scc_trim_ex.hpp:16-100)
``scc_trim_ex_residual(graph, ...)``   ``scc_trim_ex_residual`` (scc/variant/scc_trim_ex_residual.hpp:17-128)
``scc_trim_ex_light_residual``         ``scc_trim_ex_light_residual`` (scc/variant/scc_trim_ex_light_residual.hpp:16-131)
``scc_hpc_like`` / ``scc_hpc_trim_ex``  ``scc_hpc_like`` / ``scc_hpc_trim_ex`` (scc/variant/scc_hpc_*.hpp)
``scc_ispan_like``                     ``scc_ispan_like`` (scc/variant/scc_ispan_like.hpp:18-87)
``scc_variant(name, graph)``           the driver's ``--variant`` switch (tool/bench_kaspan/src/main.cpp:63-71)
``manifest.load_manifest(path)``       ``manifest::load`` + ``load_graph_from_manifest`` (scc/adapter/manifest.hpp)
``statistics.statistics_json``         ``KASPAN_STATISTIC_*`` tree written as JSON (debug/statistic.hpp:124-205)
``backward_complement_graph``          ``backward_complement_graph`` (scc/backward_complement.hpp:22-68)
``allgather_fw_sub_graph``             ``allgather_fw_sub_graph`` at one rank (scc/allgather_sub_graph.hpp:253-301)
``tarjan``                             ``tarjan`` with min-id callback (scc/tarjan.hpp:124-136)
``normalise_scc_id``                   ``normalise_scc_id`` (test/normalise_scc.hpp:10-25)
``Batch`` / ``scc_batch``              one ``scc`` call per graph, for many independent graphs at once
=====================================  ===========================================================

Every call goes through libkspn's C ABI (``include/kspn/scc.h``); device
memory is held in torch tensors (offsets as int64 = uint64 bits, ids as
int32 = uint32 bits). Differences: ``scc_id`` holds canonical ids (minimum
vertex id per component, the reference's ids after ``normalise_scc_id``);
errors raise ``KspnError`` where the reference aborts.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from ._lib import KspnError, Stats

__all__ = [
    "BidiGraph", "Context", "KspnError", "Stats", "scc", "scc_trim_ex", "scc_trim_ex_residual",
    "backward_complement_graph", "allgather_fw_sub_graph", "tarjan", "normalise_scc_id", "generate",
    "METHOD_KASPAN", "METHOD_TRIM_EX", "UNDECIDED", "planted_labels", "scc_sizes", "Batch", "generate_batch",
    "batch_plan", "scc_batch", "scc_hpc_like", "scc_hpc_trim_ex", "scc_ispan_like", "scc_trim_ex_light_residual",
    "scc_variant", "VARIANTS",
]

METHOD_KASPAN = _lib.METHOD_KASPAN
METHOD_TRIM_EX = _lib.METHOD_TRIM_EX
# the driver's --variant names (tool/bench_kaspan/src/main.cpp:63-71); None is kaspan::scc
VARIANTS = {
    None: _lib.METHOD_KASPAN,
    "async": _lib.METHOD_KASPAN,
    "hpc_like": _lib.METHOD_HPC_LIKE,
    "hpc_trim_ex": _lib.METHOD_HPC_TRIM_EX,
    "ispan_like": _lib.METHOD_ISPAN_LIKE,
    "trim_ex": _lib.METHOD_TRIM_EX,
    "trim_ex_light_residual": _lib.METHOD_TRIM_EX_LIGHT_RESIDUAL,
    "trim_ex_residual": _lib.METHOD_TRIM_EX_RESIDUAL,
}
UNDECIDED = _lib.UNDECIDED


def _require_gpu():
    if not torch.cuda.is_available():
        raise KspnError(-1, "kaspan_amd needs a HIP device (torch.cuda.is_available() is False)")


class Context:
    """Owns a kspn_ctx (scratch memory, events, pinned counters) on one device."""

    _default: dict = {}

    def __init__(self, device: int = 0):
        _require_gpu()
        self.device = int(device)
        self.ptr = C.c_void_p()
        _lib.check(_lib.lib().kspn_ctx_create(C.byref(self.ptr), self.device), "kspn_ctx_create")

    @classmethod
    def default(cls, device: int = 0) -> "Context":
        ctx = cls._default.get(device)
        if ctx is None:
            ctx = cls._default[device] = cls(device)
        return ctx

    def reserve(self, n: int, m: int):
        _lib.check(_lib.lib().kspn_ctx_reserve(self.ptr, n, m), "kspn_ctx_reserve")

    def close(self):
        if self.ptr:
            _lib.lib().kspn_ctx_destroy(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _stream_ptr(stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


def _as_u64_tensor(a, device):
    if isinstance(a, torch.Tensor):
        return a.to(device=device, dtype=torch.int64).contiguous()
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.uint64).view(np.int64)).to(device)


def _as_u32_tensor(a, device):
    if isinstance(a, torch.Tensor):
        return a.to(device=device, dtype=torch.int32).contiguous()
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.uint32).view(np.int32)).to(device)


class BidiGraph:
    """CSR + reverse CSR resident in HBM: uint64 offsets, uint32 ids.

    ``fw_off[n+1]``/``fw_adj[m]`` hold out-edges, ``bw_off``/``bw_adj`` the exact
    transpose (graph/bidi_graph_part.hpp:12-27 with KASPAN_INDEX64).
    """

    def __init__(self, n, fw_off, fw_adj, bw_off=None, bw_adj=None, device="cuda", ctx: Context | None = None):
        _require_gpu()
        self.n = int(n)
        self.device = torch.device(device)
        self.fw_off = _as_u64_tensor(fw_off, self.device)
        self.fw_adj = _as_u32_tensor(fw_adj, self.device)
        self.m = int(self.fw_adj.numel())
        if self.fw_off.numel() != self.n + 1:
            raise KspnError(-1, "fw_off must have n+1 entries")
        if bw_off is None:
            bw_off, bw_adj = backward_complement_graph(self.n, self.fw_off, self.fw_adj, ctx=ctx)
        self.bw_off = _as_u64_tensor(bw_off, self.device)
        self.bw_adj = _as_u32_tensor(bw_adj, self.device)
        if self.bw_off.numel() != self.n + 1 or self.bw_adj.numel() != self.m:
            raise KspnError(-1, "bw CSR does not match the fw CSR's shape")

    @classmethod
    def from_host(cls, n, fw_off, fw_adj, bw_off=None, bw_adj=None, device="cuda"):
        return cls(n, fw_off, fw_adj, bw_off, bw_adj, device=device)

    @classmethod
    def from_edges(cls, n, src, dst, device="cuda", ctx: Context | None = None, stream=None):
        """Build both CSRs on the GPU from an edge list; self-loops and duplicates removed."""
        ctx = ctx or Context.default(torch.device(device).index or 0)
        s = _as_u32_tensor(src, device)
        d = _as_u32_tensor(dst, device)
        m = C.c_uint64()
        ptrs = [C.c_void_p() for _ in range(4)]
        _lib.check(_lib.lib().kspn_build_csr(ctx.ptr, n, s.numel(), C.c_void_p(s.data_ptr()), C.c_void_p(d.data_ptr()),
                                             C.byref(m), *(C.byref(p) for p in ptrs), _stream_ptr(stream)),
                   "kspn_build_csr")
        return cls._adopt(n, m.value, ptrs, device)

    @classmethod
    def _adopt(cls, n, m, ptrs, device):
        """copy library-allocated CSR arrays into torch tensors and free them"""
        torch.cuda.synchronize()
        fo = torch.empty(n + 1, dtype=torch.int64, device=device)
        fa = torch.empty(max(m, 0), dtype=torch.int32, device=device)
        bo = torch.empty(n + 1, dtype=torch.int64, device=device)
        ba = torch.empty(max(m, 0), dtype=torch.int32, device=device)
        try:
            _lib.memcpy_d2d(fo.data_ptr(), ptrs[0].value, 8 * (n + 1))
            _lib.memcpy_d2d(fa.data_ptr(), ptrs[1].value, 4 * m)
            _lib.memcpy_d2d(bo.data_ptr(), ptrs[2].value, 8 * (n + 1))
            _lib.memcpy_d2d(ba.data_ptr(), ptrs[3].value, 4 * m)
        finally:
            for p in ptrs:
                _lib.lib().kspn_free(p)
        g = cls.__new__(cls)
        g.n, g.m, g.device = int(n), int(m), torch.device(device)
        g.fw_off, g.fw_adj, g.bw_off, g.bw_adj = fo, fa, bo, ba
        return g

    def view(self) -> _lib.Graph:
        return _lib.Graph(self.n, self.m, self.fw_off.data_ptr(), self.fw_adj.data_ptr() if self.m else None,
                          self.bw_off.data_ptr(), self.bw_adj.data_ptr() if self.m else None)

    def to_host(self):
        """(fw_off, fw_adj, bw_off, bw_adj) as numpy uint64/uint32"""
        return (self.fw_off.cpu().numpy().view(np.uint64), self.fw_adj.cpu().numpy().view(np.uint32),
                self.bw_off.cpu().numpy().view(np.uint64), self.bw_adj.cpu().numpy().view(np.uint32))

    def nbytes(self):
        return 2 * 8 * (self.n + 1) + 2 * 4 * self.m


def generate(kind: str, seed: int, device="cuda", ctx: Context | None = None, stream=None, **kw) -> BidiGraph:
    """Seeded graph built on the GPU; the CPU oracle reproduces the same edge stream.

    kind="gnm": n, draws; kind="rmat": scale, draws, a, b, c, permute;
    kind="lattice": side, p_bidir.
    """
    ctx = ctx or Context.default(torch.device(device).index or 0)
    p = _lib.GenParams()
    p.seed = seed
    if kind == "gnm":
        p.kind, p.p0, p.p1 = _lib.GEN_GNM, kw["n"], kw["draws"]
    elif kind == "rmat":
        p.kind, p.p0, p.p1 = _lib.GEN_RMAT, kw["scale"], kw["draws"]
        p.a, p.b, p.c = kw.get("a", 0.57), kw.get("b", 0.19), kw.get("c", 0.19)
        p.permute = int(kw.get("permute", True))
    elif kind == "lattice":
        p.kind, p.p0, p.a = _lib.GEN_LATTICE, kw["side"], kw.get("p_bidir", 0.1)
    elif kind == "planted":
        p = _planted_params(seed, **kw)
    else:
        raise KspnError(-1, f"unknown generator {kind!r}")
    n = C.c_uint32()
    m = C.c_uint64()
    ptrs = [C.c_void_p() for _ in range(4)]
    _lib.check(_lib.lib().kspn_generate(ctx.ptr, C.byref(p), C.byref(n), C.byref(m), *(C.byref(q) for q in ptrs),
                                        _stream_ptr(stream)), "kspn_generate")
    return BidiGraph._adopt(n.value, m.value, ptrs, device)


def _planted_params(seed, n, max_size=10_000_000, expo=2.1, intra_deg=4.0, inter_deg=3.0):
    p = _lib.GenParams()
    p.seed, p.kind, p.p0, p.p1 = seed, _lib.GEN_PLANTED, n, max_size
    p.a, p.b, p.c = expo, intra_deg, inter_deg
    return p


def planted_labels(seed, n, device="cuda", ctx: Context | None = None, stream=None, **kw) -> torch.Tensor:
    """Canonical labels of generate("planted", seed, n=n, ...), known by construction."""
    ctx = ctx or Context.default(torch.device(device).index or 0)
    p = _planted_params(seed, n, **kw)
    labels = torch.empty(n, dtype=torch.int32, device=device)
    _lib.check(_lib.lib().kspn_planted_labels(ctx.ptr, C.byref(p), C.c_void_p(labels.data_ptr()), _stream_ptr(stream)),
               "kspn_planted_labels")
    return labels


def _opts(method=METHOD_KASPAN, pivot=0, fwbw_rounds=1, residual_threshold=0, max_rounds=0, flags=0, alpha=14.0,
          beta=24.0, mlp_params=None, tail_threshold=0, pivot_log=None, seed=0, coop_edges=0, coop_degree=0,
          coop_workgroups=0, pull_ilp=0, expand_ept=0, sweep_wgs=0, list_wgs=0, chain_levels=0, learned_rounds=0):
    o = _lib.Opts()
    _lib.lib().kspn_opts_default(C.byref(o))
    o.method, o.pivot, o.fwbw_rounds = method, pivot, fwbw_rounds
    o.residual_threshold, o.max_rounds, o.flags = residual_threshold, max_rounds, flags
    o.tail_threshold, o.seed = tail_threshold, seed
    o.coop_edges, o.coop_degree, o.coop_workgroups = coop_edges, coop_degree, coop_workgroups
    o.pull_ilp, o.expand_ept, o.sweep_wgs, o.list_wgs = pull_ilp, expand_ept, sweep_wgs, list_wgs
    o.chain_levels = chain_levels
    o.learned_rounds = learned_rounds
    if pivot_log is not None:
        o.pivot_log = pivot_log.data_ptr()
        o.pivot_log_cap = pivot_log.numel() // _lib.LOG_WORDS
    o.alpha, o.beta = alpha, beta
    o.mlp_params = mlp_params.data_ptr() if mlp_params is not None else None
    return o


def scc(graph: BidiGraph, scc_id: torch.Tensor | None = None, *, ctx: Context | None = None, stream=None,
        profile=False, stats=True, **opts) -> tuple[torch.Tensor, Stats | None]:
    """kaspan::scc(graph, scc_id) (scc/scc.hpp:21-87) on the GPU.

    Fills ``scc_id`` (int32 device tensor of n entries, allocated if None)
    with canonical ids and returns ``(scc_id, stats)``. Keyword options map
    to ``kspn_opts`` (method, pivot, fwbw_rounds, residual_threshold,
    max_rounds, flags, alpha, beta, mlp_params). ``stats=False`` skips the
    statistics (no per-scope HIP events) and returns None for them.
    """
    ctx = ctx or Context.default(graph.device.index or 0)
    if scc_id is None:
        scc_id = torch.empty(graph.n, dtype=torch.int32, device=graph.device)
    if scc_id.numel() != graph.n or scc_id.dtype != torch.int32 or not scc_id.is_contiguous():
        raise KspnError(-1, "scc_id must be a contiguous int32 tensor with n entries")
    o = _opts(**opts)
    if profile:
        o.flags |= _lib.FLAG_PROFILE
    st = Stats() if (stats or profile) else None
    g = graph.view()
    rc = _lib.lib().kspn_scc(ctx.ptr, C.byref(g), C.c_void_p(scc_id.data_ptr()), C.byref(o),
                             C.byref(st) if st is not None else None, _stream_ptr(stream))
    _lib.check(rc, "kspn_scc")
    return scc_id, st


def scc_trim_ex(graph: BidiGraph, scc_id=None, **kw):
    """scc_trim_ex (scc/variant/scc_trim_ex.hpp:16-100): exhaustive trims around FW-BW and colouring."""
    return scc(graph, scc_id, method=METHOD_TRIM_EX, **kw)


def scc_trim_ex_residual(graph: BidiGraph, scc_id=None, **kw):
    """scc_trim_ex_residual at one rank (scc/variant/scc_trim_ex_residual.hpp:17-128): exhaustive
    trim, then everything left is the residual (the one-rank decided threshold is 0, :47)."""
    return scc(graph, scc_id, method=_lib.METHOD_TRIM_EX_RESIDUAL, **kw)


def scc_trim_ex_light_residual(graph: BidiGraph, scc_id=None, **kw):
    """scc_trim_ex_light_residual at one rank (scc/variant/scc_trim_ex_light_residual.hpp:16-131):
    exhaustive trim, the trim over decisions since (none at one rank), residual."""
    return scc(graph, scc_id, method=_lib.METHOD_TRIM_EX_LIGHT_RESIDUAL, **kw)


def scc_hpc_like(graph: BidiGraph, scc_id=None, **kw):
    """scc_hpc_like (scc/variant/scc_hpc_like.hpp:16-69): trim_1_first, FW-BW, colour rounds."""
    return scc(graph, scc_id, method=_lib.METHOD_HPC_LIKE, **kw)


def scc_hpc_trim_ex(graph: BidiGraph, scc_id=None, **kw):
    """scc_hpc_trim_ex (scc/variant/scc_hpc_trim_ex.hpp:16-76): exhaustive trim, pivot, FW-BW, colour."""
    return scc(graph, scc_id, method=_lib.METHOD_HPC_TRIM_EX, **kw)


def scc_ispan_like(graph: BidiGraph, scc_id=None, **kw):
    """scc_ispan_like (scc/variant/scc_ispan_like.hpp:18-87): trim_1_first, FW-BW, two trim_1_normal
    passes, residual."""
    return scc(graph, scc_id, method=_lib.METHOD_ISPAN_LIKE, **kw)


def scc_variant(variant: str | None, graph: BidiGraph, scc_id=None, **kw):
    """the driver's variant switch (tool/bench_kaspan/src/main.cpp:63-71): ``variant`` is one of
    ``VARIANTS`` (None: kaspan::scc). ``scc_trim_ex`` runs the trim-2 pass unless
    ``flags`` has ``FLAG_NO_TRIM2``; the other variants never do."""
    if variant not in VARIANTS:
        raise KspnError(-1, f"unknown variant {variant!r}; one of {sorted(k for k in VARIANTS if k)}")
    return scc(graph, scc_id, method=VARIANTS[variant], **kw)


def backward_complement_graph(n, fw_off: torch.Tensor, fw_adj: torch.Tensor, ctx: Context | None = None, stream=None):
    """Reverse CSR, ascending sources per row (scc/backward_complement.hpp:22-68)."""
    fw_off = _as_u64_tensor(fw_off, "cuda")
    fw_adj = _as_u32_tensor(fw_adj, "cuda")
    ctx = ctx or Context.default(fw_off.device.index or 0)
    m = fw_adj.numel()
    bw_off = torch.empty(n + 1, dtype=torch.int64, device=fw_off.device)
    bw_adj = torch.empty(m, dtype=torch.int32, device=fw_off.device)
    _lib.check(_lib.lib().kspn_transpose(ctx.ptr, n, m, C.c_void_p(fw_off.data_ptr()),
                                         C.c_void_p(fw_adj.data_ptr() if m else None), C.c_void_p(bw_off.data_ptr()),
                                         C.c_void_p(bw_adj.data_ptr() if m else None), _stream_ptr(stream)),
               "kspn_transpose")
    return bw_off, bw_adj


def tarjan(n, off: torch.Tensor, adj: torch.Tensor, labels: torch.Tensor | None = None, ctx: Context | None = None,
           stream=None):
    """GPU Tarjan (scc/tarjan.hpp:20-136), one wavefront; labels = minimum member id."""
    off = _as_u64_tensor(off, "cuda")
    adj = _as_u32_tensor(adj, "cuda")
    ctx = ctx or Context.default(off.device.index or 0)
    if labels is None:
        labels = torch.empty(n, dtype=torch.int32, device=off.device)
    m = adj.numel()
    _lib.check(_lib.lib().kspn_tarjan(ctx.ptr, n, m, C.c_void_p(off.data_ptr()), C.c_void_p(adj.data_ptr() if m else None),
                                      C.c_void_p(labels.data_ptr()), _stream_ptr(stream)), "kspn_tarjan")
    return labels


def normalise_scc_id(labels: torch.Tensor, ctx: Context | None = None, stream=None):
    """In place: each id becomes the smallest vertex carrying it (test/normalise_scc.hpp:10-25)."""
    ctx = ctx or Context.default(labels.device.index or 0)
    _lib.check(_lib.lib().kspn_canonicalize(ctx.ptr, labels.numel(), C.c_void_p(labels.data_ptr()), _stream_ptr(stream)),
               "kspn_canonicalize")
    return labels


def allgather_fw_sub_graph(n, off: torch.Tensor, adj: torch.Tensor, mask: torch.Tensor, ctx: Context | None = None,
                           stream=None):
    """Induced subgraph on mask != 0 (scc/allgather_sub_graph.hpp:253-301 at one rank).

    Returns (super_ids, sub_off, sub_adj) with sub ids ascending in global id.
    """
    off = _as_u64_tensor(off, "cuda")
    adj = _as_u32_tensor(adj, "cuda")
    ctx = ctx or Context.default(off.device.index or 0)
    mask = mask.to(device=off.device, dtype=torch.uint8).contiguous()
    m = adj.numel()
    super_ids = torch.empty(max(n, 1), dtype=torch.int32, device=off.device)
    sub_off = torch.empty(n + 1, dtype=torch.int64, device=off.device)
    sub_adj = torch.empty(max(m, 1), dtype=torch.int32, device=off.device)
    sn = C.c_uint32()
    sm = C.c_uint64()
    _lib.check(_lib.lib().kspn_subgraph(ctx.ptr, n, C.c_void_p(off.data_ptr()), C.c_void_p(adj.data_ptr() if m else None),
                                        C.c_void_p(mask.data_ptr()), C.c_void_p(super_ids.data_ptr()),
                                        C.c_void_p(sub_off.data_ptr()), C.c_void_p(sub_adj.data_ptr()), C.byref(sn),
                                        C.byref(sm), _stream_ptr(stream)), "kspn_subgraph")
    return super_ids[: sn.value], sub_off[: sn.value + 1], sub_adj[: sm.value]


def scc_sizes(labels: torch.Tensor, ctx: Context | None = None, stream=None) -> torch.Tensor:
    """sizes[v] = size of v's SCC, from canonical labels (the pivot outcome log)."""
    ctx = ctx or Context.default(labels.device.index or 0)
    sizes = torch.empty_like(labels)
    _lib.check(_lib.lib().kspn_scc_sizes(ctx.ptr, labels.numel(), C.c_void_p(labels.data_ptr()),
                                         C.c_void_p(sizes.data_ptr()), _stream_ptr(stream)), "kspn_scc_sizes")
    return sizes


# --------------------------------------------------------------------------- batches
class Batch:
    """Many independent graphs as one union graph with batch-global ids.

    ``graph`` is the union (a BidiGraph); graph i owns vertices
    ``vtx_off[i]:vtx_off[i+1]`` and edge slots ``edge_off[i]:edge_off[i+1]``.
    """

    def __init__(self, graph: BidiGraph, vtx_off: torch.Tensor, edge_off: torch.Tensor):
        self.graph = graph
        self.vtx_off = _as_u32_tensor(vtx_off, graph.device)
        self.edge_off = _as_u64_tensor(edge_off, graph.device)
        self.count = self.vtx_off.numel() - 1
        self._h_vo = self.vtx_off.cpu().numpy().view(np.uint32).astype(np.int64)
        self._h_eo = self.edge_off.cpu().numpy().view(np.uint64).astype(np.int64)
        if self.count < 1 or self._h_vo[-1] != graph.n or self._h_eo[-1] != graph.m:
            raise KspnError(-1, "vtx_off/edge_off do not match the union graph")

    @classmethod
    def from_graphs(cls, graphs) -> "Batch":
        """Pack BidiGraphs (local ids) into one union graph."""
        if not graphs:
            raise KspnError(-1, "empty batch")
        dev = graphs[0].device
        vo, eo = [0], [0]
        fo, fa, bo, ba = [torch.zeros(1, dtype=torch.int64, device=dev)], [], [torch.zeros(1, dtype=torch.int64, device=dev)], []
        for g in graphs:
            fo.append(g.fw_off[1:] + eo[-1])
            bo.append(g.bw_off[1:] + eo[-1])
            fa.append(g.fw_adj + vo[-1])
            ba.append(g.bw_adj + vo[-1])
            vo.append(vo[-1] + g.n)
            eo.append(eo[-1] + g.m)
        u = BidiGraph.__new__(BidiGraph)
        u.n, u.m, u.device = vo[-1], eo[-1], dev
        u.fw_off, u.bw_off = torch.cat(fo), torch.cat(bo)
        u.fw_adj = torch.cat(fa) if fa else torch.zeros(0, dtype=torch.int32, device=dev)
        u.bw_adj = torch.cat(ba) if ba else torch.zeros(0, dtype=torch.int32, device=dev)
        return cls(u, torch.tensor(np.array(vo, np.uint32).view(np.int32)), torch.tensor(np.array(eo, np.int64)))

    def n_of(self, i: int) -> int:
        return int(self._h_vo[i + 1] - self._h_vo[i])

    def m_of(self, i: int) -> int:
        return int(self._h_eo[i + 1] - self._h_eo[i])

    def subgraph(self, i: int) -> BidiGraph:
        """graph i alone, with local ids (copies)"""
        v0, v1 = int(self._h_vo[i]), int(self._h_vo[i + 1])
        e0, e1 = int(self._h_eo[i]), int(self._h_eo[i + 1])
        u = self.graph
        g = BidiGraph.__new__(BidiGraph)
        g.n, g.m, g.device = v1 - v0, e1 - e0, u.device
        g.fw_off = u.fw_off[v0 : v1 + 1] - e0
        g.bw_off = u.bw_off[v0 : v1 + 1] - e0
        g.fw_adj = u.fw_adj[e0:e1] - v0
        g.bw_adj = u.bw_adj[e0:e1] - v0
        return g

    def view(self) -> _lib.Batch:
        u = self.graph
        return _lib.Batch(self.count, u.n, u.m, self.vtx_off.data_ptr(), self.edge_off.data_ptr(), u.fw_off.data_ptr(),
                          u.fw_adj.data_ptr() if u.m else None, u.bw_off.data_ptr(), u.bw_adj.data_ptr() if u.m else None)


def _batch_gen(seed, count, n_min=1000, n_max=1_000_000, deg_min=2.0, deg_max=16.0, p_rmat=0.5, a=0.57, b=0.19, c=0.19):
    p = _lib.BatchGen()
    p.seed, p.count, p.n_min, p.n_max = seed, count, n_min, n_max
    p.deg_min, p.deg_max, p.p_rmat, p.a, p.b, p.c = deg_min, deg_max, p_rmat, a, b, c
    return p


def batch_plan(seed, count, **kw):
    """(kind, n, draws, seed) per graph of a generated batch (numpy arrays)."""
    p = _batch_gen(seed, count, **kw)
    kind = np.zeros(count, np.uint32)
    n = np.zeros(count, np.uint32)
    draws = np.zeros(count, np.uint64)
    gseed = np.zeros(count, np.uint64)
    ptr = lambda a: C.c_void_p(a.ctypes.data)
    _lib.check(_lib.lib().kspn_batch_plan(C.byref(p), ptr(kind), ptr(n), ptr(draws), ptr(gseed)), "kspn_batch_plan")
    return kind, n, draws, gseed


def generate_batch(seed, count, device="cuda", ctx: Context | None = None, stream=None, **kw) -> Batch:
    """Seeded batch: graph i is G(n,m) or R-MAT (probability p_rmat) with n log-uniform in
    [n_min, n_max] and average out-degree log-uniform in [deg_min, deg_max]."""
    ctx = ctx or Context.default(torch.device(device).index or 0)
    p = _batch_gen(seed, count, **kw)
    b = _lib.Batch()
    _lib.check(_lib.lib().kspn_generate_batch(ctx.ptr, C.byref(p), C.byref(b), _stream_ptr(stream)),
               "kspn_generate_batch")
    union = BidiGraph._adopt(b.n, b.m, [C.c_void_p(b.fw_off), C.c_void_p(b.fw_adj), C.c_void_p(b.bw_off),
                                        C.c_void_p(b.bw_adj)], device)
    vo = torch.empty(count + 1, dtype=torch.int32, device=device)
    eo = torch.empty(count + 1, dtype=torch.int64, device=device)
    try:
        _lib.memcpy_d2d(vo.data_ptr(), b.vtx_off, 4 * (count + 1))
        _lib.memcpy_d2d(eo.data_ptr(), b.edge_off, 8 * (count + 1))
    finally:
        _lib.lib().kspn_free(C.c_void_p(b.vtx_off))
        _lib.lib().kspn_free(C.c_void_p(b.edge_off))
    return Batch(union, vo, eo)


def scc_batch(batch: Batch, scc_id: torch.Tensor | None = None, *, ctx: Context | None = None, stream=None,
              profile=False, stats=True, **opts) -> tuple[torch.Tensor, Stats | None]:
    """scc() for every graph of the batch in one solve.

    ``scc_id[vtx_off[i] + v]`` is v's canonical id inside graph i (local ids),
    what ``scc(batch.subgraph(i))`` returns.
    """
    u = batch.graph
    ctx = ctx or Context.default(u.device.index or 0)
    if scc_id is None:
        scc_id = torch.empty(u.n, dtype=torch.int32, device=u.device)
    if scc_id.numel() != u.n or scc_id.dtype != torch.int32 or not scc_id.is_contiguous():
        raise KspnError(-1, "scc_id must be a contiguous int32 tensor with one entry per batch vertex")
    o = _opts(**opts)
    if profile:
        o.flags |= _lib.FLAG_PROFILE
    st = Stats() if (stats or profile) else None
    b = batch.view()
    _lib.check(_lib.lib().kspn_scc_batch(ctx.ptr, C.byref(b), C.c_void_p(scc_id.data_ptr()), C.byref(o),
                                         C.byref(st) if st is not None else None, _stream_ptr(stream)),
               "kspn_scc_batch")
    return scc_id, st
