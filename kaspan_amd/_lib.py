"""ctypes binding of libkspn's C ABI (include/kspn/scc.h).

The shared object is built in-tree (kaspan_amd/libkspn.so). There is no
fallback: if it is missing or cannot be loaded, every entry point raises.
torch is imported first so that libkspn binds to the HIP runtime torch has
already loaded (both carry the SONAME libamdhip64.so.7).
"""
from __future__ import annotations

import ctypes as C
import os

import torch  # noqa: F401  (loads the process's HIP runtime before libkspn)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libkspn.so")

KSPN_OK = 0
ERRORS = {
    -1: "KSPN_E_INVALID",
    -2: "KSPN_E_OOM",
    -3: "KSPN_E_HIP",
    -4: "KSPN_E_ITER_CAP",
    -5: "KSPN_E_UNIMPLEMENTED",
    -6: "KSPN_E_IO",
}
METHOD_KASPAN = 0
METHOD_TRIM_EX = 1
METHOD_HPC_LIKE = 2
METHOD_ISPAN_LIKE = 3
METHOD_HPC_TRIM_EX = 4
METHOD_TRIM_EX_RESIDUAL = 5
METHOD_TRIM_EX_LIGHT_RESIDUAL = 6
PIVOT_PLAIN = 0
PIVOT_LEARNED = 1
PIVOT_RANDOM = 2
LOG_WORDS = 8
FLAG_PROFILE = 1
FLAG_NO_TRIM2 = 2
FLAG_NO_TAIL = 4
FLAG_NO_REFINE = 8
FLAG_NO_COOP = 16
FLAG_STATS_MEMORY = 32
FLAG_FULL_BW = 64
FLAG_GRAPH = 128
FLAG_ASYNC = 256
FLAG_CHAIN = 512
FLAG_PERSISTENT = 1024
FLAG_NARROW = 2048
FLAG_PRUNE = 4096
FLAG_CHUNKED = 8192
PHASE_MAX = 12
MANIFEST_TEXT = 256
MANIFEST_PATH = 4096
GEN_GNM, GEN_RMAT, GEN_LATTICE, GEN_PLANTED = 0, 1, 2, 3
MLP_PARAMS = 1313
UNDECIDED = 0xFFFFFFFF
TARJAN_NOALLOC_MAX_N = 1 << 20

# every symbol include/kspn/*.h declares
EXPORTS = (
    "kspn_opts_default", "kspn_strerror", "kspn_last_error", "kspn_ctx_create", "kspn_ctx_destroy", "kspn_ctx_reserve",
    "kspn_scc", "kspn_scc_host", "kspn_transpose", "kspn_canonicalize", "kspn_subgraph", "kspn_tarjan",
    "kspn_generate", "kspn_build_csr", "kspn_free", "kspn_planted_labels", "kspn_pivot_features",
    "kspn_mlp_forward", "kspn_mlp_grad", "kspn_mlp_adam", "kspn_scc_sizes", "kspn_scc_batch", "kspn_batch_plan",
    "kspn_generate_batch", "kspn_manifest_parse", "kspn_manifest_load", "kspn_manifest_graph", "kspn_unpack_dense",
    "kspn_test_tail_guard", "kspn_tarjan_noalloc",
)


class KspnError(RuntimeError):
    def __init__(self, code, where=""):
        self.code = code
        super().__init__(f"{where}: {ERRORS.get(code, code)} ({code})")


class Graph(C.Structure):
    _fields_ = [
        ("n", C.c_uint32),
        ("m", C.c_uint64),
        ("fw_off", C.c_void_p),
        ("fw_adj", C.c_void_p),
        ("bw_off", C.c_void_p),
        ("bw_adj", C.c_void_p),
    ]


class Batch(C.Structure):
    _fields_ = [
        ("count", C.c_uint32),
        ("n", C.c_uint32),
        ("m", C.c_uint64),
        ("vtx_off", C.c_void_p),
        ("edge_off", C.c_void_p),
        ("fw_off", C.c_void_p),
        ("fw_adj", C.c_void_p),
        ("bw_off", C.c_void_p),
        ("bw_adj", C.c_void_p),
    ]


class BatchGen(C.Structure):
    _fields_ = [
        ("seed", C.c_uint64),
        ("count", C.c_uint32),
        ("n_min", C.c_uint32),
        ("n_max", C.c_uint32),
        ("deg_min", C.c_double),
        ("deg_max", C.c_double),
        ("p_rmat", C.c_double),
        ("a", C.c_double),
        ("b", C.c_double),
        ("c", C.c_double),
    ]


class Opts(C.Structure):
    _fields_ = [
        ("method", C.c_uint32),
        ("pivot", C.c_uint32),
        ("fwbw_rounds", C.c_uint32),
        ("residual_threshold", C.c_uint32),
        ("max_rounds", C.c_uint32),
        ("tail_threshold", C.c_uint32),
        ("flags", C.c_uint32),
        ("alpha", C.c_float),
        ("beta", C.c_float),
        ("mlp_params", C.c_void_p),
        ("pivot_log", C.c_void_p),
        ("pivot_log_cap", C.c_uint32),
        ("seed", C.c_uint32),
        ("coop_edges", C.c_uint32),
        ("coop_degree", C.c_uint32),
        ("coop_workgroups", C.c_uint32),
        ("pull_ilp", C.c_uint32),
        ("expand_ept", C.c_uint32),
        ("sweep_wgs", C.c_uint32),
        ("list_wgs", C.c_uint32), ("chain_levels", C.c_uint32), ("learned_rounds", C.c_uint32),
    ]


class KernStat(C.Structure):
    _fields_ = [("name", C.c_char * 32), ("launches", C.c_uint32), ("ms", C.c_double), ("bytes", C.c_double)]


class Phase(C.Structure):
    _fields_ = [("name", C.c_char * 32), ("ms", C.c_double), ("decided", C.c_uint64), ("memory", C.c_uint64),
                ("entries", C.c_uint32), ("has_memory", C.c_uint32), ("bytes", C.c_double), ("kernel_ms", C.c_double)]


class Stats(C.Structure):
    _fields_ = [
        ("ms_total", C.c_double), ("ms_trim", C.c_double), ("ms_fwbw", C.c_double), ("ms_color", C.c_double),
        ("ms_residual", C.c_double), ("ms_canon", C.c_double),
        ("decided_trim", C.c_uint32), ("decided_trim2", C.c_uint32), ("decided_fwbw", C.c_uint32),
        ("decided_color", C.c_uint32), ("decided_residual", C.c_uint32),
        ("pivot", C.c_uint32),
        ("fwbw_levels_push", C.c_uint32), ("fwbw_levels_pull", C.c_uint32),
        ("color_rounds", C.c_uint32), ("color_iters", C.c_uint32),
        ("residual_n", C.c_uint32),
        ("fwbw_rounds_done", C.c_uint32), ("partitions_max", C.c_uint32),
        ("pivot_log_count", C.c_uint32), ("fwbw_levels_coop", C.c_uint32),
        ("residual_m", C.c_uint64),
        ("edges_scanned", C.c_uint64),
        ("bytes_trim", C.c_double), ("bytes_fwbw", C.c_double), ("bytes_color", C.c_double),
        ("bytes_residual", C.c_double), ("bytes_canon", C.c_double),
        ("n_kern", C.c_uint32),
        ("kern", KernStat * 32),
        ("n_phase", C.c_uint32),
        ("phase", Phase * PHASE_MAX),
    ]

    def phases(self):
        """the statistics scopes in order of first entry"""
        return [{"name": p.name.decode(), "ms": p.ms, "decided": p.decided, "memory": p.memory,
                 "entries": p.entries, "has_memory": bool(p.has_memory), "bytes": p.bytes, "kernel_ms": p.kernel_ms}
                for p in self.phase[: self.n_phase]]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_ if k not in ("kern", "n_kern", "phase", "n_phase")}
        d["phases"] = self.phases()
        d["kernels"] = [
            {"name": self.kern[i].name.decode(), "launches": self.kern[i].launches, "ms": self.kern[i].ms,
             "bytes": self.kern[i].bytes}
            for i in range(self.n_kern)
        ]
        return d


class Manifest(C.Structure):
    _fields_ = [
        ("schema_version", C.c_uint32),
        ("big_endian", C.c_uint32),
        ("node_count", C.c_uint64),
        ("edge_count", C.c_uint64),
        ("contains_self_loops", C.c_uint32),
        ("contains_duplicate_edges", C.c_uint32),
        ("head_bytes", C.c_uint32),
        ("csr_bytes", C.c_uint32),
        ("code", C.c_char * MANIFEST_TEXT),
        ("name", C.c_char * MANIFEST_TEXT),
        ("fw_head_path", C.c_char * MANIFEST_PATH),
        ("fw_csr_path", C.c_char * MANIFEST_PATH),
        ("bw_head_path", C.c_char * MANIFEST_PATH),
        ("bw_csr_path", C.c_char * MANIFEST_PATH),
    ]


class GenParams(C.Structure):
    _fields_ = [
        ("kind", C.c_uint32),
        ("seed", C.c_uint64),
        ("p0", C.c_uint64),
        ("p1", C.c_uint64),
        ("a", C.c_double),
        ("b", C.c_double),
        ("c", C.c_double),
        ("permute", C.c_uint32),
    ]


_lib = None


def lib():
    """Load libkspn.so (raises if it is missing: there is no CPU fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build() (make -C kaspan_amd/csrc)")
    L = C.CDLL(LIB_PATH)
    vp, u32, u64, i32 = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int
    P = C.POINTER
    sig = {
        "kspn_opts_default": (None, [P(Opts)]),
        "kspn_strerror": (C.c_char_p, [i32]),
        "kspn_last_error": (C.c_char_p, []),
        "kspn_ctx_create": (i32, [P(vp), i32]),
        "kspn_ctx_destroy": (i32, [vp]),
        "kspn_ctx_reserve": (i32, [vp, u32, u64]),
        "kspn_scc": (i32, [vp, P(Graph), vp, P(Opts), P(Stats), vp]),
        "kspn_scc_host": (i32, [vp, u32, u64, vp, vp, vp, vp, vp, P(Opts), P(Stats)]),
        "kspn_transpose": (i32, [vp, u32, u64, vp, vp, vp, vp, vp]),
        "kspn_canonicalize": (i32, [vp, u32, vp, vp]),
        "kspn_subgraph": (i32, [vp, u32, vp, vp, vp, vp, vp, vp, P(u32), P(u64), vp]),
        "kspn_tarjan": (i32, [vp, u32, u64, vp, vp, vp, vp]),
        "kspn_generate": (i32, [vp, P(GenParams), P(u32), P(u64), P(vp), P(vp), P(vp), P(vp), vp]),
        "kspn_build_csr": (i32, [vp, u32, u64, vp, vp, P(u64), P(vp), P(vp), P(vp), P(vp), vp]),
        "kspn_free": (i32, [vp]),
        "kspn_planted_labels": (i32, [vp, P(GenParams), vp, vp]),
        "kspn_pivot_features": (i32, [vp, P(Graph), vp, u32, vp, vp]),
        "kspn_mlp_forward": (i32, [vp, vp, vp, u32, vp, vp]),
        "kspn_mlp_grad": (i32, [vp, vp, vp, vp, u32, vp, vp, vp]),
        "kspn_mlp_adam": (i32, [vp, vp, vp, vp, u32, C.c_float, vp]),
        "kspn_scc_sizes": (i32, [vp, u32, vp, vp, vp]),
        "kspn_scc_batch": (i32, [vp, P(Batch), vp, P(Opts), P(Stats), vp]),
        "kspn_batch_plan": (i32, [P(BatchGen), vp, vp, vp, vp]),
        "kspn_generate_batch": (i32, [vp, P(BatchGen), P(Batch), vp]),
        "kspn_manifest_parse": (i32, [C.c_char_p, P(Manifest)]),
        "kspn_manifest_load": (i32, [vp, P(Manifest), vp, vp, vp, vp, vp]),
        "kspn_manifest_graph": (i32, [vp, P(Manifest), P(u32), P(u64), P(vp), P(vp), P(vp), P(vp), vp]),
        "kspn_unpack_dense": (i32, [vp, vp, u64, u32, u32, u64, vp, u32, vp]),
        "kspn_test_tail_guard": (i32, [vp, P(Graph), u32, vp, u32, P(u32), vp]),
        "kspn_tarjan_noalloc": (i32, [u32, u64, vp, vp, vp, vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def check(rc, where, detail=False):
    """raise KspnError for a failing call; the library's last error message is
    appended for HIP and file errors, and for every error when `detail`"""
    if rc != KSPN_OK:
        if rc in (-3, -6) or detail:
            detail = lib().kspn_last_error()
            if detail:
                where = f"{where} [{detail.decode(errors='replace')}]"
        raise KspnError(rc, where)


_hip = None


def hip():
    """The process's HIP runtime (the one torch loaded), for raw copies."""
    global _hip
    if _hip is None:
        h = C.CDLL("libamdhip64.so.7")
        h.hipMemcpy.restype = C.c_int
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip = h
    return _hip


def memcpy_d2d(dst_ptr, src_ptr, nbytes):
    if nbytes:
        rc = hip().hipMemcpy(C.c_void_p(dst_ptr), C.c_void_p(src_ptr), nbytes, 3)
        if rc != 0:
            raise KspnError(-3, f"hipMemcpy (hip error {rc})")
