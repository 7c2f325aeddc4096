"""Shared test helpers: digests and small hand-made graphs."""
import hashlib

import numpy as np


def csr_digest(off, adj):
    """sha256 over offsets as little-endian uint64 then ids as little-endian uint32"""
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(off, dtype="<u8").tobytes())
    h.update(np.ascontiguousarray(adj, dtype="<u4").tobytes())
    return h.hexdigest()


def csr_from_pairs(n, pairs):
    """CSR in the given edge order (no sorting, no dedup)"""
    off = np.zeros(n + 1, np.uint64)
    for u, _ in pairs:
        off[u + 1] += 1
    off = np.cumsum(off).astype(np.uint64)
    adj = np.zeros(len(pairs), np.uint32)
    pos = off[:-1].copy()
    for u, v in pairs:
        adj[pos[u]] = v
        pos[u] += 1
    return off, adj


# iSpan tiny graph (extern/ispan_test/test_ispan_tiny.cpp:21-90): n=7, m=11,
# SCCs {0,2}, {1,3,4}, {5,6}
ISPAN_TINY_FW_OFF = [0, 1, 4, 5, 7, 9, 10, 11]
ISPAN_TINY_FW_ADJ = [2, 0, 2, 3, 0, 2, 4, 1, 3, 6, 5]
ISPAN_TINY_BW_OFF = [0, 2, 3, 6, 8, 9, 10, 11]
ISPAN_TINY_BW_ADJ = [1, 2, 4, 0, 1, 3, 1, 4, 3, 6, 5]
ISPAN_TINY_LABELS = [0, 1, 0, 1, 1, 5, 5]

# test_mpi/test_allgather_sub_graph.cpp:33-111
SUB_FW_OFF = [0, 2, 5, 6, 6, 9, 10]
SUB_FW_ADJ = [3, 4, 2, 4, 5, 4, 1, 3, 5, 0]
SUB_BW_OFF = [0, 1, 2, 3, 5, 8, 10]
SUB_BW_ADJ = [5, 4, 1, 0, 4, 0, 1, 2, 1, 4]
SUB_MASK = [0, 1, 1, 0, 1, 0]  # vertices {1, 2, 4}
SUB_SUPER_IDS = [1, 2, 4]
SUB_SUB_FW_OFF = [0, 2, 3, 4]
SUB_SUB_FW_ADJ = [1, 2, 2, 0]
SUB_SUB_BW_OFF = [0, 1, 2, 4]
SUB_SUB_BW_ADJ = [2, 0, 0, 1]


def cycle(n):
    return csr_from_pairs(n, [(i, (i + 1) % n) for i in range(n)]) if n > 1 else (np.zeros(n + 1, np.uint64), np.zeros(0, np.uint32))


def forward_dag(n, m, rng):
    pairs = set()
    while len(pairs) < m:
        u, v = sorted(rng.integers(0, n, 2))
        if u != v:
            pairs.add((int(u), int(v)))
    return csr_from_pairs(n, sorted(pairs))


def layered(seed, layers, width, fan=8, back=1):
    """`layers` layers of `width` vertices; each vertex has `fan` edges to
    random vertices of the next layer and `back` edges to the previous one.
    Layer j holds ids [(layers-1-j)*width, (layers-j)*width), so forward
    edges run from high ids to low: a pull sweep, which sees claims made
    earlier in the same sweep only by chance, reaches about one layer per
    sweep and leaves more than n/16 candidates for the next. Returns
    (n, src, dst) as int64 arrays."""
    rng = np.random.default_rng(seed)
    n = layers * width
    src, dst = [], []
    for j in range(layers - 1):
        hi, lo = (layers - 1 - j) * width, (layers - 2 - j) * width
        src.append(np.repeat(np.arange(width, dtype=np.int64) + hi, fan))
        dst.append(rng.integers(0, width, width * fan) + lo)
        src.append(np.repeat(np.arange(width, dtype=np.int64) + lo, back))
        dst.append(rng.integers(0, width, width * back) + hi)
    return n, np.concatenate(src), np.concatenate(dst)
