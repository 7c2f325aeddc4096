"""kspn_tarjan_noalloc: the allocation-free variant of kspn_tarjan (work
area in a fixed device array of the library) gives the same labels as
kspn_tarjan and as the oracle's Tarjan, on two calls in a row, the second
with a smaller graph, and rejects graphs above its maximum size."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def noalloc_labels(ka, g, stream=None):
    import torch

    from kaspan_amd import _lib

    lab = torch.full((g.n,), -7, dtype=torch.int32, device="cuda")
    rc = _lib.lib().kspn_tarjan_noalloc(g.n, g.m, C.c_void_p(g.fw_off.data_ptr()),
                                        C.c_void_p(g.fw_adj.data_ptr()) if g.m else None,
                                        C.c_void_p(lab.data_ptr()), stream)
    _lib.check(rc, "kspn_tarjan_noalloc")
    torch.cuda.synchronize()
    return lab.cpu().numpy().view(np.uint32)


def alloc_labels(ka, g):
    lab = ka.tarjan(g.n, g.fw_off, g.fw_adj)
    return lab.cpu().numpy().view(np.uint32)


def host(O, g):
    fo, fa, bo, ba = g.to_host()
    return O.HostGraph(g.n, fo, fa, bo, ba)


def test_noalloc_tarjan_matches_kspn_tarjan(ka, O):
    big = ka.generate("rmat", 41, scale=14, draws=8 << 14)      # 16384 vertices
    small = ka.generate("gnm", 42, n=3000, draws=6000)          # smaller, fewer edges: stale words stay behind
    for g in (big, small):
        got = noalloc_labels(ka, g)
        assert np.array_equal(got, alloc_labels(ka, g))
        assert np.array_equal(got, O.tarjan(host(O, g)))


def test_noalloc_tarjan_size_limit(ka):
    import torch

    from kaspan_amd import _lib

    n = _lib.TARJAN_NOALLOC_MAX_N + 1
    off = torch.zeros(n + 1, dtype=torch.int64, device="cuda")
    lab = torch.empty(n, dtype=torch.int32, device="cuda")
    rc = _lib.lib().kspn_tarjan_noalloc(n, 0, C.c_void_p(off.data_ptr()), None, C.c_void_p(lab.data_ptr()), None)
    assert rc == -1
    assert b"KSPN_TARJAN_NOALLOC_MAX_N" in _lib.lib().kspn_last_error()
