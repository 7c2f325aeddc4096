"""CPU checks of the C ABI: libkspn loads, exports every function the header
declares, and the ctypes structures match the C layout. No GPU calls."""
import ctypes as C
import glob
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "kspn", "scc.h")
HEADERS = sorted(glob.glob(os.path.join(ROOT, "include", "kspn", "*.h")))


def declared_functions():
    text = "".join(open(h).read() for h in HEADERS)
    return sorted(set(re.findall(r"^\s*(?:int|void|const char\*)\s+(kspn_\w+)\s*\(", text, re.M)))


def test_header_declares_the_binding_exports():
    from kaspan_amd import _lib

    assert declared_functions() == sorted(_lib.EXPORTS)


def test_library_exports_every_declared_symbol():
    from kaspan_amd import _lib

    L = _lib.lib()
    for name in declared_functions():
        assert hasattr(L, name), name
    nm = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r" T (kspn_\w+)", nm))
    assert set(declared_functions()) <= exported


def test_struct_layout_matches_header(tmp_path):
    from kaspan_amd import _lib

    src = tmp_path / "layout.c"
    src.write_text(r'''
#include <stdio.h>
#include <stddef.h>
#include "kspn/scc.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu\n", sizeof(kspn_graph), sizeof(kspn_opts), sizeof(kspn_stats),
         sizeof(kspn_gen_params), sizeof(kspn_kern_stat));
  printf("%zu %zu %zu\n", offsetof(kspn_stats, edges_scanned), offsetof(kspn_stats, n_kern), offsetof(kspn_stats, kern));
  printf("%zu %zu\n", offsetof(kspn_opts, mlp_params), offsetof(kspn_gen_params, permute));
  printf("%zu %zu %zu %zu\n", sizeof(kspn_batch), offsetof(kspn_batch, bw_adj), sizeof(kspn_batch_gen),
         offsetof(kspn_batch_gen, c));
  printf("%zu %zu %zu %zu\n", sizeof(kspn_phase), offsetof(kspn_phase, entries), offsetof(kspn_stats, n_phase),
         offsetof(kspn_stats, phase));
  printf("%zu %zu %zu %zu\n", sizeof(kspn_manifest), offsetof(kspn_manifest, csr_bytes), offsetof(kspn_manifest, name),
         offsetof(kspn_manifest, bw_csr_path));
  return 0;
}
''')
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    got = [C.sizeof(_lib.Graph), C.sizeof(_lib.Opts), C.sizeof(_lib.Stats), C.sizeof(_lib.GenParams),
           C.sizeof(_lib.KernStat), _lib.Stats.edges_scanned.offset, _lib.Stats.n_kern.offset, _lib.Stats.kern.offset,
           _lib.Opts.mlp_params.offset, _lib.GenParams.permute.offset, C.sizeof(_lib.Batch), _lib.Batch.bw_adj.offset,
           C.sizeof(_lib.BatchGen), _lib.BatchGen.c.offset,
           C.sizeof(_lib.Phase), _lib.Phase.entries.offset, _lib.Stats.n_phase.offset, _lib.Stats.phase.offset,
           C.sizeof(_lib.Manifest), _lib.Manifest.csr_bytes.offset, _lib.Manifest.name.offset,
           _lib.Manifest.bw_csr_path.offset]
    assert [int(x) for x in out] == got


def test_defaults_and_errors_without_gpu():
    from kaspan_amd import _lib

    L = _lib.lib()
    o = _lib.Opts()
    L.kspn_opts_default(C.byref(o))
    assert (o.method, o.pivot, o.fwbw_rounds, o.residual_threshold, o.max_rounds) == (0, 0, 1, 0, 0)
    assert abs(o.alpha - 14.0) < 1e-6 and abs(o.beta - 24.0) < 1e-6
    assert L.kspn_strerror(-4) == b"iteration cap reached"
    # NULL arguments are rejected before any device work
    assert L.kspn_scc(None, None, None, None, None, None) == -1
    assert L.kspn_ctx_create(None, 0) == -1


def test_ctx_create_fails_cleanly_without_device():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a device is present")
    from kaspan_amd import _lib

    p = C.c_void_p()
    assert _lib.lib().kspn_ctx_create(C.byref(p), 0) == -1
    assert not p.value


def test_batch_plan_matches_oracle(O):
    """The library's host-side batch plan and the oracle's are the same numbers."""
    import kaspan_amd as ka

    for seed, count, kw in [(0x6B5350AA00000004, 4096, {}), (7, 50, dict(n_min=1, n_max=5, deg_min=0.5, deg_max=3.0)),
                            (8, 300, dict(n_min=100, n_max=100, p_rmat=1.0)), (9, 10, dict(p_rmat=0.0))]:
        got = ka.batch_plan(seed, count, **kw)
        ref = O.batch_plan(seed, count, **kw)
        for a, b in zip(got, ref):
            assert np.array_equal(a, b)
    from kaspan_amd import _lib

    bad = _lib.BatchGen()
    bad.count, bad.n_min, bad.n_max, bad.deg_min, bad.deg_max = 4, 10, 5, 1.0, 2.0  # n_min > n_max
    z = np.zeros(4, np.uint64)
    ptr = lambda a: C.c_void_p(a.ctypes.data)
    assert _lib.lib().kspn_batch_plan(C.byref(bad), ptr(z), ptr(z), ptr(z), ptr(z)) == -1
    assert _lib.lib().kspn_scc_batch(None, None, None, None, None, None) == -1
