"""Writes tests/golden/generators.json: for small seeded configurations of
each generator, m after dedup, the first edges and a checksum of the CSR
as produced by the CPU oracle. The GPU generators must reproduce them
bit-for-bit; the CPU tests re-derive them to catch generator drift.

Run: python tests/golden/make_generator_fixtures.py
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import oracle as O  # noqa: E402

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
from helpers import csr_digest  # noqa: E402

CASES = [
    {"kind": "gnm", "seed": 0x6B5350AA00000000, "n": 4096, "draws": 32768},
    {"kind": "gnm", "seed": 7, "n": 1000, "draws": 3000},
    {"kind": "rmat", "seed": 0x6B5350AA00000001, "scale": 12, "draws": 16 << 12, "permute": True},
    {"kind": "rmat", "seed": 11, "scale": 10, "draws": 8 << 10, "permute": False},
    {"kind": "lattice", "seed": 0x6B5350AA00000003, "side": 12, "p_bidir": 0.1},
]


def build(case):
    k = case["kind"]
    if k == "gnm":
        s, d = O.gen_gnm(case["seed"], case["n"], case["draws"])
        n = case["n"]
    elif k == "rmat":
        s, d = O.gen_rmat(case["seed"], case["scale"], case["draws"], permute=case["permute"])
        n = 1 << case["scale"]
    else:
        s, d = O.gen_lattice(case["seed"], case["side"], case["p_bidir"])
        n = case["side"] ** 3
    g = O.csr_from_edges(n, s, d)
    return n, s, d, g


def main():
    out = []
    for case in CASES:
        n, s, d, g = build(case)
        out.append(dict(case, n_vertices=n, raw_edges=int(len(s)), m=g.m, first_edges=[[int(a), int(b)] for a, b in zip(s[:8], d[:8])],
                        fw_digest=csr_digest(g.fw_off, g.fw_adj), bw_digest=csr_digest(g.bw_off, g.bw_adj),
                        components=int(len(np.unique(O.tarjan(g))))))
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "generators.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main()
