"""The benchmark workloads: seeded synthetic graphs of the five configurations
(and manifest files), each with its solve call. Shared by bench.py and the
autotuner so both measure the same inputs."""
from __future__ import annotations

CONFIGS = {
    # name: (generator kind, kwargs, seed, description)
    "gnm64k": ("gnm", dict(n=65536, draws=524288), 0x6B5350AA00000000,
               "G(n,m) n=65536 m=524288 draws, self-loops/duplicates removed"),
    "rmat24": ("rmat", dict(scale=24, draws=16 << 24, a=0.57, b=0.19, c=0.19, permute=True), 0x6B5350AA00000001,
               "R-MAT scale 24, edge factor 16, a=.57 b=.19 c=.19 d=.05, permuted ids, self-loops/duplicates removed"),
    "planted50m": ("planted", dict(n=50_000_000, max_size=10_000_000, expo=2.1, intra_deg=4.0, inter_deg=3.0),
                   0x6B5350AA00000002,
                   "planted SCCs, n=5e7, power-law sizes (exponent 2.1, 1..1e7), Hamiltonian cycle + chords to "
                   "out-degree 4 inside, 3 forward inter-component edges per vertex"),
    "lattice256": ("lattice", dict(side=256, p_bidir=0.1), 0x6B5350AA00000003,
                   "3-D lattice 256^3, each edge one random direction, 10% bidirectional"),
    "batch4096": ("batch", dict(count=4096, n_min=1000, n_max=1_000_000, deg_min=2.0, deg_max=16.0, p_rmat=0.5),
                  0x6B5350AA00000004,
                  "4096 independent graphs in one batch: G(n,m) or R-MAT(.57/.19/.19/.05) with p=.5, n log-uniform "
                  "in [1e3, 1e6], average out-degree log-uniform in [2, 16], deduplicated, concatenated CSR with "
                  "per-graph vertex/edge offsets"),
}


class Workload:
    """The graph (or batch) of one configuration, and its solve call."""

    def __init__(self, ka, name, manifest=None):
        self.ka = ka
        self.batch = None
        if manifest:
            from kaspan_amd import manifest as M

            mf = M.parse_manifest(manifest)
            self.name = f"manifest:{mf.name}"
            self.kind, self.kw, self.seed = "manifest", {}, None
            self.desc = (f"manifest graph {mf.name} (n={mf.node_count}, m={mf.edge_count}, {mf.head_bytes}/"
                         f"{mf.csr_bytes}-byte {mf.endian}-endian arrays)")
            self.g = M.load_manifest(mf)
        else:
            self.name = name
            self.kind, self.kw, self.seed, self.desc = CONFIGS[name]
            if self.kind == "batch":
                bkw = dict(self.kw)
                count = bkw.pop("count")
                self.batch = ka.generate_batch(self.seed, count, **bkw)
                self.g = self.batch.graph
            else:
                self.g = ka.generate(self.kind, self.seed, **self.kw)

    def solve(self, lab, **o):
        if self.batch is not None:
            return self.ka.scc_batch(self.batch, lab, **o)
        return self.ka.scc(self.g, lab, **o)
