"""The autotuner over a small option space: every candidate gives the same
labels (the tuner checks and raises otherwise), and the tuned options
reproduce Tarjan's labels."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_autotune_small_space(ka, O):
    from kaspan_amd import autotune

    g = ka.generate("rmat", 41, scale=14, draws=16 << 14)
    space = {"alpha": [14.0, 1e-9], "coop_edges": [1 << 20, 1], "fwbw_rounds": [1, 4], "method": [0, 1]}
    best, ms, trials = autotune.tune(g, space=space, reps=2)
    assert ms > 0 and len(trials) == 1 + 4
    assert set(best) >= set(space)
    lab, _ = ka.scc(g, **best)
    fo, fa, bo, ba = g.to_host()
    ref = O.tarjan(O.HostGraph(g.n, fo, fa, bo, ba))
    assert np.array_equal(lab.cpu().numpy().view(np.uint32), ref)


def test_autotune_batch(ka):
    from kaspan_amd import autotune

    b = ka.generate_batch(5, 16, n_min=100, n_max=5000, deg_min=1.0, deg_max=6.0)
    best, ms, trials = autotune.tune(b, space={"alpha": [14.0, 1e-9], "fwbw_rounds": [1, 3]}, reps=1)
    assert len(trials) == 3 and ms > 0
