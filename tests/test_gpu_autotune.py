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


@pytest.mark.parametrize("shape", [dict(pull_ilp=1), dict(pull_ilp=4), dict(expand_ept=4), dict(expand_ept=16),
                                   dict(sweep_wgs=2, list_wgs=1), dict(pull_ilp=4, expand_ept=16, sweep_wgs=16)])
def test_launch_shapes_keep_labels(ka, O, shape):
    """every compiled launch shape gives Tarjan's labels, single graph and multi-pivot rounds"""
    g = ka.generate("rmat", 43, scale=15, draws=16 << 15)
    fo, fa, bo, ba = g.to_host()
    ref = O.tarjan(O.HostGraph(g.n, fo, fa, bo, ba))
    for rounds in (1, 8):
        lab, _ = ka.scc(g, fwbw_rounds=rounds, alpha=4.0, **shape)
        assert np.array_equal(lab.cpu().numpy().view(np.uint32), ref), (shape, rounds)


def test_launch_shapes_rejected(ka):
    g = ka.generate("gnm", 44, n=1000, draws=4000)
    for bad in (dict(pull_ilp=3), dict(expand_ept=32), dict(sweep_wgs=65)):
        with pytest.raises(ka.KspnError):
            ka.scc(g, **bad)
