"""The autotuner end to end on a small graph: tune() with a reduced space on
R-MAT scale 12 returns parameters from that space or the defaults it starts
from; an engine with them
applied (autotune.apply, as bench.py and the CK run commands do) holds them
and matches the CPU oracle."""
import numpy as np
import pytest

import autotune
import bcmi
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def test_tune_reduced_space_then_oracle():
    g = bcmi.Graph.rmat(12, 16, seed=3)
    src = g.sample_sources(640, seed=4)
    space = {"chunk": [256, 512], "grid": [256, 1024], "alpha_batched": [125, 1000], "steps": [4, 8]}
    p, score, pscore = autotune.tune(g, src, batches=4, plain=4, space=space,
                                     plain_space={"grid_plain": [256, 1024]}, passes=1, log=lambda m: None)
    assert score > 0 and pscore > 0
    assert set(p) == set(autotune.APPLIED)
    # a coordinate keeps its starting value (the default) unless a value of the
    # space beats it by more than the minimum gain
    for k, vals in space.items():
        assert p[k] in vals or p[k] == autotune.DEFAULTS[k], k
    assert p["grid_plain"] in (256, 1024, autotune.DEFAULTS["grid_plain"])
    for k in set(autotune.APPLIED) - set(space) - {"grid_plain"}:
        assert p[k] == autotune.DEFAULTS[k], k
    key = autotune.graph_key("rmat", scale=12, edgefactor=16)
    e = bcmi.Engine()
    assert autotune.apply(e, {key: p}, key) == p
    for k in autotune.APPLIED:
        assert e.get_param(k) == p[k], k
    e.upload(g)
    ref = O.brandes(g.off, g.adj, src[:160], nthreads=4)
    for mode in (bcmi.MODE_BATCHED, bcmi.MODE_PLAIN):
        e.reset()
        e.run(src[:160], mode=mode)
        assert bcmi.rel_err(e.bc(), ref) <= 1e-9, mode
    e.close()
