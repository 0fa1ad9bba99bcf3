"""Deep graphs (hub BFS depth >= 64) run the sweeps of one call on several
workspaces and streams at once; their BC arrays are summed at the end.  The
result must match the oracle and the one-stream run, in both sweep modes and
with peeling."""
import numpy as np
import pytest

import bcmi
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", [bcmi.MODE_BATCHED, bcmi.MODE_PLAIN])
def test_streams_match_oracle_and_single_stream(mode):
    g = bcmi.Graph.block_chain(40, 128, 6.0, seed=11)   # depth ~ 40 blocks x 4 levels
    n_src = 300 if mode == bcmi.MODE_BATCHED else 24
    src = g.sample_sources(n_src, seed=4)
    ref = O.brandes(g.off, g.adj, src, nthreads=4)
    out = {}
    for streams in (1, 4):
        e = bcmi.Engine(streams=streams)
        e.upload(g)
        e.reset()
        st = e.run(src, mode=mode)
        out[streams] = e.bc()
        assert st["sources_run"] == len(src)
        assert bcmi.rel_err(out[streams], ref) <= 1e-9, streams
        # a second call on the same workspaces accumulates on top
        e.run(src[:5], mode=mode)
        ref2 = ref + O.brandes(g.off, g.adj, src[:5], nthreads=4)
        assert bcmi.rel_err(e.bc(), ref2) <= 1e-9
        e.close()
    assert bcmi.rel_err(out[4], out[1]) <= 1e-12


def test_streams_with_peeling_expand():
    g = bcmi.Graph.block_chain(40, 128, 6.0, seed=12).with_pendant_trees(0.3, seed=3)
    src = g.sample_sources(200, seed=8)
    ref = O.bc_pipeline(g.off, g.adj, src, heuristic=1, expand=True, nthreads=4)
    e = bcmi.Engine(streams=4)
    e.upload(g, heuristic=bcmi.H_PEEL)
    e.reset()
    e.run(src, mode=bcmi.MODE_BATCHED, expand=True)
    assert bcmi.rel_err(e.bc(), ref) <= 1e-9
    e.close()
