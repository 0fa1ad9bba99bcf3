"""Component labels from the lock-free union-find on the device.

Every access to the parent array is atomic, a parent id only ever decreases
(a link is a CAS from a root to a smaller root; path compression stores a
grandparent, an ancestor with a smaller id), so one union pass gives the
final forest and the labels agree with every arc.  The upload still checks
every arc and repeats the union if one disagrees; this test shows that the
repeat is never taken: on 43 graphs (R-MAT of scales 8 to 16 with several
edge factors, random directed graphs, forests with extra edges, block chains,
peeled R-MAT with pendant trees) the labels take one pass and equal the
oracle's components (weak components of the directed graphs: the oracle runs
on their symmetric closure)."""
import numpy as np
import pytest

import bcmi
import graphs as G
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _graphs():
    out = []
    for i, scale in enumerate(range(8, 17)):
        for ef in (1, 4, 16):
            if scale >= 15 and ef == 16:
                continue
            out.append((f"rmat{scale}-ef{ef}", bcmi.Graph.rmat(scale, ef, seed=100 + 7 * i + ef), bcmi.H_NONE, None))
    for seed in range(6):
        e, n = G.random_graph(3000, 2500 + 500 * seed, 40 + seed)
        out.append((f"directed{seed}", bcmi.Graph.from_edges(e, n, directed=True), bcmi.H_NONE,
                    bcmi.Graph.from_edges(e, n)))
        e, n = G.random_forest_plus(4000, 60 + seed)
        out.append((f"forest{seed}", bcmi.Graph.from_edges(e, n), bcmi.H_NONE, None))
    for seed in range(3):
        out.append((f"chain{seed}", bcmi.Graph.block_chain(50 + 40 * seed, 64, 5.0, seed=seed + 3), bcmi.H_NONE, None))
        core = bcmi.Graph.rmat(12, 4, seed=70 + seed)
        out.append((f"peeled{seed}", core.with_pendant_trees(0.4, seed=80 + seed), bcmi.H_PEEL, None))
    return out


def test_labels_take_one_pass():
    gs = _graphs()
    assert len(gs) == 43
    e = bcmi.Engine()
    for name, g, h, sym in gs:
        e.upload(g, heuristic=h)
        assert e.get_param("label_passes") == 1, name
        u = g if sym is None else sym
        lab, _ = O.components(u.off, u.adj)
        assert np.array_equal(e.labels(), lab), name
    e.close()
