"""The served-query workload: a pool of graphs of uneven size and a stream of
BC queries with Poisson arrivals.

Pool (32 graphs by default): every fourth graph is a bridged-block chain of
U{64..max_blocks} Erdos-Renyi blocks of 256 vertices (average degree 8, one
bridge edge between consecutive blocks: high diameter, path counts beyond
2^64 past ~64 blocks); the others are R-MAT graphs of scale U{12..20} and
edge factor U{8..32} (a/b/c/d = .57/.19/.19/.05, ids permuted).  A query is
(graph, U{16..1024} distinct sampled sources of degree > 0); arrivals are a
seeded Poisson process.  bench.py's serving run and the serving parity test
draw the same pool and queries from the same seed.
"""
import numpy as np

import bcmi


def make_pool(seed, n_graphs=32, max_blocks=2048):
    """Returns (graphs, specs, rng); rng continues the seeded stream that
    make_queries draws from."""
    rng = np.random.default_rng(seed + 7)
    graphs, specs = [], []
    for i in range(n_graphs):
        if i % 4 == 3:
            nb = int(rng.integers(64, max_blocks + 1))
            graphs.append(bcmi.Graph.block_chain(nb, 256, 8.0, seed=seed + 100 + i))
            specs.append(("chain", nb))
        else:
            sc = int(rng.integers(12, 21))
            ef = int(rng.integers(8, 33))
            graphs.append(bcmi.Graph.rmat(sc, ef, seed=seed + 100 + i))
            specs.append(("rmat", sc, ef))
    return graphs, specs, rng


def make_queries(rng, graphs, seed, n_queries, rate):
    """Returns (graph index per query, sources per query, inter-arrival gaps in s)."""
    gsel = rng.integers(0, len(graphs), n_queries)
    ks = rng.integers(16, 1025, n_queries)
    gaps = rng.exponential(1.0 / rate, n_queries)
    srcs = [graphs[gi].sample_sources(min(int(k), int((graphs[gi].degrees() > 0).sum())), seed=seed + 1000 + i)
            for i, (gi, k) in enumerate(zip(gsel, ks))]
    return [int(x) for x in gsel], srcs, gaps
