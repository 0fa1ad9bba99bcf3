"""Small deterministic test graphs (edge lists) shared by the tests."""
import numpy as np


def path(n):
    return np.stack([np.arange(n - 1), np.arange(1, n)], 1), n


def star(k):
    return np.stack([np.zeros(k, int), np.arange(1, k + 1)], 1), k + 1


def cycle(n):
    return np.stack([np.arange(n), (np.arange(n) + 1) % n], 1), n


def complete(n):
    i, j = np.triu_indices(n, 1)
    return np.stack([i, j], 1), n


def grid(a, b):
    e = []
    for i in range(a):
        for j in range(b):
            v = i * b + j
            if j + 1 < b:
                e.append((v, v + 1))
            if i + 1 < a:
                e.append((v, v + b))
    return np.array(e), a * b


def hypercube(d):
    n = 1 << d
    e = [(v, v ^ (1 << k)) for v in range(n) for k in range(d) if v < v ^ (1 << k)]
    return np.array(e), n


def diamonds(k):
    """k diamonds in series: sigma from vertex 0 to the end is 2^k."""
    e = []
    for i in range(k):
        a = 3 * i
        e += [(a, a + 1), (a, a + 2), (a + 1, a + 3), (a + 2, a + 3)]
    return np.array(e), 3 * k + 1


def block_chain(nblocks, block, avgdeg, seed):
    """Bridged-block chain: nblocks random blocks of `block` vertices with
    about block*avgdeg/2 edges each, consecutive blocks joined by one edge
    between random vertices of each (path counts multiply along the chain)."""
    rng = np.random.default_rng(seed)
    m = int(block * avgdeg / 2)
    e = []
    for b in range(nblocks):
        e.append(b * block + rng.integers(0, block, size=(m, 2)))
        if b + 1 < nblocks:
            e.append(np.array([[b * block + rng.integers(0, block), (b + 1) * block + rng.integers(0, block)]]))
    return np.concatenate(e), nblocks * block


def random_graph(n, m, seed):
    rng = np.random.default_rng(seed)
    return rng.integers(0, n, size=(m, 2)), n


def random_forest_plus(n, seed):
    """Random sparse graph with many pendant trees, isolated edges and stars."""
    rng = np.random.default_rng(seed)
    e = []
    core = max(3, n // 3)
    for _ in range(core * 2):
        e.append(tuple(rng.integers(0, core, 2)))
    nxt = core
    while nxt < n - 2:
        kind = rng.integers(0, 4)
        if kind == 0:  # isolated edge
            e.append((nxt, nxt + 1)); nxt += 2
        elif kind == 1:  # star
            k = int(rng.integers(2, 5))
            if nxt + k >= n:
                break
            for t in range(k):
                e.append((nxt, nxt + 1 + t))
            nxt += k + 1
        else:  # tree hanging at a core vertex
            par = [int(rng.integers(0, core))]
            for _ in range(int(rng.integers(1, 6))):
                if nxt >= n:
                    break
                p = par[int(rng.integers(0, len(par)))]
                e.append((p, nxt)); par.append(nxt); nxt += 1
    return np.array(e), n
