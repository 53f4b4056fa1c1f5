"""Small test graphs with closed-form BFS depths (test helpers)."""
import itertools

import numpy as np

from enterprise_amd import HostGraph


def edges_graph(n, edges):
    if len(edges) == 0:
        return HostGraph.from_edges(n, np.zeros(0, np.int32), np.zeros(0, np.int32))
    e = np.asarray(edges, dtype=np.int64)
    return HostGraph.from_edges(n, e[:, 0], e[:, 1])


def path(n):
    return edges_graph(n, [(i, i + 1) for i in range(n - 1)]), lambda r: np.abs(np.arange(n) - r)


def cycle(n):
    def depth(r):
        d = np.abs(np.arange(n) - r)
        return np.minimum(d, n - d)
    return edges_graph(n, [(i, (i + 1) % n) for i in range(n)]), depth


def star(n):
    g = edges_graph(n, [(0, i) for i in range(1, n)])

    def depth(r):
        d = np.full(n, 2 if r != 0 else 1)
        d[0] = 1 if r != 0 else 0
        d[r] = 0
        return d
    return g, depth


def complete(n):
    g = edges_graph(n, list(itertools.combinations(range(n), 2)))

    def depth(r):
        d = np.ones(n, dtype=np.int64)
        d[r] = 0
        return d
    return g, depth


def hypercube(k):
    n = 1 << k
    edges = [(v, v ^ (1 << b)) for v in range(n) for b in range(k) if v < v ^ (1 << b)]
    popcount = np.vectorize(lambda x: bin(int(x)).count("1"))
    return edges_graph(n, edges), lambda r: popcount(np.arange(n) ^ r)


def binary_tree(levels):
    n = (1 << levels) - 1
    edges = [(v, (v - 1) // 2) for v in range(1, n)]

    def anc(v):
        out = []
        while True:
            out.append(v)
            if v == 0:
                return out
            v = (v - 1) // 2

    def depth(r):
        ar = anc(r)
        pos = {a: i for i, a in enumerate(ar)}
        d = np.empty(n, dtype=np.int64)
        for v in range(n):
            for i, a in enumerate(anc(v)):
                if a in pos:
                    d[v] = i + pos[a]
                    break
        return d
    return edges_graph(n, edges), depth


def grid(w, h):
    g = HostGraph.grid(w, h, delete_ppm=0, seed=0, keep_lcc=False)

    def depth(r):
        x0, y0 = r % w, r // w
        v = np.arange(w * h)
        return np.abs(v % w - x0) + np.abs(v // w - y0)
    return g, depth


def disjoint_paths(n1, n2):
    n = n1 + n2
    edges = [(i, i + 1) for i in range(n1 - 1)] + [(n1 + i, n1 + i + 1) for i in range(n2 - 1)]

    def depth(r):
        d = np.full(n, -1, dtype=np.int64)
        if r < n1:
            d[:n1] = np.abs(np.arange(n1) - r)
        else:
            d[n1:] = np.abs(np.arange(n1, n) - r)
        return d
    return edges_graph(n, edges), depth


def directed_with_sinks(n_src, n_sink, deg, seed):
    """Directed CSR: vertices 0..n_src-1 each have `deg` distinct out-edges
    to random targets over all vertices (no loops); the n_sink vertices after
    them have in-edges only."""
    rng = np.random.default_rng(seed)
    n = n_src + n_sink
    rows = []
    for v in range(n_src):
        t = rng.choice(n - 1, size=deg, replace=False)
        t[t >= v] += 1
        rows.append(np.sort(t).astype(np.int32))
    beg = np.zeros(n + 1, dtype=np.int64)
    beg[1:n_src + 1] = np.cumsum([len(r) for r in rows])
    beg[n_src + 1:] = beg[n_src]
    return beg, np.concatenate(rows)


CLOSED_FORMS = {
    "path300": lambda: path(300),
    "cycle257": lambda: cycle(257),
    "star5000": lambda: star(5000),
    "complete40": lambda: complete(40),
    "hypercube10": lambda: hypercube(10),
    "bintree11": lambda: binary_tree(11),
    "grid37x23": lambda: grid(37, 23),
    "disjoint": lambda: disjoint_paths(100, 77),
}
