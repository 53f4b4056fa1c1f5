"""Writes the golden fixtures under tests/golden/ (run in the build container;
the fixtures are committed, this script is not needed at test time).

- small_graphs.npz: small CSR graphs (random, disconnected, with isolated
  vertices, and a seeded R-MAT scale 8) with BFS depths from
  scipy.sparse.csgraph.shortest_path(unweighted=True) for several roots. scipy
  is an implementation independent of the oracle and of the GPU library.
- hashes.json: SHA-256 of the generator's CSR arrays for fixed seeds and of the
  scale-16 R-MAT depth array for the seeded root (config 1), so any change of
  the generator, root picker or BFS result is caught.
"""
import hashlib
import json
import os
import sys

import numpy as np
import scipy.sparse as sp
from scipy.sparse.csgraph import shortest_path

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from enterprise_amd import HostGraph  # noqa: E402


def scipy_depths(beg_pos, adj, roots):
    n = beg_pos.size - 1
    a = sp.csr_matrix((np.ones(adj.size), adj, beg_pos), shape=(n, n))
    d = shortest_path(a, unweighted=True, directed=True, indices=roots)
    return np.where(np.isinf(d), -1, d).astype(np.int32)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def main():
    rng = np.random.default_rng(20261202)
    cases = {}
    # Erdos-Renyi style graphs of varying density.
    for name, n, ne in [("er_sparse", 300, 280), ("er_mid", 300, 900), ("er_dense", 120, 3000)]:
        src = rng.integers(0, n, ne)
        dst = rng.integers(0, n, ne)
        cases[name] = HostGraph.from_edges(n, src, dst)
    # Two components plus isolated vertices.
    src = np.r_[np.arange(0, 49), np.arange(60, 99)]
    dst = np.r_[np.arange(1, 50), np.arange(61, 100)]
    cases["two_paths_isolated"] = HostGraph.from_edges(110, src, dst)
    cases["rmat8"] = HostGraph.rmat(8, 16, seed=8)
    arrays = {}
    for name, g in cases.items():
        deg = g.degrees()
        nz = np.flatnonzero(deg > 0)
        roots = np.unique(np.r_[nz[:2], nz[-1:], rng.choice(nz, 3, replace=False)]).astype(np.int32)
        arrays[f"{name}__beg_pos"] = g.beg_pos.copy()
        arrays[f"{name}__adj"] = g.adj.copy()
        arrays[f"{name}__roots"] = roots
        arrays[f"{name}__depth"] = scipy_depths(g.beg_pos, g.adj, roots)
    np.savez_compressed(os.path.join(HERE, "small_graphs.npz"), **arrays)

    hashes = {}
    for scale, seed in [(10, 3), (16, 16)]:
        g = HostGraph.rmat(scale, 16, seed=seed)
        hashes[f"rmat{scale}_seed{seed}"] = {"n": g.n, "m": g.m, "beg_pos": sha(g.beg_pos),
                                             "adj": sha(g.adj)}
    g = HostGraph.rmat(16, 16, seed=16)
    root = int(g.pick_roots(1, 1001)[0])
    d = scipy_depths(g.beg_pos, g.adj, [root])[0]
    hashes["config1"] = {"root": root, "depth": sha(d), "levels": int(d.max()) + 1,
                         "reached": int((d >= 0).sum())}
    gg = HostGraph.grid(64, 64, delete_ppm=20000, seed=4096, keep_lcc=True)
    hashes["grid64_lcc"] = {"n": gg.n, "m": gg.m, "beg_pos": sha(gg.beg_pos), "adj": sha(gg.adj)}
    with open(os.path.join(HERE, "hashes.json"), "w") as f:
        json.dump(hashes, f, indent=1, sort_keys=True)
    print("wrote", sorted(arrays)[:3], "...", list(hashes))


if __name__ == "__main__":
    main()
