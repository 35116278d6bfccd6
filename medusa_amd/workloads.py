"""The synthetic workloads, built by the library's own generators.

Each builder returns host CSR arrays (row_ptr, col, weight-or-None) plus the
run parameters.  Seeds are fixed per workload so every run sees the same graph.
"""
import math

import numpy as np

from . import csr_from_arcs, gen_chung_lu, gen_erdos_renyi, gen_lattice, gen_rmat, gen_weights


def er_graph(n=65_536, m=1_048_576, seed=1):
    """Erdos-Renyi G(n, m): uniform endpoints, self-loops dropped, duplicates kept."""
    src, dst = gen_erdos_renyi(n, m, seed)
    row_ptr, col, _ = csr_from_arcs(n, src, dst)
    return row_ptr, col


def rmat_graph(scale=20, edge_factor=16, seed=2, a=0.45, b=0.15, c=0.15, noise=0.1):
    """R-MAT, quadrant probabilities a/b/c/d with +-noise per level, duplicates and self-loops kept."""
    n = 1 << scale
    src, dst = gen_rmat(scale, edge_factor * n, a, b, c, noise, seed)
    row_ptr, col, _ = csr_from_arcs(n, src, dst)
    return row_ptr, col


def lattice_graph(rows=1400, cols=1400, seed=3, p_keep=0.7, p_diag=0.02, wmin=1, wmax=1000):
    """Road-network-like lattice with diagonal shortcuts, both directions, uint32 weights."""
    src, dst, w = gen_lattice(rows, cols, p_keep, p_diag, wmin, wmax, seed)
    row_ptr, col, wo = csr_from_arcs(rows * cols, src, dst, w)
    return row_ptr, col, wo


def lattice_source(rows=1400, cols=1400):
    return (rows // 2) * cols + cols // 2


def chung_lu_graph(n=2_400_000, m=5_000_000, seed=5, exponent=2.0, max_degree=100_000, p_zero=0.4):
    """Chung-Lu: Zipf(exponent) out-weights capped at max_degree, p_zero of vertices weight 0,
    uniform destinations."""
    src, dst = gen_chung_lu(n, m, exponent, max_degree, p_zero, seed)
    row_ptr, col, _ = csr_from_arcs(n, src, dst)
    return row_ptr, col


def chung_lu_batch(count=128, lo_log2=10, hi_log2=18, edge_factor=8, seed=5):
    """`count` independent Chung-Lu graphs with n log-uniform in [2^lo, 2^hi]; returns a list of
    (row_ptr, col) and the sizes."""
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(lo_log2 * math.log(2), hi_log2 * math.log(2), size=count)).astype(np.int64)
    graphs = []
    for i, n in enumerate(sizes):
        n = int(n)
        src, dst = gen_chung_lu(n, edge_factor * n, 2.0, min(100_000, n - 1), 0.4, seed * 1000 + i)
        graphs.append(csr_from_arcs(n, src, dst)[:2])
    return graphs, sizes


def weighted(row_ptr, col, seed=0, wmin=1, wmax=1000):
    return gen_weights(len(col), wmin, wmax, seed)


def max_outdeg_vertex(row_ptr):
    """Smallest id among vertices of maximum out-degree (the reference's BFS root rule,
    Algorithm/main.cu:209-212)."""
    deg = np.diff(row_ptr.astype(np.int64))
    return int(np.argmax(deg))


def mbfs_sources(row_ptr, k=64, seed=4):
    """k distinct sources uniform among vertices with out-degree >= 1."""
    deg = np.diff(row_ptr.astype(np.int64))
    cand = np.flatnonzero(deg >= 1)
    rng = np.random.default_rng(seed)
    return np.sort(rng.choice(cand, size=min(k, len(cand)), replace=False)).astype(np.uint32)
