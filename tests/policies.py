"""Training of the learned scheduler's policies for the GPU tests (not a test module).

Each policy is trained on the GPU from the plain engine's own traced runs on graphs other than
the ones it is then run on.  Results never depend on a policy's choices; the tests use these
to run every configuration with the variant it names and to measure what the policy gains.
"""
import numpy as np

import medusa_amd as md
from medusa_amd import workloads as wl


def direction_samples(G, src, runs=1):
    """Per-superstep (record, push µs, pull µs) from push-only and pull-only traced runs (the
    median of `runs` runs each); BFS visits the same frontier sequence in either direction."""
    tp, tl = [], []
    for _ in range(runs):
        G.bfs(src, direction=md.DIR_PUSH, timing=1, trace=1)
        tp.append(G.trace())
        G.bfs(src, direction=md.DIR_PULL, timing=1, trace=1)
        tl.append(G.trace())
    assert len(tp[0]) == len(tl[0])
    assert np.array_equal(tp[0][:, :7], tl[0][:, :7])
    push_us = np.median(np.stack([t[:, 8] for t in tp]), 0) * 1e3
    pull_us = np.median(np.stack([t[:, 8] for t in tl]), 0) * 1e3
    return tp[0], push_us, pull_us


def direction_rows(records, action):
    return np.stack([md.policy_features(md.POLICY_DIRECTION, r, action) for r in records])


def train_direction_policy():
    """BFS push/pull cost model from RMAT scale 14-17, Erdos-Renyi and two Chung-Lu graphs."""
    Xs, ys = [], []
    graphs = [wl.rmat_graph(scale=s, edge_factor=16, seed=10 + s) for s in (14, 15, 16, 17)]
    graphs += [wl.er_graph(2 ** 15, 2 ** 19, seed=21), wl.chung_lu_graph(200_000, 420_000, seed=22),
               wl.chung_lu_graph(1_200_000, 2_500_000, seed=23)]
    for rp, col in graphs:
        with md.Graph(rp, col) as G:
            for src in (wl.max_outdeg_vertex(rp), 3):
                rec, push_us, pull_us = direction_samples(G, src)
                for a, t in ((0, push_us), (1, pull_us)):
                    Xs.append(direction_rows(rec, a))
                    ys.append(np.log2(1.0 + t))
    X = np.concatenate(Xs).astype(np.float32)
    y = np.concatenate(ys).astype(np.float32)
    pol = md.Policy(md.POLICY_DIRECTION, seed=7)
    pol.train(X, y, iters=3000, lr=3e-3)
    return pol


def train_mbfs_policy():
    """64-source mask engine push/pull cost model from its own traces on RMAT scale 14-16."""
    Xs, ys = [], []
    for sc in (14, 15, 16):
        rp, col = wl.rmat_graph(scale=sc, edge_factor=16, seed=10 + sc)
        with md.Graph(rp, col) as H:
            srcs = wl.mbfs_sources(rp, 64, seed=4)
            for act, d in ((0, md.DIR_PUSH), (1, md.DIR_PULL)):
                H.mbfs(srcs, direction=d, timing=1, trace=1)
                tr = H.trace()
                Xs.append(np.stack([md.policy_features(md.POLICY_DIRECTION, r, act) for r in tr]))
                ys.append(np.log2(1.0 + tr[:, 8] * 1e3))
    pol = md.Policy(md.POLICY_DIRECTION, seed=9)
    pol.train(np.concatenate(Xs).astype(np.float32), np.concatenate(ys).astype(np.float32), iters=2000, lr=3e-3)
    return pol


def variant_samples(G, steps):
    """Per-superstep feature rows, log-cost targets and costs (µs) of each PageRank variant,
    from one traced run per variant (every variant runs the same supersteps)."""
    rows, ys, cost = [], [], {}
    for v in (md.PR_PULL, md.PR_PUSH, md.PR_BINNED):
        G.pagerank(0.85, steps, variant=v)  # layouts built outside the traced run
        G.pagerank(0.85, steps, variant=v, trace=1)
        tr = G.trace()
        assert len(tr) == steps and np.all(tr[:, 7] == v)
        cost[v] = tr[:, 8] * 1e3
        for r in tr:
            rows.append(md.policy_features(md.POLICY_VARIANT, r, r[7]))
            ys.append(np.log2(1.0 + r[8] * 1e3))
    return rows, ys, cost


def variant_training_graphs():
    g = [wl.rmat_graph(scale=sc, edge_factor=16, seed=60 + sc) for sc in (11, 12, 13, 14, 15, 17, 18, 19)]
    g += [wl.rmat_graph(scale=17, edge_factor=4, seed=61), wl.rmat_graph(scale=15, edge_factor=64, seed=62),
          wl.er_graph(2 ** 14, 2 ** 18, seed=63), wl.er_graph(2 ** 17, 2 ** 21, seed=64),
          wl.chung_lu_graph(100_000, 210_000, seed=65), wl.chung_lu_graph(400_000, 840_000, seed=66)]
    return g


def variant_training_set():
    Xs, ys = [], []
    for rp, col in variant_training_graphs():
        with md.Graph(rp, col) as G:
            x, y, _ = variant_samples(G, 6)
            Xs += x
            ys += y
    return np.array(Xs, np.float32), np.array(ys, np.float32)


def train_variant_policy(X=None, y=None):
    """PageRank kernel-variant cost model from the variant training graphs."""
    if X is None:
        X, y = variant_training_set()
    pol = md.Policy(md.POLICY_VARIANT, seed=4)
    pol.train(X, y, iters=3000, lr=3e-3)
    return pol


def delta_samples(G, src, repeats=2):
    """One training row per fixed-width run (features at the run's start, whole-run cost per
    vertex) for every width the policy chooses among, `repeats` runs each."""
    G.sssp(src, trace=1)
    wmean = G.trace()[0, 2]  # the engine's sampled mean weight
    Xs, ys = [], []
    for d in md.delta_candidates(wmean):
        for _ in range(repeats):
            G.sssp(src, delta=d, trace=1)
            x, y = md.delta_run_sample(G.trace())
            Xs.append(x)
            ys.append(y)
    return Xs, ys


def _train_delta(Xs, ys):
    pol = md.Policy(md.POLICY_DELTA, seed=3)
    pol.train(np.array(Xs, np.float32), np.array(ys, np.float32), iters=3000, lr=3e-3)
    return pol


def train_delta_policy_rmat(scale=20, seeds=(41, 42, 43)):
    """SSSP bucket-width cost model from RMAT graphs of the given scale (other seeds than any
    graph it is then run on), weights U[1,1000]."""
    Xs, ys = [], []
    for sd in seeds:
        rp, col = wl.rmat_graph(scale=scale, edge_factor=16, seed=sd)
        w = wl.weighted(rp, col, seed=sd + 100)
        with md.Graph(rp, col, w) as G:
            x, y = delta_samples(G, wl.max_outdeg_vertex(rp))
            Xs += x
            ys += y
    return _train_delta(Xs, ys)


def train_delta_policy():
    """SSSP bucket-width cost model from three smaller lattices."""
    Xs, ys = [], []
    for rows, cols, seed in ((200, 200, 31), (400, 300, 32), (500, 500, 33)):
        rp, col, w = wl.lattice_graph(rows, cols, seed=seed)
        with md.Graph(rp, col, w) as G:
            x, y = delta_samples(G, wl.lattice_source(rows, cols))
            Xs += x
            ys += y
    return _train_delta(Xs, ys)
