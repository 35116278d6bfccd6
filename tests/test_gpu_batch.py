"""Packed batches of independent graphs and multi-source BFS.

A batch is one handle whose vertex ranges are disjoint graphs.  Every graph's BFS
levels must equal the oracle's BFS on that graph alone, and every graph's PageRank
must be bit-identical to the oracle's FP32 restatement on that graph alone (its own
n, its own dangling mass) and within L1 <= 1e-6 of the FP64 oracle.
"""
import numpy as np
import pytest

import medusa_amd as md
from medusa_amd import workloads as wl
from conftest import csr_from_edges, random_graph

pytestmark = pytest.mark.gpu
L1_TOL = 1e-6


@pytest.fixture(autouse=True)
def _announce(request):
    print(f"\n[case] {request.node.nodeid}", flush=True)


def _bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


def _mixed_batch():
    """Uneven graphs including the edge cases of the per-graph combine: no edges at all,
    one vertex, every vertex dangling, more than one 4096-block of dangling vertices,
    in-degree above 64 (wave path), self loops and duplicates."""
    graphs = []
    for i, n in enumerate([1, 7, 300, 2500, 9000, 20000]):
        graphs.append(random_graph(n, 6 * n, 100 + i)[:2])
    graphs.append(csr_from_edges(50, np.zeros((0, 2), np.int64))[:2])           # no edges
    star = [(k, 0) for k in range(1, 400)]                                        # hub with in-degree 399
    graphs.append(csr_from_edges(400, star)[:2])
    rng = np.random.default_rng(9)
    src = rng.integers(0, 1500, 30000)
    dst = rng.integers(0, 12000, 30000)                                           # ~10500 dangling
    graphs.append(csr_from_edges(12000, np.stack([src, dst], 1))[:2])
    small, _ = wl.chung_lu_batch(count=12, lo_log2=6, hi_log2=12, edge_factor=8, seed=11)
    return graphs + small


def test_pack_graphs_layout():
    graphs = _mixed_batch()
    rp, col, w, begin = md.pack_graphs(graphs)
    assert begin[0] == 0 and begin[-1] == len(rp) - 1
    for k, (grp, gcol) in enumerate(graphs):
        b = begin[k]
        n = len(grp) - 1
        assert np.array_equal(rp[b:b + n + 1] - rp[b], grp)
        assert np.array_equal(col[rp[b]:rp[b + n]] - b, gcol)


def test_packed_pagerank_per_graph_bits(oracle):
    graphs = _mixed_batch()
    with md.Graph.packed(graphs) as G:
        begin = G.batch()
        assert len(begin) == len(graphs) + 1
        for steps in (1, 20):
            for v in (md.PR_PULL, md.PR_PUSH, md.PR_BINNED):
                for use_graph in (0, 1):
                    r, st = G.pagerank(0.85, steps, variant=v, use_graph=use_graph)
                    for k, (grp, gcol) in enumerate(graphs):
                        part = r[begin[k]:begin[k + 1]]
                        want = oracle.pagerank_f32(grp, gcol, 0.85, steps)
                        assert np.array_equal(_bits(part), _bits(want)), (k, steps, v, use_graph)
        r, _ = G.pagerank(0.85, 20)
        for k, (grp, gcol) in enumerate(graphs):
            part = r[begin[k]:begin[k + 1]].astype(np.float64)
            want64 = oracle.pagerank_f64(grp, gcol, 0.85, 20)
            assert np.abs(part - want64).sum() <= L1_TOL, k
            assert abs(part.sum() - 1.0) < 1e-5, k


def test_packed_pagerank_equals_single_runs():
    graphs = _mixed_batch()[:9]
    with md.Graph.packed(graphs) as G:
        r, _ = G.pagerank(0.85, 15)
        begin = G.batch()
    for k, (grp, gcol) in enumerate(graphs):
        with md.Graph(grp, gcol) as H:
            rk, _ = H.pagerank(0.85, 15)
        assert np.array_equal(_bits(r[begin[k]:begin[k + 1]]), _bits(rk)), k


def test_packed_pagerank_quarter_bins_large_items(oracle):
    """A batch large enough that the binned plan's bins overflow one round of the device's
    resident combine workgroups by a short last round (RMAT scale 19 plus a small graph: 1122
    full-size bins for 1024 workgroups on a 256-CU device), so the vertices of that round's bins
    are planned into quarter-size bins.  The small graph lies in that tail and has vertices whose
    slots exceed a quarter-size bin (wave vertices of in-degree 2100-4000, lane groups of
    in-degree 40-64), each of which gets a bin of its own; every graph's ranks stay bit-identical
    to the oracle's FP32 restatement."""
    big = wl.rmat_graph(scale=19)
    rng = np.random.default_rng(13)
    n = 3000
    edges = [np.stack([rng.integers(0, n, deg), np.full(deg, v)], 1) for v, deg in ((0, 4000), (1, 2100), (2, 3000))]
    for v in range(3, 200):
        d = int(rng.integers(40, 65))
        edges.append(np.stack([rng.integers(0, n, d), np.full(d, v)], 1))
    edges.append(np.stack([rng.integers(0, n, 5000), rng.integers(200, n, 5000)], 1))
    graphs = [big, csr_from_edges(n, np.concatenate(edges))[:2]]
    with md.Graph.packed(graphs) as G:
        begin = G.batch()
        for steps in (1, 5):
            for use_graph in (0, 1):
                r, _ = G.pagerank(0.85, steps, variant=md.PR_BINNED, use_graph=use_graph)
                for k, (grp, gcol) in enumerate(graphs):
                    want = oracle.pagerank_f32(grp, gcol, 0.85, steps)
                    assert np.array_equal(_bits(r[begin[k]:begin[k + 1]]), _bits(want)), (k, steps, use_graph)


def test_set_batch_back_to_plain_graph(oracle):
    graphs = _mixed_batch()[:6]
    rp, col, _, begin = md.pack_graphs(graphs)
    with md.Graph(rp, col) as G:
        plain, _ = G.pagerank(0.85, 10)
        G.set_batch(begin)
        G.pagerank(0.85, 10)
        G.set_batch([0, G.n])
        again, _ = G.pagerank(0.85, 10)
        assert np.array_equal(_bits(plain), _bits(again))
        assert np.array_equal(_bits(plain), _bits(oracle.pagerank_f32(rp, col, 0.85, 10)))


@pytest.mark.parametrize("direction", [md.DIR_PUSH, md.DIR_PULL, md.DIR_AUTO])
def test_packed_bfs_per_graph_levels(oracle, direction):
    graphs = _mixed_batch()
    with md.Graph.packed(graphs) as G:
        begin = G.batch()
        local = [wl.max_outdeg_vertex(grp) for grp, _ in graphs]
        sources = np.array([begin[k] + s for k, s in enumerate(local)], np.uint32)
        for timing in (0, 1):
            lv, st = G.bfs_multi(sources, direction=direction, timing=timing)
            reached = 0
            for k, (grp, gcol) in enumerate(graphs):
                want = oracle.bfs_queue(grp, gcol, local[k])
                assert np.array_equal(lv[begin[k]:begin[k + 1]], want), (k, timing)
                reached += int((want >= 0).sum())
            assert st.reached == reached


def test_bfs_multi_is_nearest_source(oracle):
    rp, col, _ = random_graph(30000, 90000, 21)
    rng = np.random.default_rng(3)
    sources = rng.integers(0, 30000, 40).astype(np.uint32)
    sources = np.concatenate([sources, sources[:5]])  # duplicates are allowed
    want = None
    for s in np.unique(sources):
        lv = oracle.bfs_queue(rp, col, int(s)).astype(np.int64)
        lv[lv < 0] = np.iinfo(np.int64).max
        want = lv if want is None else np.minimum(want, lv)
    want[want == np.iinfo(np.int64).max] = -1
    with md.Graph(rp, col) as G:
        for d in (md.DIR_PUSH, md.DIR_PULL, md.DIR_AUTO):
            lv, _ = G.bfs_multi(sources, direction=d)
            assert np.array_equal(lv, want.astype(np.int32)), d
        one, _ = G.bfs_multi(sources[:1])
        assert np.array_equal(one, oracle.bfs_queue(rp, col, int(sources[0])))


def test_set_batch_rejects_crossing_edges():
    rp, col, _ = csr_from_edges(6, [(0, 1), (1, 2), (2, 3), (4, 5)])
    with md.Graph(rp, col) as G:
        with pytest.raises(md.MedusaError):
            G.set_batch([0, 2, 6])      # edge 1->2 leaves graph 0
        with pytest.raises(md.MedusaError):
            G.set_batch([0, 4, 4, 6])   # empty graph
        with pytest.raises(md.MedusaError):
            G.set_batch([0, 4, 5])      # does not end at n
        G.set_batch([0, 4, 6])
        assert list(G.batch()) == [0, 4, 6]


def test_chung_lu_batch_config(oracle, direction_policy, variant_policy):
    """The 128-graph Chung-Lu batch (n log-uniform in 2^10..2^18, edge factor 8, seed 5):
    direction-optimising BFS per graph bit-exact (heuristic, learned direction, autotuned launch
    shape), 20-step PageRank per graph within L1 1e-6 of FP64 (bit-exact against the FP32
    restatement on every 8th graph; the learned kernel-variant policy gives the same bits)."""
    graphs, sizes = wl.chung_lu_batch()
    assert len(graphs) == 128 and sizes.min() >= 2 ** 10 - 1 and sizes.max() <= 2 ** 18
    with md.Graph.packed(graphs) as G:
        begin = G.batch()
        local = [wl.max_outdeg_vertex(grp) for grp, _ in graphs]
        sources = np.array([begin[k] + s for k, s in enumerate(local)], np.uint32)
        lv, st = G.bfs_multi(sources)
        for k, (grp, gcol) in enumerate(graphs):
            assert np.array_equal(lv[begin[k]:begin[k + 1]], oracle.bfs_queue(grp, gcol, local[k])), k
        lv_l, _ = G.bfs_multi(sources, policy=direction_policy)
        assert np.array_equal(lv_l, lv), "learned direction"
        G.autotune(md.TUNE_BFS, sources, reps=1)
        for pol in (None, direction_policy):
            lv_t, _ = G.bfs_multi(sources, policy=pol)
            assert np.array_equal(lv_t, lv), ("autotuned", pol is not None)
        r, _ = G.pagerank(0.85, 20)
        r_v, _ = G.pagerank(0.85, 20, policy=variant_policy)
        assert np.array_equal(_bits(r_v), _bits(r)), "variant policy"
        worst = 0.0
        for k, (grp, gcol) in enumerate(graphs):
            part = r[begin[k]:begin[k + 1]]
            l1 = np.abs(part.astype(np.float64) - oracle.pagerank_f64(grp, gcol, 0.85, 20)).sum()
            worst = max(worst, l1)
            if k % 8 == 0:
                assert np.array_equal(_bits(part), _bits(oracle.pagerank_f32(grp, gcol, 0.85, 20))), k
        assert worst <= L1_TOL, worst
        print(f"[chung_lu batch] {G.n} vertices, {G.m} arcs, worst per-graph L1 {worst:.3g}", flush=True)
