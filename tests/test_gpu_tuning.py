"""Launch shapes, the autotuner and the device-driven frontier loops.

The shape of a launch (hub threshold, expansion workgroups) and the direction chosen per
superstep (heuristic, fixed, or any policy, even an untrained one) must never change
results: every variant is compared bit-exact against the oracle.
"""
import numpy as np
import pytest

import medusa_amd as md
from medusa_amd import workloads as wl
from conftest import csr_from_edges, random_graph

pytestmark = pytest.mark.gpu

SHAPES = [(1, 2), (1, 7), (16, 64), (128, 512), (2048, 2048), (1 << 30, 8192)]


@pytest.fixture(autouse=True)
def _announce(request):
    print(f"\n[case] {request.node.nodeid}", flush=True)


@pytest.fixture(scope="module")
def skewed():
    rp, col = wl.rmat_graph(scale=16, edge_factor=16, seed=41)
    return rp, col


def test_bfs_every_shape_same_levels(oracle, skewed):
    rp, col = skewed
    src = wl.max_outdeg_vertex(rp)
    want = oracle.bfs_queue(rp, col, src)
    with md.Graph(rp, col) as G:
        for large, grid in SHAPES:
            for d in (md.DIR_PUSH, md.DIR_PULL, md.DIR_AUTO):
                for timing in (0, 1):
                    lv, st = G.bfs(src, direction=d, large_degree=large, grid=grid, timing=timing)
                    assert np.array_equal(lv, want), (large, grid, d, timing)


def test_mbfs_every_shape_same_levels(oracle, skewed):
    rp, col = skewed
    srcs = wl.mbfs_sources(rp, 64, seed=5)
    want = oracle.mbfs(rp, col, srcs)
    with md.Graph(rp, col) as G:
        for large, grid in SHAPES:
            for d in (md.DIR_PUSH, md.DIR_PULL, md.DIR_AUTO):
                for timing in (0, 1):
                    out, st = G.mbfs(srcs, direction=d, large_degree=large, grid=grid, timing=timing)
                    assert np.array_equal(out, want), (large, grid, d, timing)


def test_mbfs_device_and_host_drivers_agree(oracle):
    rp, col, _ = random_graph(20000, 120000, 17)
    srcs = np.arange(0, 20000, 500, dtype=np.uint32)[:40]
    with md.Graph(rp, col) as G:
        for d in (md.DIR_AUTO, md.DIR_PUSH, md.DIR_PULL):
            a, sa = G.mbfs(srcs, direction=d)
            b, sb = G.mbfs(srcs, direction=d, timing=1)
            assert np.array_equal(a, b)
            assert (sa.supersteps, sa.push_steps, sa.pull_steps, sa.edges_traversed, sa.reached) == \
                   (sb.supersteps, sb.push_steps, sb.pull_steps, sb.edges_traversed, sb.reached)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_any_policy_keeps_results_exact(oracle, skewed, seed):
    """An untrained network makes arbitrary direction choices: levels must not change."""
    rp, col = skewed
    pol = md.Policy(md.POLICY_DIRECTION, seed=seed)
    src = wl.max_outdeg_vertex(rp)
    srcs = wl.mbfs_sources(rp, 64, seed=seed)
    with md.Graph(rp, col) as G:
        assert np.array_equal(G.bfs(src, policy=pol)[0], oracle.bfs_queue(rp, col, src))
        assert np.array_equal(G.bfs(src, policy=pol, trace=1)[0], oracle.bfs_queue(rp, col, src))
        want = oracle.mbfs(rp, col, srcs)
        assert np.array_equal(G.mbfs(srcs, policy=pol)[0], want)
        assert np.array_equal(G.mbfs(srcs, policy=pol, trace=1)[0], want)


def test_mbfs_saturated_levels(oracle):
    """A path longer than 255: levels saturate at 0xFF and the flag is raised."""
    n = 300
    rp = np.arange(n + 1, dtype=np.uint32)
    rp[-1] = n - 1
    col = np.arange(1, n, dtype=np.uint32)
    with md.Graph(rp, col) as G:
        for timing in (0, 1):
            out, st = G.mbfs([0, 10], timing=timing)
            want = oracle.mbfs(rp, col, np.array([0, 10], np.uint32))
            assert np.array_equal(out, want)
            assert st.reserved[0] == 1


def test_autotune_bfs(oracle, skewed):
    rp, col = skewed
    src = wl.max_outdeg_vertex(rp)
    want = oracle.bfs_queue(rp, col, src)
    with md.Graph(rp, col) as G:
        t = G.autotune(md.TUNE_BFS, [src], reps=2)
        assert t.candidates >= 30 and t.best_ms <= t.default_ms
        assert (t.large_degree, t.grid) != (0, 0)
        lv, _ = G.bfs(src)                 # runs with the tuned shape
        assert np.array_equal(lv, want)
        G.autotune(md.TUNE_BFS, [src], reps=0)
        assert np.array_equal(G.bfs(src)[0], want)
        print(f"[tune bfs] large={t.large_degree} grid={t.grid} best {t.best_ms:.3f} ms "
              f"default {t.default_ms:.3f} ms", flush=True)


def test_autotune_mbfs_and_batch(oracle):
    graphs, _ = wl.chung_lu_batch(count=24, lo_log2=8, hi_log2=14, edge_factor=8, seed=12)
    with md.Graph.packed(graphs) as G:
        begin = G.batch()
        local = [wl.max_outdeg_vertex(grp) for grp, _ in graphs]
        sources = np.array([begin[k] + s for k, s in enumerate(local)], np.uint32)
        t = G.autotune(md.TUNE_BFS, sources, reps=1)
        assert t.best_ms <= t.default_ms
        lv, _ = G.bfs_multi(sources)
        for k, (grp, gcol) in enumerate(graphs):
            assert np.array_equal(lv[begin[k]:begin[k + 1]], oracle.bfs_queue(grp, gcol, local[k])), k
        rp, col = md.pack_graphs(graphs)[:2]
        srcs = sources[:64]
        tm = G.autotune(md.TUNE_MBFS, srcs, reps=1)
        assert tm.best_ms <= tm.default_ms
        assert np.array_equal(G.mbfs(srcs)[0], oracle.mbfs(rp, col, srcs))


def test_bad_grid_is_an_error():
    rp, col, _ = random_graph(100, 400, 3)
    with md.Graph(rp, col) as G:
        with pytest.raises(md.MedusaError):
            G.bfs(0, grid=1)
        with pytest.raises(md.MedusaError):
            G.mbfs([0], grid=70000)
        with pytest.raises(md.MedusaError):
            G.autotune(7, [0])


def test_first_replay_sized_from_previous_run(oracle):
    """The device loops size their first replay from the previous run on the handle: a long
    traversal after a short one (more slots than predicted) and a short one after a long one
    (idle slots) must both give exact levels."""
    n = 300
    path = [(i, i + 1) for i in range(199)]              # 200-vertex path: 200 supersteps
    star = [(200, j) for j in range(201, 300)]           # star: 2 supersteps
    rp, col, _ = csr_from_edges(n, path + star)
    with md.Graph(rp, col) as G:
        for src in (200, 0, 200, 0, 150, 200):
            lv, st = G.bfs(src)
            assert np.array_equal(lv, oracle.bfs_queue(rp, col, src)), src
            out, _ = G.mbfs([src, 100])
            assert np.array_equal(out, oracle.mbfs(rp, col, np.array([src, 100], np.uint32))), src
