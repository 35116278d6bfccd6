"""GPU parity for PageRank.

Two bars: the FP32 ranks are bit-identical to the oracle's FP32 restatement of
the engine's canonical compensated-combine order (so every kernel variant and
lane mapping is pinned), and within L1 <= 1e-6 of the FP64 oracle.
"""
import json
import pathlib

import numpy as np
import pytest

import medusa_amd as md
from medusa_amd import workloads as wl
from conftest import csr_from_edges, random_graph

pytestmark = pytest.mark.gpu
GOLDEN = json.loads((pathlib.Path(__file__).parent / "golden" / "golden.json").read_text())
L1_TOL = 1e-6  # FP32 + Kahan vs FP64, summed over all vertices (ranks sum to 1)


@pytest.fixture(autouse=True)
def _announce(request):
    print(f"\n[case] {request.node.nodeid}", flush=True)


def _bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


def _check(G, rp, col, oracle, steps=30, variants=(md.PR_PULL, md.PR_PUSH, md.PR_BINNED)):
    want32 = oracle.pagerank_f32(rp, col, 0.85, steps)
    want64 = oracle.pagerank_f64(rp, col, 0.85, steps)
    for v in variants:
        for use_graph in (0, 1):
            r, st = G.pagerank(0.85, steps, variant=v, use_graph=use_graph)
            assert np.array_equal(_bits(r), _bits(want32)), (v, use_graph)
            assert st.supersteps == steps
    r, _ = G.pagerank(0.85, steps)
    l1 = np.abs(r.astype(np.float64) - want64).sum()
    assert l1 <= L1_TOL, l1
    return r, want64


@pytest.mark.parametrize("name", sorted(GOLDEN))
def test_pagerank_golden(oracle, name):
    g = GOLDEN[name]
    rp, col, _ = csr_from_edges(g["n"], g["edges"], g["weights"])
    with md.Graph(rp, col) as G:
        for steps, want in g["pagerank_engine"].items():
            r, _ = G.pagerank(0.85, int(steps))
            np.testing.assert_allclose(r, want, rtol=2e-6)
            assert np.array_equal(_bits(r), _bits(oracle.pagerank_f32(rp, col, 0.85, int(steps))))


@pytest.mark.parametrize("seed", range(3))
def test_pagerank_random(oracle, seed):
    n = 5000 * (seed + 1)
    rp, col, _ = random_graph(n, 10 * n, seed)
    with md.Graph(rp, col) as G:
        _check(G, rp, col, oracle, steps=20)


def test_pagerank_all_lane_classes(oracle):
    """In-degrees spanning every lane-group class (1..64), the wave class (>64) and the
    multi-block case (>4096), plus many dangling vertices."""
    rng = np.random.default_rng(11)
    n = 30000
    parts = []
    for deg in (1, 2, 3, 5, 9, 17, 33, 64, 65, 300, 4096, 4097, 9000):
        v = rng.integers(0, n)
        parts.append(np.stack([rng.integers(0, n, deg), np.full(deg, v)], 1))
    parts.append(np.stack([rng.integers(0, n // 3, 60000), rng.integers(0, n, 60000)], 1))
    rp, col, _ = csr_from_edges(n, np.concatenate(parts))
    assert (np.diff(rp.astype(np.int64)) == 0).mean() > 0.3  # dangling-heavy
    with md.Graph(rp, col) as G:
        _check(G, rp, col, oracle, steps=15)


def test_pagerank_edge_cases(oracle):
    # no edges: every vertex dangling
    with md.Graph(np.zeros(6, np.uint32), np.array([], np.uint32)) as G:
        r, _ = G.pagerank(0.85, 5)
        np.testing.assert_allclose(r, 0.2, rtol=1e-7)
    rp, col, _ = csr_from_edges(3, [(0, 1), (1, 2), (2, 0)])
    with md.Graph(rp, col) as G:
        r, _ = G.pagerank(0.85, 0)
        assert np.array_equal(_bits(r), _bits(np.full(3, np.float32(1 / 3))))
        r, st = G.pagerank(0.85, 7)
        assert np.array_equal(_bits(r), _bits(oracle.pagerank_f32(rp, col, 0.85, 7)))
        with pytest.raises(md.MedusaError):
            G.pagerank(1.5, 3)


def test_pagerank_binned_shapes(oracle):
    """Shapes at the edges of the binned variant's plan: no edges at all (nothing to send),
    every in-edge into one hub of in-degree above 4096 with nothing dangling (no message and no
    dangling block, so the send launch is skipped), a hub plus ordinary vertices, and a graph
    whose bins cut runs of equal in-degree."""
    n = 5000
    shapes = [csr_from_edges(6, np.zeros((0, 2), np.int64))]
    shapes.append(csr_from_edges(n, np.stack([np.arange(n), np.zeros(n, np.int64)], 1)))
    rng = np.random.default_rng(5)
    hub = np.stack([rng.integers(0, n, 6000), np.full(6000, 7)], 1)
    rest = np.stack([rng.integers(0, n, 20000), rng.integers(0, n, 20000)], 1)
    shapes.append(csr_from_edges(n, np.concatenate([hub, rest])))
    big = np.stack([rng.integers(0, 40000, 400000), rng.integers(0, 40000, 400000) % 2000], 1)
    shapes.append(csr_from_edges(40000, big))
    for rp, col, _ in shapes:
        with md.Graph(rp, col) as G:
            for steps in (1, 6):
                want = oracle.pagerank_f32(rp, col, 0.85, steps)
                for use_graph in (0, 1):
                    r, _ = G.pagerank(0.85, steps, variant=md.PR_BINNED, use_graph=use_graph)
                    assert np.array_equal(_bits(r), _bits(want)), (len(rp) - 1, steps, use_graph)


def test_pagerank_damping_and_repeat(oracle):
    rp, col, _ = random_graph(8000, 50000, 4)
    with md.Graph(rp, col) as G:
        for d in (0.5, 0.85, 0.99):
            r1, _ = G.pagerank(d, 12, use_graph=1)
            r2, _ = G.pagerank(d, 12)
            assert np.array_equal(_bits(r1), _bits(r2))
            assert np.array_equal(_bits(r1), _bits(oracle.pagerank_f32(rp, col, d, 12)))


def test_pagerank_er_config(oracle):
    rp, col = wl.er_graph()
    with md.Graph(rp, col) as G:
        _check(G, rp, col, oracle, steps=30)


@pytest.mark.slow
def test_pagerank_rmat20_config(oracle):
    rp, col = wl.rmat_graph()
    with md.Graph(rp, col) as G:
        r, r64 = _check(G, rp, col, oracle, steps=30, variants=(md.PR_PULL, md.PR_BINNED))
        assert abs(r.astype(np.float64).sum() - 1.0) < 1e-5
        # FP64 residual between the last two iterates is reported and small
        _, st = G.pagerank(0.85, 30)
        assert 0 < st.residual_l1 < 1e-2


def test_chung_lu_config(oracle, direction_policy, variant_policy):
    """The Chung-Lu configuration at full size: 2.4 M vertices, 5 M arcs, Zipf(2.0) out-weights
    capped at 100,000, 40 % weight 0 (seed 5): direction-optimising BFS bit-exact with the
    heuristic, the learned direction and the autotuned launch shape; 20-step PageRank with the
    default variant and with the learned kernel-variant policy bit-exact against the FP32
    restatement and within L1 1e-6 of FP64."""
    rp, col = wl.chung_lu_graph()
    src = wl.max_outdeg_vertex(rp)
    want = oracle.bfs_queue(rp, col, src)
    want32 = oracle.pagerank_f32(rp, col, 0.85, 20)
    with md.Graph(rp, col) as G:
        lv, st = G.bfs(src)
        assert np.array_equal(lv, want)
        assert st.push_steps > 0 and st.pull_steps > 0
        lv, _ = G.bfs(src, policy=direction_policy)
        assert np.array_equal(lv, want), "learned direction"
        G.autotune(md.TUNE_BFS, [src], reps=1)
        for pol in (None, direction_policy):
            lv, _ = G.bfs(src, policy=pol)
            assert np.array_equal(lv, want), ("autotuned", pol is not None)
        r, _ = G.pagerank(0.85, 20, use_graph=1)
        assert np.array_equal(_bits(r), _bits(want32))
        l1 = np.abs(r.astype(np.float64) - oracle.pagerank_f64(rp, col, 0.85, 20)).sum()
        assert l1 <= L1_TOL, l1
        for use_graph in (0, 1):
            r, _ = G.pagerank(0.85, 20, use_graph=use_graph, policy=variant_policy)
            assert np.array_equal(_bits(r), _bits(want32)), ("variant policy", use_graph)


def test_pagerank_more_than_2_24_vertices(oracle):
    """n above 2^24, where n and 1/n are no longer exact in FP32: the ranks stay bit-identical
    to the FP32 restatement (which rounds n the same way) and still sum to 1."""
    n = (1 << 24) + 4099
    m = 2 * n
    rng = np.random.default_rng(24)
    src = np.sort(rng.integers(0, n, m))
    col = rng.integers(0, n, m).astype(np.uint32)
    rp = np.zeros(n + 1, np.uint32)
    rp[1:] = np.cumsum(np.bincount(src, minlength=n))
    with md.Graph(rp, col) as G:
        r, st = G.pagerank(0.85, 5)
    assert st.supersteps == 5
    assert np.array_equal(_bits(r), _bits(oracle.pagerank_f32(rp, col, 0.85, 5)))
    assert abs(r.astype(np.float64).sum() - 1.0) < 1e-4


@pytest.mark.slow
def test_pagerank_rmat22(oracle):
    """RMAT scale 22 (4,194,304 vertices, 67,108,864 arcs): 10 supersteps bit-identical to
    the FP32 restatement, within L1 1e-6 of FP64."""
    rp, col = wl.rmat_graph(scale=22, edge_factor=16, seed=22)
    with md.Graph(rp, col) as G:
        r, st = G.pagerank(0.85, 10, use_graph=1)
    assert st.supersteps == 10
    assert np.array_equal(_bits(r), _bits(oracle.pagerank_f32(rp, col, 0.85, 10)))
    l1 = np.abs(r.astype(np.float64) - oracle.pagerank_f64(rp, col, 0.85, 10)).sum()
    assert l1 <= L1_TOL, l1
