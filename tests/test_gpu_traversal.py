"""GPU parity for BFS, SSSP and multi-source BFS: bit-exact against the CPU oracle."""
import json
import pathlib

import numpy as np
import pytest

import medusa_amd as md
from medusa_amd import workloads as wl
from conftest import csr_from_edges, random_graph

pytestmark = pytest.mark.gpu
GOLDEN = json.loads((pathlib.Path(__file__).parent / "golden" / "golden.json").read_text())
U32MAX = 0xFFFFFFFF


@pytest.fixture(autouse=True)
def _announce(request):
    print(f"\n[case] {request.node.nodeid}", flush=True)


def _dirs():
    return [md.DIR_PUSH, md.DIR_PULL, md.DIR_AUTO]


# ---- BFS -------------------------------------------------------------------------

@pytest.mark.parametrize("name", sorted(GOLDEN))
def test_bfs_golden(name):
    g = GOLDEN[name]
    rp, col, _ = csr_from_edges(g["n"], g["edges"], g["weights"])
    with md.Graph(rp, col) as G:
        for s, want in g["bfs"].items():
            for d in _dirs():
                lv, st = G.bfs(int(s), direction=d)
                assert lv.tolist() == want, (s, d)
                assert st.reached == sum(1 for x in want if x >= 0)


@pytest.mark.parametrize("seed", range(4))
def test_bfs_random_vs_oracle(oracle, seed):
    n = 3000 + 500 * seed
    rp, col, _ = random_graph(n, 8 * n, seed)
    with md.Graph(rp, col) as G:
        for src in (0, n - 1, seed * 97 % n):
            want = oracle.bfs_queue(rp, col, src)
            for d in _dirs():
                lv, _ = G.bfs(src, direction=d)
                assert np.array_equal(lv, want), (src, d)


def test_bfs_hubs_use_large_path(oracle):
    # a few hubs with thousands of out-edges, and a small threshold so the
    # workgroup-split kernel handles them
    rng = np.random.default_rng(3)
    n = 20000
    src = np.concatenate([rng.integers(0, n, 60000), np.repeat([5, 77, 901], 9000)])
    dst = rng.integers(0, n, len(src))
    rp, col, _ = csr_from_edges(n, np.stack([src, dst], 1))
    with md.Graph(rp, col) as G:
        for large in (0, 64, 1):
            for s in (5, 3):
                lv, _ = G.bfs(s, direction=md.DIR_PUSH, large_degree=large)
                assert np.array_equal(lv, oracle.bfs_queue(rp, col, s)), (large, s)


def test_bfs_edge_cases(oracle):
    with md.Graph(np.array([0, 0], np.uint32), np.array([], np.uint32)) as G:
        lv, st = G.bfs(0)
        assert lv.tolist() == [0] and st.supersteps == 1
    # isolated source in a larger graph, self-loops and duplicates
    rp, col, _ = csr_from_edges(6, [(1, 1), (1, 2), (1, 2), (2, 3), (3, 1)])
    with md.Graph(rp, col) as G:
        assert G.bfs(0)[0].tolist() == [0, -1, -1, -1, -1, -1]
        assert G.bfs(1)[0].tolist() == [-1, 0, 1, 2, -1, -1]
        with pytest.raises(md.MedusaError):
            G.bfs(6)


def test_bfs_er_config(oracle):
    rp, col = wl.er_graph()
    want = oracle.bfs_queue(rp, col, 0)
    sweep, steps = oracle.bfs_sweep(rp, col, 0)
    assert np.array_equal(want, sweep)
    with md.Graph(rp, col) as G:
        for d in _dirs():
            lv, st = G.bfs(0, direction=d)
            assert np.array_equal(lv, want), d
            assert st.supersteps == steps  # one superstep per level plus the empty one
        out = np.zeros_like(want)
        deg = np.diff(rp.astype(np.int64))
        assert G.bfs(0, out=out)[1].edges_traversed == deg[want >= 0].sum()


@pytest.mark.slow
def test_bfs_rmat20(oracle):
    rp, col = wl.rmat_graph()
    with md.Graph(rp, col) as G:
        for s in (0, wl.max_outdeg_vertex(rp), 123457):
            want = oracle.bfs_queue(rp, col, s)
            for d in _dirs():
                lv, _ = G.bfs(s, direction=d)
                assert np.array_equal(lv, want), (s, d)


def test_bfs_device_output_and_repeatability(oracle):
    torch = pytest.importorskip("torch")
    rp, col, _ = random_graph(5000, 40000, 9)
    with md.Graph(rp, col) as G:
        out = torch.empty(5000, dtype=torch.int32, device="cuda")
        G.bfs(17, out=out)
        a = out.cpu().numpy()
        b, _ = G.bfs(17)
        assert np.array_equal(a, b)
        assert np.array_equal(a, oracle.bfs_queue(rp, col, 17))


# ---- SSSP ------------------------------------------------------------------------

@pytest.mark.parametrize("name", sorted(GOLDEN))
def test_sssp_golden(name):
    g = GOLDEN[name]
    rp, col, w = csr_from_edges(g["n"], g["edges"], g["weights"])
    with md.Graph(rp, col, w) as G:
        for s, want in g["sssp"].items():
            want = [U32MAX if x is None else x for x in want]
            for delta in (0, 1, 3, U32MAX):
                d, _ = G.sssp(int(s), delta=delta)
                assert d.tolist() == want, (s, delta)


@pytest.mark.parametrize("seed", range(3))
def test_sssp_random_vs_dijkstra(oracle, seed):
    n = 4000
    rp, col, _ = random_graph(n, 6 * n, seed)
    w = np.random.default_rng(seed).integers(1, 1001, len(col)).astype(np.uint32)
    with md.Graph(rp, col, w) as G:
        for src in (0, 1234):
            want = oracle.sssp_dijkstra(rp, col, w, src)
            for delta in (0, 1, 50, 1000, U32MAX):
                d, _ = G.sssp(src, delta=delta)
                assert np.array_equal(d, want), (src, delta)


def test_sssp_large_weights_and_unweighted_error(oracle):
    rp, col, w = csr_from_edges(4, [(0, 1), (1, 2), (0, 2), (2, 3)], [0x7FFFFFFF, 0x7FFFFFFF, 0xFFFFFF00, 0xF0])
    with md.Graph(rp, col, w) as G:
        d, _ = G.sssp(0)
        assert d.tolist() == oracle.sssp_dijkstra(rp, col, w, 0).tolist()
    with md.Graph(rp, col) as G:
        with pytest.raises(md.MedusaError) as e:
            G.sssp(0)
        assert e.value.code == -4


def test_sssp_lattice_small(oracle):
    rp, col, w = wl.lattice_graph(64, 64, seed=3)
    src = wl.lattice_source(64, 64)
    want = oracle.sssp_dijkstra(rp, col, w, src)
    assert np.array_equal(want, oracle.sssp_bsp(rp, col, w, src)[0])
    with md.Graph(rp, col, w) as G:
        for delta in (0, 1, 500, 5000, U32MAX):
            d, _ = G.sssp(src, delta=delta)
            assert np.array_equal(d, want), delta


@pytest.mark.slow
def test_sssp_lattice_config(oracle, delta_policy):
    """The 1400x1400 lattice at full size: the default width, a fixed width and the learned width
    (the policy trained on smaller lattices; device- and host-driven) all give Dijkstra's
    distances."""
    rp, col, w = wl.lattice_graph()
    src = wl.lattice_source()
    want = oracle.sssp_dijkstra(rp, col, w, src)
    deg = np.diff(rp.astype(np.int64))
    with md.Graph(rp, col, w) as G:
        for delta in (0, 2000):
            d, st = G.sssp(src, delta=delta)
            assert np.array_equal(d, want), delta
            assert st.edges_traversed == deg[want != U32MAX].sum()
        for timing in (0, 1):
            d, st = G.sssp(src, policy=delta_policy, timing=timing)
            assert np.array_equal(d, want), ("learned width", timing)
            assert st.edges_traversed == deg[want != U32MAX].sum()


# ---- multi-source BFS ----------------------------------------------------------------

def test_mbfs_small(oracle):
    rp, col, _ = random_graph(3000, 15000, 5)
    with md.Graph(rp, col) as G:
        for srcs in ([0], [0, 0, 5], list(range(0, 3000, 47))[:64], [2999] * 3):
            want = oracle.mbfs(rp, col, np.array(srcs, np.uint32))
            for d in _dirs():
                out, st = G.mbfs(srcs, direction=d)
                assert np.array_equal(out, want), (len(srcs), d)


def test_mbfs_matches_plain_bfs(oracle):
    rp, col = wl.er_graph(8192, 65536, seed=1)
    srcs = wl.mbfs_sources(rp, 64, seed=4)
    with md.Graph(rp, col) as G:
        out, st = G.mbfs(srcs)
        for i, s in enumerate(srcs):
            lv, _ = G.bfs(int(s))
            assert np.array_equal(out[i], np.where(lv < 0, 255, lv).astype(np.uint8))


@pytest.mark.slow
def test_mbfs_rmat20_config(oracle, mbfs_policy):
    """RMAT-20 with 64 sources at full size: the heuristic engine, the learned direction (policy
    trained on smaller RMAT graphs) and the autotuned launch shape (with and without the policy)
    all give the oracle's per-source levels."""
    rp, col = wl.rmat_graph()
    srcs = wl.mbfs_sources(rp, 64, seed=4)
    want = oracle.mbfs(rp, col, srcs)
    deg = np.diff(rp.astype(np.int64))
    reached = sum(deg[want[i] != 255].sum() for i in range(len(srcs)))
    with md.Graph(rp, col) as G:
        out, st = G.mbfs(srcs)
        assert np.array_equal(out, want)
        assert st.edges_traversed == reached
        out, st = G.mbfs(srcs, policy=mbfs_policy)
        assert np.array_equal(out, want), "learned direction"
        t = G.autotune(md.TUNE_MBFS, srcs, reps=1)
        assert t.large_degree > 0 and t.grid > 0
        for pol in (None, mbfs_policy):
            out, st = G.mbfs(srcs, policy=pol)
            assert np.array_equal(out, want), ("autotuned", pol is not None)
            assert st.edges_traversed == reached


@pytest.mark.parametrize("far_limit", [1, 7, 64])
def test_sssp_far_queue_limit_rebuilds(oracle, far_limit):
    """A far queue capped far below its need is rebuilt from the distances at every split;
    distances stay exact on both drivers, for every bucket width, and with local rounds (whose
    split slot relaxes the vertices it moves, also after a rescan) on a random graph and a
    lattice."""
    rp, col, _ = random_graph(6000, 36000, 31)
    w = np.random.default_rng(31).integers(1, 1001, len(col)).astype(np.uint32)
    lrp, lcol, lw = wl.lattice_graph(60, 70, seed=13)
    for rp, col, w, src in ((rp, col, w, 7), (lrp, lcol, lw, wl.lattice_source(60, 70))):
        want = oracle.sssp_dijkstra(rp, col, w, src)
        with md.Graph(rp, col, w) as G:
            for lr in (0, 32, U32MAX):
                for delta in (0, 50, 400):
                    for timing in (0, 1):
                        d, _ = G.sssp(src, delta=delta, far_limit=far_limit, timing=timing, local_rounds=lr)
                        assert np.array_equal(d, want), (lr, delta, timing)


def test_sssp_zero_weights_duplicates_and_self_loops(oracle):
    """The reference draws weights from rand32() % 1000, so 0 occurs; duplicates and self
    loops are kept."""
    rng = np.random.default_rng(8)
    n, m = 4000, 30000
    src = rng.integers(0, n, m)
    dst = rng.integers(0, n, m)
    src[:500] = dst[:500]                         # self loops
    src[500:1500] = src[1500:2500]                # duplicate arcs
    dst[500:1500] = dst[1500:2500]
    w = rng.integers(0, 1000, m).astype(np.uint32)
    w[::7] = 0
    rp, col, wc = csr_from_edges(n, np.stack([src, dst], 1), w)
    want = oracle.sssp_dijkstra(rp, col, wc, 11)
    with md.Graph(rp, col, wc) as G:
        for delta in (0, 1, 37, 0xFFFFFFFF):
            for timing in (0, 1):
                d, _ = G.sssp(11, delta=delta, timing=timing)
                assert np.array_equal(d, want), (delta, timing)


@pytest.mark.parametrize("case", ["lattice", "random", "zero_weights", "hubs"])
def test_sssp_local_rounds(oracle, case):
    """Workgroups relaxing the vertices they lowered themselves for up to 1, 3 or 64 further
    rounds, or none, in both loop drivers: distances stay Dijkstra's.  The cases cover a
    thin frontier (lattice), a dense random graph, zero weights with duplicate arcs and self
    loops (offers that tie), and hubs wider than one near-queue slice (never kept locally)."""
    rng = np.random.default_rng(21)
    if case == "lattice":
        rp, col, w = wl.lattice_graph(120, 150, seed=9)
        src = wl.lattice_source(120, 150)
    elif case == "random":
        rp, col, _ = random_graph(6000, 60000, 21)
        w = rng.integers(1, 1001, len(col)).astype(np.uint32)
        src = 3
    elif case == "zero_weights":
        n, m = 3000, 9000
        e = rng.integers(0, n, (m, 2))
        e[:300, 0] = e[:300, 1]
        e[300:900] = e[900:1500]
        w = rng.integers(0, 20, m).astype(np.uint32)
        w[::3] = 0
        rp, col, w = csr_from_edges(n, e, w)
        src = 5
    else:
        n = 20000
        hub = np.stack([np.zeros(3000, np.int64), rng.integers(1, n, 3000)], 1)
        rest = np.stack([rng.integers(0, n, 40000), rng.integers(0, n, 40000)], 1)
        rp, col, _ = csr_from_edges(n, np.concatenate([hub, rest]))
        w = rng.integers(1, 1001, len(col)).astype(np.uint32)
        src = 0
    want = oracle.sssp_dijkstra(rp, col, w, src)
    with md.Graph(rp, col, w) as G:
        for lr in (1, 3, 64, U32MAX, 0):
            for delta in ((0, 300, U32MAX) if case == "lattice" else (0, 1, 300, U32MAX)):
                for timing in (0, 1):
                    d, _ = G.sssp(src, delta=delta, timing=timing, local_rounds=lr)
                    assert np.array_equal(d, want), (lr, delta, timing)
        with pytest.raises(md.MedusaError):
            G.sssp(src, local_rounds=65)


def test_sssp_sparse_power_law_local_rounds(oracle):
    """A sparse skewed graph (Chung-Lu, mean out-degree about 2, so local rounds are on by
    default) whose hubs span many near-queue slices and are never kept locally, and whose
    frontier grows past the queue length up to which the rounds run: distances stay Dijkstra's
    with the default rounds, with none, and with the widest bucket."""
    rp, col = wl.chung_lu_graph(n=200_000, m=420_000, seed=17, max_degree=20_000)
    w = wl.weighted(rp, col, seed=17)
    src = wl.max_outdeg_vertex(rp)
    want = oracle.sssp_dijkstra(rp, col, w, src)
    with md.Graph(rp, col, w) as G:
        for kw in ({}, {"local_rounds": U32MAX}, {"delta": U32MAX}, {"local_rounds": 64, "delta": 300},
                   {"timing": 1}):
            d, st = G.sssp(src, **kw)
            assert np.array_equal(d, want), kw
            assert st.reached == int((want != U32MAX).sum())


def test_sssp_wide_bucket_split_moves_most_vertices(oracle):
    """A sparse graph of more than 2^20 vertices with local rounds on and a bucket so wide that
    the first split moves almost every vertex below the threshold: the split appends the
    vertices it moves and its local rounds append the ones they lower again, so one slot may
    append a vertex twice.  Distances stay Dijkstra's for several wide widths."""
    rp, col = wl.chung_lu_graph(n=1_200_000, m=2_500_000, seed=29, max_degree=5_000)
    w = wl.weighted(rp, col, seed=29)
    src = wl.max_outdeg_vertex(rp)
    want = oracle.sssp_dijkstra(rp, col, w, src)
    with md.Graph(rp, col, w) as G:
        for delta in (20_000, 200_000, 1 << 30):
            d, st = G.sssp(src, delta=delta, local_rounds=64)
            assert np.array_equal(d, want), delta
            assert st.reached == int((want != U32MAX).sum())


def test_sssp_hub_source_slices(oracle):
    """A source whose out-degree spans many near-queue slices, plus hubs inside the graph."""
    rng = np.random.default_rng(12)
    n = 50000
    hub = np.stack([np.zeros(20000, np.int64), rng.integers(1, n, 20000)], 1)
    hubs2 = np.stack([np.repeat(rng.integers(0, n, 8), 3000), rng.integers(0, n, 24000)], 1)
    rest = rng.integers(0, n, (150000, 2))
    rp, col, _ = csr_from_edges(n, np.concatenate([hub, hubs2, rest]))
    w = rng.integers(1, 1001, len(col)).astype(np.uint32)
    want = oracle.sssp_dijkstra(rp, col, w, 0)
    with md.Graph(rp, col, w) as G:
        for delta in (0, 100, 0xFFFFFFFF):
            for timing in (0, 1):
                d, _ = G.sssp(0, delta=delta, timing=timing)
                assert np.array_equal(d, want), (delta, timing)


@pytest.mark.parametrize("case", ["n1_empty", "n1_selfloop", "n2_empty", "n3_chain"])
def test_tiny_graphs_every_engine(oracle, case):
    """The smallest inputs through every engine and driver: one vertex with and without a
    self loop, two isolated vertices, a 3-chain."""
    edges, n = {"n1_empty": ([], 1), "n1_selfloop": ([(0, 0)], 1), "n2_empty": ([], 2),
                "n3_chain": ([(0, 1), (1, 2)], 3)}[case]
    rp, col, w = csr_from_edges(n, np.array(edges, np.int64).reshape(-1, 2), [7] * len(edges))
    w = np.asarray(w if w is not None else [], np.uint32)
    with md.Graph(rp, col, w) as G:
        for src in range(n):
            want = oracle.bfs_queue(rp, col, src)
            for d in _dirs():
                for timing in (0, 1):
                    assert np.array_equal(G.bfs(src, direction=d, timing=timing)[0], want)
            dwant = oracle.sssp_dijkstra(rp, col, w, src)
            for timing in (0, 1):
                assert np.array_equal(G.sssp(src, timing=timing)[0], dwant)
            srcs = np.array([src, src], np.uint32)
            mwant = oracle.mbfs(rp, col, srcs)
            for timing in (0, 1):
                assert np.array_equal(G.mbfs(srcs, timing=timing)[0], mwant)
        r, _ = G.pagerank(0.85, 4)
        assert np.array_equal(r.view(np.uint32), oracle.pagerank_f32(rp, col, 0.85, 4).view(np.uint32))


def test_one_handle_from_several_threads(oracle):
    """Calls on one handle from several host threads are serialised by the handle's lock:
    every thread gets its own source's exact levels, distances and ranks."""
    import threading
    n = 20000
    rp, col, _ = random_graph(n, 8 * n, 21)
    w = np.random.default_rng(21).integers(1, 1001, len(col)).astype(np.uint32)
    sources = [0, 17, 999, 4321, 12345, 19999]
    want_lv = {s: oracle.bfs_queue(rp, col, s) for s in sources}
    want_d = {s: oracle.sssp_dijkstra(rp, col, w, s) for s in sources}
    errors = []
    with md.Graph(rp, col, w) as G:
        want_pr = G.pagerank(steps=10)[0]

        def work(s):
            try:
                for _ in range(3):
                    lv, _ = G.bfs(s)
                    d, _ = G.sssp(s)
                    pr, _ = G.pagerank(steps=10)
                    mb, _ = G.mbfs([s, (s + 1) % n])
                    if not (np.array_equal(lv, want_lv[s]) and np.array_equal(d, want_d[s])
                            and np.array_equal(pr, want_pr) and np.array_equal(mb[0], np.where(
                                want_lv[s] < 0, 0xFF, want_lv[s]).astype(np.uint8))):
                        errors.append(s)
            except Exception as e:  # reported below
                errors.append((s, repr(e)))

        threads = [threading.Thread(target=work, args=(s,)) for s in sources]
        for t in threads:
            t.start()
        for t in threads:
            t.join(timeout=120)
        assert not any(t.is_alive() for t in threads)
    assert errors == []


@pytest.mark.slow
def test_rmat22_four_times_the_configs(oracle):
    """RMAT scale 22 (4,194,304 vertices, 67,108,864 arcs), four times the RMAT-20 configs:
    BFS, SSSP and 8-source MBFS stay bit-exact against the oracle."""
    rp, col = wl.rmat_graph(scale=22, edge_factor=16, seed=22)
    w = wl.weighted(rp, col, seed=23)
    src = wl.max_outdeg_vertex(rp)
    want = oracle.bfs_queue(rp, col, src)
    with md.Graph(rp, col, w) as G:
        for d in _dirs():
            lv, _ = G.bfs(src, direction=d)
            assert np.array_equal(lv, want), d
        dist, _ = G.sssp(src)
        assert np.array_equal(dist, oracle.sssp_dijkstra(rp, col, w, src))
        srcs = wl.mbfs_sources(rp, 8, seed=24)
        out, _ = G.mbfs(srcs)
        assert np.array_equal(out, oracle.mbfs(rp, col, srcs))
