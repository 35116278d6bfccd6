"""Pins the CPU oracle: golden fixtures, closed forms, invariants and brute force.

The reference ships no tests or golden vectors (its datasets are not in the repository
either, .MISSING_LARGE_BLOBS), so every oracle function is checked here against something
other than its own formula.
"""
import json
import pathlib

import numpy as np
import pytest

from conftest import csr_from_edges, random_graph

GOLDEN = json.loads((pathlib.Path(__file__).parent / "golden" / "golden.json").read_text())
U32MAX = 0xFFFFFFFF


def _golden_csr(g):
    return csr_from_edges(g["n"], g["edges"], g["weights"])


# ---- golden fixtures ------------------------------------------------------------

@pytest.mark.parametrize("name", sorted(GOLDEN))
def test_golden_bfs(oracle, name):
    g = GOLDEN[name]
    rp, col, _ = _golden_csr(g)
    for s, want in g["bfs"].items():
        assert oracle.bfs_queue(rp, col, int(s)).tolist() == want
        lv, steps = oracle.bfs_sweep(rp, col, int(s))
        assert lv.tolist() == want
        assert steps == max(want) + 1  # the last superstep finds nothing


@pytest.mark.parametrize("name", sorted(GOLDEN))
def test_golden_sssp(oracle, name):
    g = GOLDEN[name]
    rp, col, w = _golden_csr(g)
    for s, want in g["sssp"].items():
        want = [U32MAX if x is None else x for x in want]
        assert oracle.sssp_dijkstra(rp, col, w, int(s)).tolist() == want
        assert oracle.sssp_bellman_ford(rp, col, w, int(s)).tolist() == want
        assert oracle.sssp_bsp(rp, col, w, int(s))[0].tolist() == want


@pytest.mark.parametrize("name", sorted(GOLDEN))
def test_golden_pagerank(oracle, name):
    g = GOLDEN[name]
    rp, col, _ = _golden_csr(g)
    for steps, want in g["pagerank_engine"].items():
        r64 = oracle.pagerank_f64(rp, col, 0.85, int(steps))
        np.testing.assert_allclose(r64, want, rtol=1e-13, atol=1e-15)
        r32 = oracle.pagerank_f32(rp, col, 0.85, int(steps))
        np.testing.assert_allclose(r32, want, rtol=2e-6)
    for steps, want in g["pagerank_reference"].items():
        rr = oracle.pagerank_ref(rp, col, int(steps))
        np.testing.assert_allclose(rr, want, rtol=2e-6)


# ---- closed forms ---------------------------------------------------------------

def test_bfs_closed_forms(oracle):
    n = 50
    path = [(i, i + 1) for i in range(n - 1)]
    rp, col, _ = csr_from_edges(n, path)
    assert oracle.bfs_queue(rp, col, 0).tolist() == list(range(n))
    assert oracle.bfs_queue(rp, col, 10).tolist() == [-1] * 10 + list(range(n - 10))
    cyc = path + [(n - 1, 0)]
    rp, col, _ = csr_from_edges(n, cyc)
    assert oracle.bfs_queue(rp, col, 7).tolist() == [(v - 7) % n for v in range(n)]
    # bidirectional grid: BFS level = Manhattan distance
    R, C = 9, 13
    e = []
    for r in range(R):
        for c in range(C):
            v = r * C + c
            if c + 1 < C:
                e += [(v, v + 1), (v + 1, v)]
            if r + 1 < R:
                e += [(v, v + C), (v + C, v)]
    rp, col, _ = csr_from_edges(R * C, e)
    lv = oracle.bfs_queue(rp, col, 4 * C + 6)
    want = [abs(r - 4) + abs(c - 6) for r in range(R) for c in range(C)]
    assert lv.tolist() == want
    # complete digraph: everything at level 1
    k = 12
    rp, col, _ = csr_from_edges(k, [(a, b) for a in range(k) for b in range(k) if a != b])
    assert oracle.bfs_queue(rp, col, 3).tolist() == [1 if v != 3 else 0 for v in range(k)]


def test_sssp_closed_forms(oracle):
    n = 40
    w = np.arange(1, n, dtype=np.uint32) * 3
    rp, col, wc = csr_from_edges(n, [(i, i + 1) for i in range(n - 1)], w)
    d = oracle.sssp_dijkstra(rp, col, wc, 0)
    assert d.tolist() == [0] + np.cumsum(w).tolist()
    # unit-weight grid: distance = Manhattan distance
    R, C = 7, 11
    e = []
    for r in range(R):
        for c in range(C):
            v = r * C + c
            if c + 1 < C:
                e += [(v, v + 1), (v + 1, v)]
            if r + 1 < R:
                e += [(v, v + C), (v + C, v)]
    rp, col, wc = csr_from_edges(R * C, e, np.ones(len(e), np.uint32))
    d = oracle.sssp_dijkstra(rp, col, wc, 0)
    assert d.tolist() == [r + c for r in range(R) for c in range(C)]


def _dense_google_power(rp, col, d, steps):
    """Dense-matrix power iteration r' = G r (independent formulation of the engine formula)."""
    n = len(rp) - 1
    A = np.zeros((n, n))
    outdeg = np.diff(rp.astype(np.int64))
    for u in range(n):
        for e in range(rp[u], rp[u + 1]):
            A[col[e], u] += 1.0 / outdeg[u]
    for u in range(n):
        if outdeg[u] == 0:
            A[:, u] = 1.0 / n  # dangling column spreads uniformly
    G = d * A + (1 - d) / n
    r = np.full(n, 1.0 / n)
    for _ in range(steps):
        r = G @ r
    return r, G


def test_pagerank_closed_forms(oracle):
    n = 17
    rp, col, _ = csr_from_edges(n, [(i, (i + 1) % n) for i in range(n)])
    np.testing.assert_allclose(oracle.pagerank_f64(rp, col, 0.85, 30), 1.0 / n, rtol=1e-14)
    np.testing.assert_allclose(oracle.pagerank_ref(rp, col, 100), 1.0, rtol=1e-6)
    k = 9
    rp, col, _ = csr_from_edges(k, [(a, b) for a in range(k) for b in range(k) if a != b])
    np.testing.assert_allclose(oracle.pagerank_f64(rp, col, 0.85, 30), 1.0 / k, rtol=1e-14)
    # no edges: every vertex dangling, ranks stay uniform
    rp, col, _ = csr_from_edges(5, np.zeros((0, 2)))
    np.testing.assert_allclose(oracle.pagerank_f64(rp, col, 0.85, 7), 0.2, rtol=1e-15)
    np.testing.assert_allclose(oracle.pagerank_f32(rp, col, 0.85, 7), 0.2, rtol=1e-7)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_pagerank_f64_matches_dense_and_stationary(oracle, seed):
    rp, col, _ = random_graph(60, 200, seed)
    for steps in (1, 5, 30):
        r_dense, G = _dense_google_power(rp, col, 0.85, steps)
        np.testing.assert_allclose(oracle.pagerank_f64(rp, col, 0.85, steps), r_dense, rtol=1e-12, atol=1e-15)
    # many steps converge to the stationary vector of G (eigenvalue 1)
    w, V = np.linalg.eig(G)
    pi = np.real(V[:, np.argmin(np.abs(w - 1))])
    pi /= pi.sum()
    np.testing.assert_allclose(oracle.pagerank_f64(rp, col, 0.85, 300), pi, rtol=1e-9)


@pytest.mark.parametrize("seed", [4, 5])
def test_pagerank_mass_and_f32_accuracy(oracle, seed):
    rp, col, _ = random_graph(3000, 30000, seed)
    # add dangling vertices: drop the out-edges of every 7th vertex
    r64 = oracle.pagerank_f64(rp, col, 0.85, 30)
    assert abs(r64.sum() - 1.0) < 1e-12
    r32 = oracle.pagerank_f32(rp, col, 0.85, 30)
    l1 = np.abs(r32.astype(np.float64) - r64).sum()
    assert l1 <= 1e-6, l1


def test_pagerank_f32_order_matters_and_is_pinned(oracle):
    """The FP32 restatement is a specific summation order: it must differ from a naive
    float32 sum on some vertex of a graph with hubs, yet stay within the L1 bound."""
    rng = np.random.default_rng(7)
    n = 5000
    src = rng.integers(0, n, 80000)
    dst = np.where(rng.random(80000) < 0.5, rng.integers(0, 4, 80000), rng.integers(0, n, 80000))
    rp, col, _ = csr_from_edges(n, np.stack([src, dst], 1))
    r64 = oracle.pagerank_f64(rp, col, 0.85, 20)
    r32 = oracle.pagerank_f32(rp, col, 0.85, 20)
    assert np.abs(r32 - r64).sum() <= 1e-6
    # hub in-degrees exceed one 4096-element block: compensated sum beats a naive one
    cptr, csrc, _ = oracle.csc(rp, col)
    outdeg = np.diff(rp.astype(np.int64))
    contrib = np.where(outdeg > 0, r64 / np.maximum(outdeg, 1), 0).astype(np.float32)
    hub = int(np.argmax(np.diff(cptr.astype(np.int64))))
    seg = contrib[csrc[cptr[hub]:cptr[hub + 1]]]
    assert len(seg) > 4096
    exact = np.sum(seg.astype(np.float64))
    naive = np.float32(0)
    for x in seg:
        naive = np.float32(naive + x)
    assert abs(float(naive) - exact) > 0  # naive float32 accumulation loses bits on a hub


# ---- invariants and brute force ---------------------------------------------------

@pytest.mark.parametrize("seed", range(6))
def test_bfs_invariants_and_paths(oracle, seed):
    n = 400
    rp, col, _ = random_graph(n, 900 + 300 * seed, seed)
    src = seed * 37 % n
    lv = oracle.bfs_queue(rp, col, src)
    sweep, _ = oracle.bfs_sweep(rp, col, src)
    assert np.array_equal(lv, sweep)
    assert lv[src] == 0
    for u in range(n):
        for e in range(rp[u], rp[u + 1]):
            v = col[e]
            if lv[u] >= 0:
                assert lv[v] >= 0 and lv[v] <= lv[u] + 1
    cptr, csrc, _ = oracle.csc(rp, col)
    for v in range(n):
        if lv[v] > 0:
            assert any(lv[csrc[i]] == lv[v] - 1 for i in range(cptr[v], cptr[v + 1]))


@pytest.mark.parametrize("seed", range(4))
def test_bfs_brute_force_floyd(oracle, seed):
    n = 30
    rp, col, _ = random_graph(n, 60, seed)
    D = np.full((n, n), np.inf)
    np.fill_diagonal(D, 0)
    for u in range(n):
        for e in range(rp[u], rp[u + 1]):
            D[u, col[e]] = min(D[u, col[e]], 1)
    for k in range(n):
        D = np.minimum(D, D[:, [k]] + D[[k], :])
    for s in range(n):
        want = np.where(np.isinf(D[s]), -1, D[s]).astype(np.int32)
        assert np.array_equal(oracle.bfs_queue(rp, col, s), want)


@pytest.mark.parametrize("seed", range(5))
def test_sssp_cross_and_invariants(oracle, seed):
    n = 300
    rp, col, _ = random_graph(n, 1500, seed)
    w = np.random.default_rng(seed).integers(1, 1001, len(col)).astype(np.uint32)
    d = oracle.sssp_dijkstra(rp, col, w, 0)
    assert np.array_equal(d, oracle.sssp_bellman_ford(rp, col, w, 0))
    assert np.array_equal(d, oracle.sssp_bsp(rp, col, w, 0)[0])
    dd = d.astype(np.int64)
    reached = d != U32MAX
    for u in np.flatnonzero(reached):
        for e in range(rp[u], rp[u + 1]):
            assert dd[col[e]] <= dd[u] + w[e]
    cptr, csrc, slot = oracle.csc(rp, col)
    # tight parent edge for every reached non-source vertex
    wc = np.empty_like(w)
    wc[slot] = w
    for v in np.flatnonzero(reached):
        if v == 0:
            continue
        ins = range(cptr[v], cptr[v + 1])
        assert any(reached[csrc[i]] and dd[csrc[i]] + wc[i] == dd[v] for i in ins)


def test_sssp_overflow_detected(oracle):
    rp, col, w = csr_from_edges(3, [(0, 1), (1, 2)], [0xFFFFFFF0, 0x20])
    with pytest.raises(OverflowError):
        oracle.sssp_dijkstra(rp, col, w, 0)


def test_mbfs_equals_independent_bfs(oracle):
    rp, col, _ = random_graph(700, 2500, 11)
    srcs = np.array([0, 5, 5, 699, 100], np.uint32)
    out = oracle.mbfs(rp, col, srcs)
    for i, s in enumerate(srcs):
        lv = oracle.bfs_queue(rp, col, int(s))
        want = np.where(lv < 0, 255, lv).astype(np.uint8)
        assert np.array_equal(out[i], want)


def test_csc_and_msgdst_bruteforce(oracle):
    rp, col, _ = random_graph(200, 1000, 3)
    cptr, csrc, slot = oracle.csc(rp, col)
    src = np.repeat(np.arange(200), np.diff(rp.astype(np.int64)))
    order = np.lexsort((np.arange(len(col)), src, col))  # by dst, then src, then edge index
    assert np.array_equal(csrc, src[order].astype(np.uint32))
    assert np.array_equal(np.diff(cptr.astype(np.int64)), np.bincount(col, minlength=200))
    inv = np.empty_like(order)
    inv[order] = np.arange(len(order))
    assert np.array_equal(slot, inv.astype(np.uint32))


def test_segmented_scan_definition(oracle):
    rng = np.random.default_rng(0)
    vals = rng.random(1000).astype(np.float32)
    heads = np.zeros(1000, np.uint32)
    starts = np.sort(rng.choice(np.arange(1, 1000), 60, replace=False))
    heads[0] = 1
    heads[starts] = 1
    mins = oracle.segmented_scan(vals, heads, "min")
    sums = oracle.segmented_scan(vals, heads, "add")
    bounds = np.concatenate([[0], starts, [1000]])
    for a, b in zip(bounds[:-1], bounds[1:]):
        assert mins[b - 1] == vals[a:b].min()
        np.testing.assert_allclose(sums[b - 1], vals[a:b].astype(np.float64).sum(), rtol=1e-5)
        assert np.all(np.diff(mins[a:b]) <= 0)


def test_planted_mistake_is_caught(oracle):
    """A dropped dangling term (a plausible mistake) must fail the dense-matrix check."""
    rp, col, _ = csr_from_edges(4, [(0, 1), (1, 2), (2, 0)])  # vertex 3 dangling and isolated
    r_ok = oracle.pagerank_f64(rp, col, 0.85, 10)
    r_dense, _ = _dense_google_power(rp, col, 0.85, 10)
    np.testing.assert_allclose(r_ok, r_dense, rtol=1e-12)
    r_bad = r_dense.copy()
    # recompute without redistributing dangling mass
    n = 4
    r = np.full(n, 0.25)
    for _ in range(10):
        s = np.zeros(n)
        s[1] += r[0]
        s[2] += r[1]
        s[0] += r[2]
        r = 0.15 / n + 0.85 * s
    r_bad = r
    assert not np.allclose(r_bad, r_dense, rtol=1e-6)


@pytest.mark.parametrize("seed", range(3))
def test_sssp_rounds_predicts_reference_path(oracle, seed):
    """The superstep count derived from distances equals the count the reference CPU path runs."""
    rp, col, _ = random_graph(500, 2500, seed)
    w = np.random.default_rng(seed).integers(1, 50, len(col)).astype(np.uint32)
    d, steps = oracle.sssp_bsp(rp, col, w, 3)
    assert oracle.sssp_rounds(rp, col, w, 3, d) == steps
    d2, s2 = oracle.sssp_bsp(rp, col, w, 3, max_steps=2)
    assert s2 == min(2, steps)


def test_mlp_oracle_gradient_matches_finite_differences(oracle):
    """The FP64 MLP gradient (chain rule) against central differences of the FP64 loss."""
    rng = np.random.default_rng(3)
    P = oracle.MLP_PARAMS
    params = (rng.standard_normal(P) * 0.3).astype(np.float32)
    X = rng.random((40, 16)).astype(np.float32)
    y = rng.standard_normal(40).astype(np.float32)
    loss, g = oracle.mlp_loss_grad(params, X, y)
    assert loss > 0
    idx = np.concatenate([rng.choice(P, 60, replace=False), [P - 1, 0, 512]])
    checked = 0
    for p in idx:
        h = np.float32(1e-3 * max(1.0, abs(float(params[p]))))
        up, dn = params.copy(), params.copy()
        up[p] = params[p] + h
        dn[p] = params[p] - h
        lu, _ = oracle.mlp_loss_grad(up, X, y)
        ld, _ = oracle.mlp_loss_grad(dn, X, y)
        fd = (lu - ld) / (float(up[p]) - float(dn[p]))
        # a kink of the leaky ReLU inside [p-h, p+h] makes the difference quotient inexact
        if abs(fd - g[p]) <= 1e-6 + 1e-4 * abs(g[p]):
            checked += 1
    assert checked >= len(idx) - 3, (checked, len(idx))


def test_mlp_oracle_forward_consistent_with_loss(oracle):
    rng = np.random.default_rng(4)
    params = (rng.standard_normal(oracle.MLP_PARAMS) * 0.2).astype(np.float32)
    X = rng.random((25, 16)).astype(np.float32)
    y = rng.standard_normal(25).astype(np.float32)
    out = oracle.mlp_forward(params, X)
    loss, _ = oracle.mlp_loss_grad(params, X, y)
    assert abs(loss - np.mean((out - y.astype(np.float64)) ** 2)) < 1e-12
    # a zero network predicts b3
    z = np.zeros(oracle.MLP_PARAMS, np.float32)
    z[-1] = 0.5
    assert np.all(oracle.mlp_forward(z, X) == 0.5)
