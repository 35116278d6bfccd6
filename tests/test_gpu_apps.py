"""The example applications written against Medusa's user API (include/medusa/medusa.h):
BFS, SSSP, PageRank (reference formula) and multi-source BFS, run as separate
processes on the GPU and checked against the CPU oracle."""
import pathlib
import subprocess

import numpy as np
import pytest

from conftest import random_graph

pytestmark = pytest.mark.gpu
BIN = pathlib.Path(__file__).resolve().parent.parent / "medusa_amd" / "apps" / "bin"


@pytest.fixture(autouse=True)
def _announce(request):
    print(f"\n[case] {request.node.nodeid}", flush=True)


def write_csr(path, rp, col, w=None):
    with open(path, "wb") as f:
        f.write(b"MDSCSR01")
        f.write(np.uint32(len(rp) - 1).tobytes())
        f.write(np.uint64(len(col)).tobytes())
        f.write(np.uint32(0 if w is None else 1).tobytes())
        f.write(np.asarray(rp, np.uint32).tobytes())
        f.write(np.asarray(col, np.uint32).tobytes())
        if w is not None:
            f.write(np.asarray(w, np.uint32).tobytes())


def write_gtgraph(path, rp, col, w):
    n = len(rp) - 1
    with open(path, "w") as f:
        f.write("c generated\n")
        f.write(f"p sp {n} {len(col)}\n")
        for u in range(n):
            for e in range(rp[u], rp[u + 1]):
                f.write(f"a {u + 1} {col[e] + 1} {w[e]}\n")


def run(app, *args, env=None, ok=True):
    import os
    e = dict(os.environ, **(env or {}))
    r = subprocess.run([str(BIN / app), *map(str, args)], capture_output=True, text=True, timeout=120, env=e)
    if ok:
        assert r.returncode == 0, r.stderr
        return r.stdout
    return r


def test_bfs_app(tmp_path, oracle):
    rp, col, _ = random_graph(5000, 30000, 1)
    g = tmp_path / "g.csr"
    write_csr(g, rp, col)
    for root in (0, 4321):
        out = tmp_path / "lv.bin"
        run("bfs_app", g, out, root)
        assert np.array_equal(np.fromfile(out, np.int32), oracle.bfs_queue(rp, col, root))
    # default root: first vertex of maximum out-degree
    run("bfs_app", g, out)
    root = int(np.argmax(np.diff(rp.astype(np.int64))))
    assert np.array_equal(np.fromfile(out, np.int32), oracle.bfs_queue(rp, col, root))


def test_sssp_app_csr_and_gtgraph(tmp_path, oracle):
    rp, col, _ = random_graph(3000, 15000, 2)
    w = np.random.default_rng(2).integers(1, 1001, len(col)).astype(np.uint32)
    want = oracle.sssp_dijkstra(rp, col, w, 7)
    wantf = np.where(want == 0xFFFFFFFF, np.float32(np.finfo(np.float32).max), want.astype(np.float32))
    g = tmp_path / "g.csr"
    write_csr(g, rp, col, w)
    out = tmp_path / "d.bin"
    run("sssp_app", g, out, 7)
    assert np.array_equal(np.fromfile(out, np.float32), wantf)
    gt = tmp_path / "g.gr"
    write_gtgraph(gt, rp, col, w)
    run("sssp_app", gt, out, 7)
    assert np.array_equal(np.fromfile(out, np.float32), wantf)


def test_pagerank_app_reference_formula(tmp_path, oracle):
    rp, col, _ = random_graph(4000, 40000, 3)
    g = tmp_path / "g.csr"
    write_csr(g, rp, col)
    out = tmp_path / "r.bin"
    run("pagerank_app", g, out, 100)
    got = np.fromfile(out, np.float32)
    want = oracle.pagerank_ref(rp, col, 100)
    # same formula; the combiner's summation order differs from the sequential CPU sum
    np.testing.assert_allclose(got, want, rtol=2e-5)


def test_mbfs_app(tmp_path, oracle):
    rp, col, _ = random_graph(4000, 20000, 4)
    g = tmp_path / "g.csr"
    write_csr(g, rp, col)
    srcs = [0, 17, 17, 3999, 250]
    out = tmp_path / "m.bin"
    run("mbfs_app", g, out, *srcs)
    got = np.fromfile(out, np.uint8).reshape(len(srcs), -1)
    assert np.array_equal(got, oracle.mbfs(rp, col, np.array(srcs, np.uint32)))


@pytest.mark.parametrize("kind", ["int", "float"])
def test_combiner_ops_app(tmp_path, kind):
    """Every user-API combiner operator (the reference's CUDPP ADD/MULTIPLY/MIN/MAX) over
    each vertex's in-edge weights; vertices without in-edges read the init value -1."""
    rng = np.random.default_rng(5)
    n, m = 3000, 6000
    src = rng.integers(0, n, m)
    dst = rng.integers(0, n, m)
    from conftest import csr_from_edges
    w = rng.integers(1, 4, m).astype(np.uint32)   # small weights keep products in range
    rp, col, wc = csr_from_edges(n, np.stack([src, dst], 1), w)
    g = tmp_path / "g.csr"
    write_csr(g, rp, col, wc)
    dt = np.int32 if kind == "int" else np.float32
    indeg = np.bincount(col, minlength=n)
    vals = wc.astype(np.float64)
    for op in ("add", "min", "max", "mul"):
        out = tmp_path / f"{op}.bin"
        run(f"combine_{kind}_app", g, out, op)
        got = np.fromfile(out, dt).astype(np.float64)
        want = np.full(n, -1.0)
        for v in np.flatnonzero(indeg):
            xs = vals[col == v]
            want[v] = {"add": xs.sum(), "min": xs.min(), "max": xs.max(), "mul": np.prod(xs)}[op]
        if kind == "float" and op in ("add", "mul"):
            np.testing.assert_allclose(got, want, rtol=1e-6)
        else:
            assert np.array_equal(got, want), op


def test_edgelist_and_message_functors_app(tmp_path):
    """FunctorHolder<EDGELIST> walks each vertex's out-edges (CSR order) and sends messages;
    FunctorHolder<MESSAGE> rewrites every message slot; the ADD combiner and a VERTEX functor
    collect the result."""
    rng = np.random.default_rng(9)
    n, m = 4000, 20000
    src = rng.integers(0, n, m)
    dst = rng.integers(0, n, m)
    from conftest import csr_from_edges
    w = rng.integers(0, 50, m).astype(np.uint32)
    rp, col, wc = csr_from_edges(n, np.stack([src, dst], 1), w)
    g = tmp_path / "g.csr"
    write_csr(g, rp, col, wc)
    out = tmp_path / "f.bin"
    run("functors_app", g, out)
    got = np.fromfile(out, np.int32)
    wsum, result = got[:n], got[n:]
    rows = np.repeat(np.arange(n), np.diff(rp.astype(np.int64)))
    want_wsum = np.bincount(rows, weights=wc, minlength=n).astype(np.int64)
    want_res = np.bincount(col, weights=2 * wc.astype(np.int64) + 1, minlength=n).astype(np.int64)
    want_res[np.bincount(col, minlength=n) == 0] = -1
    assert np.array_equal(wsum, want_wsum)
    assert np.array_equal(result, want_res)


def _sections(path, count):
    a = np.fromfile(path, np.int32)
    out, i = [], 0
    for _ in range(count):
        k = int(a[i])
        out.append(a[i + 1:i + 1 + k])
        i += 1 + k
    assert i == len(a)
    return out


@pytest.mark.parametrize("seed,hub", [(11, 0), (12, 20000)])
def test_queue_operations_app(tmp_path, oracle, seed, hub):
    """QFunctorHolder<...>::Run for every QAPIType: a BFS built from VQ2EQ + EQ2EQ (a functor
    claiming the destination's level) must give the oracle's levels, and each other operation
    must produce exactly the multiset of items its definition gives (order is unspecified).
    `hub` adds a vertex with that many out-edges, so one activated vertex fills many blocks'
    worth of queue entries."""
    from conftest import csr_from_edges
    rng = np.random.default_rng(seed)
    n, m = 5000, 40000
    edges = np.stack([rng.integers(0, n, m), rng.integers(0, n, m)], 1)
    if hub:
        edges = np.concatenate([edges, np.stack([np.full(hub, 17), rng.integers(0, n, hub)], 1)])
    w = rng.integers(0, 50, len(edges)).astype(np.uint32)
    rp, col, wc = csr_from_edges(n, edges, w)
    g = tmp_path / "g.csr"
    write_csr(g, rp, col, wc)
    out = tmp_path / "q.bin"
    src, light = 17, 3
    run("queues_app", g, out, src, light)
    (levels, reached, even, mark, va2eq, ea2aq, ea2eq, eq2vq, hits, steps) = _sections(out, 10)

    want = oracle.bfs_queue(rp, col, src)
    assert np.array_equal(levels, want)
    deg = np.diff(rp.astype(np.int64))
    esrc = np.repeat(np.arange(n), deg)
    with_edges = want[(want >= 0) & (deg > 0)]
    assert int(steps[0]) == (int(with_edges.max()) + 1 if len(with_edges) else 0)

    def edges_of(vs):
        vs = np.asarray(vs, np.int64)
        if len(vs) == 0:
            return np.zeros(0, np.int64)
        return np.concatenate([np.arange(rp[v], rp[v + 1], dtype=np.int64) for v in vs])

    srt = np.sort
    assert np.array_equal(srt(reached), np.flatnonzero(want >= 0))
    ev = np.flatnonzero((want >= 0) & (want % 2 == 0))
    assert np.array_equal(srt(even), ev)
    assert np.array_equal(mark, np.isin(np.arange(n), ev).astype(np.int32))
    assert np.array_equal(srt(va2eq), srt(edges_of(np.flatnonzero(want == 1))))
    lt = np.flatnonzero(wc < light)
    assert np.array_equal(srt(ea2aq), srt(col[lt].astype(np.int64)))
    want_ea2eq = edges_of(col[lt])
    assert np.array_equal(srt(ea2eq), srt(want_ea2eq))
    odd = want_ea2eq[wc[want_ea2eq] % 2 == 1]
    assert np.array_equal(srt(eq2vq), srt(col[odd].astype(np.int64)))
    assert np.array_equal(hits, np.bincount(esrc[want_ea2eq], minlength=n).astype(np.int32))


def test_bfs_app_loops_and_layouts(tmp_path, oracle):
    """The BFS app's Medusa_Exec has the reference's branch structure (MGMH_EDGE / MG_EDGE /
    EDGE by GPU count and hops).  Its levels equal the oracle's with the device loop
    (Medusa_Run), with the reference's host loop (ResetToExecute / RetriveToExecute per
    superstep), over repeated runs, and built for the MEG and HY layouts; an ELL build is refused
    at InitConfig, as the reference refuses it."""
    rp, col, _ = random_graph(20000, 200000, 21)
    g = tmp_path / "g.csr"
    write_csr(g, rp, col)
    root = 5
    want = oracle.bfs_queue(rp, col, root)
    steps = None
    for app, env in (("bfs_app", {}), ("bfs_app", {"MEDUSA_HOST_LOOP": "1"}),
                     ("bfs_app", {"MEDUSA_APP_REPEAT": "3"}), ("bfs_app", {"MEDUSA_HOST_LOOP": "1", "MEDUSA_APP_REPEAT": "2"}),
                     ("bfs_meg_app", {}), ("bfs_hy_app", {})):
        out = tmp_path / f"{app}_{len(env)}.bin"
        text = run(app, g, out, root, env=env)
        assert np.array_equal(np.fromfile(out, np.int32), want), (app, env)
        s = int(text.split("supersteps=")[1].split()[0])
        assert steps in (None, s), (app, env, text)  # every loop runs the same supersteps
        steps = s
        assert "loop_ms=" in text
    r = run("bfs_ell_app", g, tmp_path / "ell.bin", root, ok=False)
    assert r.returncode != 0 and "ELL" in r.stderr


@pytest.mark.parametrize("env", [{}, {"MEDUSA_HOST_LOOP": "1"}, {"MEDUSA_APP_REPEAT": "2"}])
def test_sssp_and_pagerank_apps_every_loop(tmp_path, oracle, env):
    """SSSP (vote-ended) and PageRank (fixed supersteps) through the device loop, the host
    loop and repeated runs give the same results as the oracle."""
    rp, col, w = random_graph(6000, 48000, 23)
    w = np.random.default_rng(23).integers(1, 100, len(col)).astype(np.uint32)
    g = tmp_path / "g.csr"
    write_csr(g, rp, col, w)
    out = tmp_path / "d.bin"
    run("sssp_app", g, out, 0, env=env)
    want = oracle.sssp_dijkstra(rp, col, w, 0)
    wantf = np.where(want == 0xFFFFFFFF, np.float32(np.finfo(np.float32).max), want.astype(np.float32))
    assert np.array_equal(np.fromfile(out, np.float32), wantf)
    g2 = tmp_path / "g2.csr"
    write_csr(g2, rp, col)
    out = tmp_path / "r.bin"
    text = run("pagerank_app", g2, out, 20, env=env)
    assert "supersteps=20" in text
    np.testing.assert_allclose(np.fromfile(out, np.float32), oracle.pagerank_ref(rp, col, 20), rtol=2e-5)
