"""Writes tests/golden/golden.json: tiny hand-worked graphs with expected outputs.

BFS levels and SSSP distances below were worked out by hand (and are asserted
here against a brute-force recomputation).  PageRank values are computed with
exact rational arithmetic (fractions.Fraction) from the two stated formulas:
  engine:    r0 = 1/n, r' = (1-d)/n + d*(sum_{u->v} r(u)/outdeg(u) + D/n), D = dangling mass
  reference: r0 = 1,   r' = 0.15 + 0.85*sum_{u->v} r(u)/outdeg(u)   (MPG/UserOperators.h:16-43,
             MPG/main.cu:72; dangling mass dropped)
No reference code is run (the reference is CUDA and cannot be built here).
"""
import json
import pathlib
from fractions import Fraction

HERE = pathlib.Path(__file__).resolve().parent

GRAPHS = {
    # 0->1 twice (duplicate), self-loop 2->2, cycle 3->4->5->3, separate component 6->7
    "g8": {
        "n": 8,
        "edges": [[0, 1], [0, 2], [1, 3], [2, 3], [3, 4], [4, 5], [5, 3], [6, 7], [2, 2], [0, 1]],
        "weights": [4, 1, 1, 5, 2, 3, 1, 7, 9, 2],
        "bfs": {"0": [0, 1, 1, 2, 3, 4, -1, -1], "6": [-1, -1, -1, -1, -1, -1, 0, 1],
                "3": [-1, -1, -1, 0, 1, 2, -1, -1]},
        "sssp": {"0": [0, 2, 1, 3, 5, 8, None, None], "2": [None, None, 0, 5, 7, 10, None, None]},
    },
    "two": {"n": 2, "edges": [[0, 1]], "weights": [5], "bfs": {"0": [0, 1], "1": [-1, 0]},
            "sssp": {"0": [0, 5], "1": [None, 0]}},
    # star: leaves 1..5 point to the centre 0; centre points back to leaf 1
    "star": {"n": 6, "edges": [[1, 0], [2, 0], [3, 0], [4, 0], [5, 0], [0, 1]], "weights": [1, 2, 3, 4, 5, 6],
             "bfs": {"0": [0, 1, -1, -1, -1, -1], "3": [1, 2, -1, 0, -1, -1]},
             "sssp": {"0": [0, 6, None, None, None, None], "5": [5, 11, None, None, None, 0]}},
    "single": {"n": 1, "edges": [], "weights": [], "bfs": {"0": [0]}, "sssp": {"0": [0]}},
}


def brute_levels(n, edges, src):
    # repeated relaxation with unit weights
    INF = 10 ** 9
    d = [INF] * n
    d[src] = 0
    for _ in range(n + 1):
        for u, v in edges:
            if d[u] + 1 < d[v]:
                d[v] = d[u] + 1
    return [x if x < INF else -1 for x in d]


def brute_dist(n, edges, w, src):
    INF = 10 ** 18
    d = [INF] * n
    d[src] = 0
    for _ in range(n + 1):
        for (u, v), c in zip(edges, w):
            if d[u] + c < d[v]:
                d[v] = d[u] + c
    return [x if x < INF else None for x in d]


def pagerank_engine(n, edges, d, steps):
    d = Fraction(d)
    outdeg = [0] * n
    for u, _ in edges:
        outdeg[u] += 1
    r = [Fraction(1, n)] * n
    for _ in range(steps):
        D = sum(r[u] for u in range(n) if outdeg[u] == 0)
        s = [Fraction(0)] * n
        for u, v in edges:
            s[v] += r[u] / outdeg[u]
        r = [(1 - d) / n + d * (s[v] + D / n) for v in range(n)]
    return r


def pagerank_reference(n, edges, steps):
    outdeg = [0] * n
    for u, _ in edges:
        outdeg[u] += 1
    r = [Fraction(1)] * n
    for _ in range(steps):
        s = [Fraction(0)] * n
        for u, v in edges:
            s[v] += r[u] / outdeg[u]
        r = [Fraction(15, 100) + Fraction(85, 100) * s[v] for v in range(n)]
    return r


def main():
    out = {}
    for name, g in GRAPHS.items():
        n, edges, w = g["n"], g["edges"], g["weights"]
        for s, lv in g["bfs"].items():
            assert brute_levels(n, edges, int(s)) == lv, (name, s)
        for s, dist in g["sssp"].items():
            assert brute_dist(n, edges, w, int(s)) == dist, (name, s)
        rec = dict(g)
        rec["pagerank_engine"] = {
            str(steps): [float(x) for x in pagerank_engine(n, edges, Fraction(85, 100), steps)] for steps in (1, 2, 30)
        }
        rec["pagerank_reference"] = {
            str(steps): [float(x) for x in pagerank_reference(n, edges, steps)] for steps in (1, 2, 30)
        }
        out[name] = rec
    # hand check of the two-node graph after one engine step: r(0) = 0.075 + 0.85*0.25, r(1) = 0.075 + 0.85*0.75
    assert out["two"]["pagerank_engine"]["1"] == [0.2875, 0.7125]
    (HERE / "golden.json").write_text(json.dumps(out, indent=1, sort_keys=True))
    print("wrote", HERE / "golden.json")


if __name__ == "__main__":
    main()
