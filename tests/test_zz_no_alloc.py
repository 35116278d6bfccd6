"""medusa_bfs_noalloc, the BFS entry point that allocates no device memory (its frontier
queues, bitmap and control block sit in a static per-GPU array), checked against medusa_bfs
and the oracle through the C ABI: two calls in a row, the second on a smaller graph, so the
second run starts from the scratch contents the first one left."""
import numpy as np
import pytest

import medusa_amd as md
from medusa_amd import workloads as wl
from conftest import random_graph

pytestmark = pytest.mark.gpu

MAX_VERTICES = 1 << 21  # MEDUSA_NOALLOC_MAX_VERTICES


def test_bfs_noalloc_matches_bfs(oracle):
    import torch

    cases = [wl.rmat_graph(scale=20, edge_factor=16, seed=2), random_graph(5000, 40000, 3)[:2]]
    for i, (rp, col) in enumerate(cases):
        n = len(rp) - 1
        src = wl.max_outdeg_vertex(rp)
        print(f"[case] noalloc call {i}: n={n} m={len(col)} source={src}", flush=True)
        with md.Graph(rp, col) as G:
            want, st_want = G.bfs(src)
            out = torch.full((n,), 7, dtype=torch.int32, device="cuda")
            free0 = torch.cuda.mem_get_info()[0]
            st = G.bfs_noalloc(src, out)
            free1 = torch.cuda.mem_get_info()[0]
            got = out.cpu().numpy()
            assert np.array_equal(got, want), i
            assert np.array_equal(got, oracle.bfs_queue(rp, col, src)), i
            assert st.supersteps == st_want.supersteps and st.reached == st_want.reached
            assert st.edges_traversed == st_want.edges_traversed
            print(f"[case] noalloc call {i}: device free bytes before/after {free0} {free1}", flush=True)
            assert free1 == free0, "the no-allocation entry point changed the device's free memory"


def test_bfs_noalloc_rejects_what_it_cannot_serve():
    import torch

    n = MAX_VERTICES + 1
    with md.Graph(np.zeros(n + 1, np.uint32), np.zeros(0, np.uint32)) as G:
        out = torch.empty(n, dtype=torch.int32, device="cuda")
        with pytest.raises(md.MedusaError, match="scratch limit"):
            G.bfs_noalloc(0, out)
    rp, col, _ = random_graph(100, 400, 1)
    with md.Graph(rp, col) as G:
        with pytest.raises(md.MedusaError, match="device memory"):
            G.bfs_noalloc(0, np.zeros(100, np.int32))  # host memory
        with pytest.raises(md.MedusaError, match="out of range"):
            G.bfs_noalloc(100, torch.empty(100, dtype=torch.int32, device="cuda"))
