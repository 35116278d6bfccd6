"""Benchmark: PageRank on an RMAT scale-20 graph (1,048,576 vertices, 16,777,216 arcs),
30 supersteps, FP32 with compensated combine, on one MI355X.

A step is one full PageRank run (30 supersteps) over the HBM-resident graph; the
graph is built and uploaded before the timed region.  Prints one JSON line with
GTEPS (30*m per run / time), the roofline of the dominant kernel (the pull
superstep) measured with HIP events on the library's stream, and the reference
CPU path (message-array push + per-vertex segment sum, OpenMP) timed on the
host cores.
"""
import argparse
import json
import os
import pathlib
import re
import sys
import time

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
STEPS_PR = 30


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    p.add_argument("--scale", type=int, default=20)
    p.add_argument("--profile", action="store_true",
                   help="short run for hardware counters: 3 PageRank runs (non-graph launches) and one "
                        "256 MiB device copy as the byte-count calibration; prints no JSON")
    return p.parse_args()


def profile_pass(G, out, torch):
    """Kernels for a counter pass: the pull kernel 90 times, then a plain copy of known size."""
    for _ in range(3):
        G.pagerank(0.85, STEPS_PR, out=out)
    torch.cuda.synchronize()
    a = torch.ones(64 << 20, dtype=torch.float32, device="cuda")  # 256 MiB
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    print("profile pass done: 90 pull launches, copy of 268435456 bytes", flush=True)


def traffic_from_profiles():
    """HBM bytes per launch of the pull kernel from the committed rocprofv3 PMC summary."""
    f = ROOT / "profiles" / "pagerank_pull_rmat20_traffic.md"
    if not f.exists():
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", f.read_text())
    return float(m.group(1)) if m else None


def main():
    a = parse()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this engine runs on one GPU; multi-GPU partitioning is not part of this build",
              file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    import medusa_amd as md
    from medusa_amd import workloads as wl

    if md.device_count() < 1:
        print("bench.py: no GPU visible", file=sys.stderr)
        sys.exit(1)
    rp, col = wl.rmat_graph(scale=a.scale, edge_factor=16, seed=2)
    n, m = len(rp) - 1, len(col)
    G = md.Graph(rp, col)
    out = torch.empty(n, dtype=torch.float32, device="cuda")
    if a.profile:
        profile_pass(G, out, torch)
        G.close()
        return

    # warmup (also builds the pull layout and the captured superstep graph)
    for _ in range(max(1, a.warmup)):
        G.pagerank(0.85, STEPS_PR, out=out, use_graph=1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    dev_ms = 0.0
    for _ in range(a.steps):
        _, st = G.pagerank(0.85, STEPS_PR, out=out, use_graph=1)
        dev_ms += st.device_ms
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    wall = t1 - t0
    ms_per_step = wall / a.steps * 1e3
    gteps = a.steps * STEPS_PR * m / wall / 1e9
    rank_sum = float(out.double().sum().item())

    # roofline: HIP events around every pull-kernel launch on the library's stream
    kms, launches, kbytes = 0.0, 0, 0
    for _ in range(3):
        _, st = G.pagerank(0.85, STEPS_PR, out=out, timing=1)
        kms += st.kernel_ms
        launches += st.kernel_launches
        kbytes += st.kernel_bytes
    avg_launch_s = kms / launches / 1e3
    bytes_per_launch = kbytes / launches
    achieved = bytes_per_launch / avg_launch_s / 1e9
    roofline = {
        "bound": "hbm",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": traffic_from_profiles(),
        "kernel": "k_pr_main<false> (pull superstep)",
        "kernel_avg_us": round(avg_launch_s * 1e6, 2),
        "alg_bytes_per_launch": int(bytes_per_launch),
    }

    cpu = None
    if not a.no_cpu:
        import oracle

        threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
        steps_cpu = 3
        oracle.pagerank_ref(rp, col, 1, threads)  # first touch
        reps, tc = 0, 0.0
        while reps < 3 and tc < 10.0:
            c0 = time.perf_counter()
            oracle.pagerank_ref(rp, col, steps_cpu, threads)
            tc += time.perf_counter() - c0
            reps += 1
        cpu = {
            "value": round(reps * steps_cpu * m / tc / 1e9, 4),
            "unit": "GTEPS",
            "cores": threads,
            "kind": "port",
            "sample": f"reference CPU path restated (message-array push + per-vertex segment sum, OpenMP), "
                      f"same RMAT-20 graph, {reps}x{steps_cpu} of the 30 supersteps",
        }

    line = {
        "metric": "GTEPS",
        "value": round(gteps, 3),
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"pagerank_rmat{a.scale}",
            "vertices": n,
            "arcs": m,
            "supersteps": STEPS_PR,
            "damping": 0.85,
            "graph": "RMAT a/b/c/d=0.45/0.15/0.15/0.25, +-10% per-level noise, seed 2",
        },
        "device_ms_per_step": round(dev_ms / a.steps, 4),
        "hbm_gbs_alg": round(st.alg_bytes / (dev_ms / a.steps / 1e3) / 1e9, 1) if dev_ms else None,
        "rank_sum": round(rank_sum, 7),
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    G.close()


if __name__ == "__main__":
    main()
