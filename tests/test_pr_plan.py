"""The binned PageRank planner on the host (medusa_amd/csrc/pr_bins.h, compiled here with g++):
send tiles enlarged to fit one round of resident send workgroups only when that takes at most
twice the default tile, and the combine's short last round of bins re-planned into quarter-size
bins.  Every plan must pass the planner's own index checks (check_pr_bins) and place every
message; the GPU tests check that the ranks stay bit-identical."""
import json
import pathlib
import shutil
import subprocess

import numpy as np
import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
TILE = 16384  # kPrTileEdges
BIN = 8192  # kPrBinValues


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not found")
    exe = tmp_path_factory.mktemp("plan") / "pr_plan_check"
    subprocess.run([gxx, "-O2", "-std=c++17", "-iquote", str(ROOT / "medusa_amd" / "csrc"), "-o", str(exe),
                    str(ROOT / "tests" / "pr_plan_check.cpp")], check=True, timeout=120)
    return exe


def _internal_csc(parts):
    """CSC of packed graphs (lists of (src, dst) arrays, one per graph) in the PageRank internal
    order: per graph, vertices by in-degree descending, then id."""
    begin = [0]
    srcs, dsts = [], []
    for n, src, dst in parts:
        srcs.append(src + begin[-1])
        dsts.append(dst + begin[-1])
        begin.append(begin[-1] + n)
    src, dst, N = np.concatenate(srcs), np.concatenate(dsts), begin[-1]
    indeg = np.bincount(dst, minlength=N)
    perm = np.concatenate([np.arange(b, e)[np.lexsort((np.arange(b, e), -indeg[b:e]))]
                           for b, e in zip(begin[:-1], begin[1:])])
    iperm = np.empty(N, np.int64)
    iperm[perm] = np.arange(N)
    d, s = iperm[dst], iperm[src]
    order = np.lexsort((s, d))
    cptr = np.zeros(N + 1, np.uint32)
    cptr[1:] = np.cumsum(np.bincount(d, minlength=N))
    return cptr, s[order].astype(np.uint32), np.array(begin, np.uint32)


def _plan(checker, tmp_path, csc, send_slots, combine_slots):
    names = []
    for k, a in enumerate(csc):
        p = tmp_path / f"a{k}.u32"
        a.astype(np.uint32).tofile(p)
        names.append(str(p))
    r = subprocess.run([str(checker), *names, str(send_slots), str(combine_slots)], capture_output=True, text=True,
                       timeout=120, check=True)
    out = json.loads(r.stdout)
    assert out["err"] == "", out
    assert out["check0"] == "" and out["check1"] == "", out
    assert out["nmsg0"] == out["nmsg1"] == int(csc[0][-1]), out
    return out


@pytest.fixture(scope="module")
def random_csc():
    rng = np.random.default_rng(21)
    n, m = 65536, 1_000_000
    return _internal_csc([(n, rng.integers(0, n, m), rng.integers(0, n, m))])


def test_send_tiles_enlarged_to_one_round(checker, tmp_path, random_csc):
    base = _plan(checker, tmp_path, random_csc, 0, 0)
    assert base["tile_edges1"] == TILE and base["tiles1"] == base["tiles0"]
    slots = base["tiles0"] - 5  # a few tiles more than one round
    out = _plan(checker, tmp_path, random_csc, slots, 0)
    assert TILE < out["tile_edges1"] <= 2 * TILE and out["tile_edges1"] % 4 == 0
    assert out["tiles1"] <= slots and out["max_tile1"] <= out["tile_edges1"]


def test_send_tiles_kept_when_round_fits_or_tiles_would_grow_past_twice(checker, tmp_path, random_csc):
    base = _plan(checker, tmp_path, random_csc, 0, 0)
    fits = _plan(checker, tmp_path, random_csc, base["tiles0"] + 1, 0)
    assert fits["tile_edges1"] == TILE and fits["tiles1"] == base["tiles0"]
    far = _plan(checker, tmp_path, random_csc, base["tiles0"] // 3, 0)  # would need tiles of about 3x
    assert far["tile_edges1"] == TILE and far["tiles1"] == base["tiles0"]


def test_short_last_combine_round_split_into_quarter_bins(checker, tmp_path, random_csc):
    base = _plan(checker, tmp_path, random_csc, 0, 0)
    b0 = base["bins0"]
    slots = b0 - 3  # a last round of 3 bins
    out = _plan(checker, tmp_path, random_csc, 0, slots)
    assert out["bins1"] > b0
    assert out["tail_bins"] == out["bins1"] - slots and out["tail_bins"] >= 3 * 2
    assert out["tail_max_slots"] <= BIN // 4 and out["tail_over_quarter"] == 0


def test_long_last_combine_round_kept(checker, tmp_path, random_csc):
    b0 = _plan(checker, tmp_path, random_csc, 0, 0)["bins0"]
    slots = (2 * b0) // 3  # last round of about half the slots: more than a quarter
    assert b0 % slots > slots // 4
    out = _plan(checker, tmp_path, random_csc, 0, slots)
    assert out["bins1"] == b0


def test_quarter_bins_hold_larger_items_alone(checker, tmp_path):
    """A packed batch whose last graph (in the tail) has wave vertices of in-degree 2100-4000 and
    lane groups of in-degree 40-64, whose slots exceed a quarter-size bin: each gets a bin of its
    own, within the full bin size."""
    rng = np.random.default_rng(13)
    n0, m0 = 65536, 1_000_000
    big = (n0, rng.integers(0, n0, m0), rng.integers(0, n0, m0))
    n1 = 3000
    src, dst = [], []
    for v, deg in ((0, 4000), (1, 2100), (2, 3000)):
        src.append(rng.integers(0, n1, deg))
        dst.append(np.full(deg, v))
    for v in range(3, 200):
        d = int(rng.integers(40, 65))
        src.append(rng.integers(0, n1, d))
        dst.append(np.full(d, v))
    src.append(rng.integers(0, n1, 5000))
    dst.append(rng.integers(200, n1, 5000))
    csc = _internal_csc([big, (n1, np.concatenate(src), np.concatenate(dst))])
    b0 = _plan(checker, tmp_path, csc, 0, 0)["bins0"]
    slots = b0 - 4  # the last graph's bins form the short last round
    out = _plan(checker, tmp_path, csc, 0, slots)
    assert out["bins1"] > b0
    assert out["tail_over_quarter"] >= 3 and out["tail_max_slots"] <= BIN
    assert out["tail_over_quarter_shared"] == 0
