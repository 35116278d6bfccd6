"""Host-side checks that need no GPU: the C-ABI library loads and exports every
declared symbol, generators are deterministic and follow their distributions,
the CSR builder is a stable sort, and argument errors come back as status codes."""
import ctypes
import pathlib
import re

import numpy as np
import pytest

import medusa_amd as md
from conftest import csr_from_edges

HEADER = pathlib.Path(__file__).resolve().parent.parent / "include" / "medusa_c.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(medusa_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = md.lib()
    decl = declared_functions()
    assert len(decl) >= 19
    for name in decl:
        assert hasattr(L, name), name
    assert set(decl) == set(md.EXPORTED)
    assert b"gfx950" in L.medusa_version()


def test_struct_layouts_match_header():
    # medusa_opts: 6 u32/i32 fields, a pointer, use_graph, 7 reserved
    assert ctypes.sizeof(md.Opts) == 6 * 4 + 8 + 4 + 7 * 4
    assert md.Opts.policy.offset == 24
    assert ctypes.sizeof(md.Stats) == 8 * 2 + 8 * 4 + 4 * 4 + 8 + 4 * 8


def test_generators_deterministic():
    a = md.gen_erdos_renyi(1000, 20000, 9)
    b = md.gen_erdos_renyi(1000, 20000, 9)
    c = md.gen_erdos_renyi(1000, 20000, 10)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert not np.array_equal(a[0], c[0])
    r1 = md.gen_rmat(12, 100_000, seed=2)
    r2 = md.gen_rmat(12, 100_000, seed=2)
    assert np.array_equal(r1[0], r2[0]) and np.array_equal(r1[1], r2[1])


def test_erdos_renyi_properties():
    n, m = 4096, 200_000
    src, dst = md.gen_erdos_renyi(n, m, 1)
    assert np.all(src != dst)
    assert np.all(src < n) and np.all(dst < n)
    # about m/n self-loops were dropped
    assert m - 4 * m // n < len(src) <= m
    # uniform endpoints: each vertex's out-degree is near m/n
    deg = np.bincount(src, minlength=n)
    assert abs(deg.mean() - len(src) / n) < 1e-9
    assert deg.std() < 2 * np.sqrt(m / n)


def test_rmat_properties():
    scale, m = 14, 16 << 14
    src, dst = md.gen_rmat(scale, m, 0.45, 0.15, 0.15, 0.1, seed=2)
    assert len(src) == m and src.max() < (1 << scale) and dst.max() < (1 << scale)
    # row bit = 1 with probability c+d = 0.4 at every level (noise averages out)
    for bit in (0, 7, 13):
        frac = ((src >> bit) & 1).mean()
        assert abs(frac - 0.40) < 0.01, (bit, frac)
        frac = ((dst >> bit) & 1).mean()  # column bit 1 with probability b+d = 0.4
        assert abs(frac - 0.40) < 0.01, (bit, frac)
    deg = np.bincount(src, minlength=1 << scale)
    # vertex 0 takes row bit 0 at every level: expected out-degree m * (a+b)^scale
    expect0 = m * 0.6 ** scale
    assert abs(deg[0] - expect0) < 0.3 * expect0, (deg[0], expect0)
    assert deg.max() > 8 * deg.mean()  # skewed


def test_lattice_properties():
    R, C = 60, 80
    src, dst, w = md.gen_lattice(R, C, 0.7, 0.02, 1, 1000, seed=3)
    assert len(src) % 2 == 0
    # arcs come in reverse pairs with equal weights
    assert np.array_equal(src[0::2], dst[1::2]) and np.array_equal(dst[0::2], src[1::2])
    assert np.array_equal(w[0::2], w[1::2])
    assert w.min() >= 1 and w.max() <= 1000
    rs, cs, rd, cd = src // C, src % C, dst // C, dst % C
    step = np.abs(rs.astype(int) - rd) + np.abs(cs.astype(int) - cd)
    assert set(np.unique(step)) <= {1, 2}
    assert np.all(np.abs(rs.astype(int) - rd) <= 1) and np.all(np.abs(cs.astype(int) - cd) <= 1)
    lattice_edges = (step == 1).sum() // 2
    possible = R * (C - 1) + (R - 1) * C
    assert abs(lattice_edges / possible - 0.7) < 0.03
    diag = (step == 2).sum() // 2
    assert abs(diag / ((R - 1) * C) - 0.02 * (C - 1) / C) < 0.01


def test_chung_lu_properties():
    n, m = 200_000, 420_000
    src, dst = md.gen_chung_lu(n, m, 2.0, 100_000, 0.4, seed=5)
    deg = np.bincount(src, minlength=n)
    zero = (deg == 0).mean()
    assert 0.55 < zero < 0.68, zero
    assert deg.max() <= m and deg.max() > 1000
    assert np.all(dst < n)


def test_csr_from_arcs_is_stable_sort():
    rng = np.random.default_rng(0)
    n, m = 300, 5000
    src = rng.integers(0, n, m).astype(np.uint32)
    dst = rng.integers(0, n, m).astype(np.uint32)
    w = rng.integers(1, 100, m).astype(np.uint32)
    rp, col, wo = md.csr_from_arcs(n, src, dst, w)
    rp2, col2, w2 = csr_from_edges(n, np.stack([src, dst], 1), w)
    assert np.array_equal(rp, rp2) and np.array_equal(col, col2) and np.array_equal(wo, w2)


def test_argument_errors_are_status_codes():
    with pytest.raises(md.MedusaError) as e:
        md.csr_from_arcs(3, [0, 5], [1, 1])
    assert e.value.code == -1 and "out of range" in str(e.value)
    with pytest.raises(md.MedusaError) as e:
        md.Graph(np.array([0, 2, 1], np.uint32), np.array([0, 1], np.uint32))
    assert e.value.code == -1
    with pytest.raises(md.MedusaError) as e:
        md.Graph(np.array([0, 1], np.uint32), np.array([7], np.uint32))
    assert e.value.code == -1
    L = md.lib()
    st = md.Stats()
    rc = L.medusa_pagerank(None, 0.85, 3, None, None, ctypes.byref(st))
    assert rc == -1
    assert L.medusa_graph_destroy(None) == 0


def test_no_gpu_is_an_error_not_a_fallback():
    if md.device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(md.MedusaError) as e:
        md.Graph(np.array([0, 1, 1], np.uint32), np.array([1], np.uint32))
    assert e.value.code == -2


def test_pack_graphs_host_logic():
    a = (np.array([0, 1, 2], np.uint32), np.array([1, 0], np.uint32), np.array([5, 6], np.uint32))
    b = (np.array([0, 0], np.uint32), np.zeros(0, np.uint32), np.zeros(0, np.uint32))
    c = (np.array([0, 2, 2, 3], np.uint32), np.array([1, 2, 0], np.uint32), np.array([7, 8, 9], np.uint32))
    rp, col, w, begin = md.pack_graphs([a, b, c])
    assert list(begin) == [0, 2, 3, 6]
    assert list(rp) == [0, 1, 2, 2, 4, 4, 5]
    assert list(col) == [1, 0, 4, 5, 3]
    assert list(w) == [5, 6, 7, 8, 9]
    with pytest.raises(md.MedusaError):
        md.pack_graphs([a, b[:2]])  # weights on some graphs only


def test_delta_samples_charge_the_rest_of_the_run():
    """Width-policy rows: each phase is charged the run's time from that phase to the end per
    vertex still to settle.  Hand-worked trace: phase ms 4, 1, 1 and settled 2, 8, 10 give
    6 ms / 20, 2 ms / 18 and 1 ms / 10 vertices; a record without a width is dropped."""
    tr = np.zeros((4, md.TRACE_FIELDS))
    tr[:, 5], tr[:, 6] = 1000, 3000          # n, m
    tr[:, 2] = 100                           # mean weight
    tr[:, 3] = [0, 400, 800, 1200]           # thresholds
    tr[:, 4] = [2, 8, 10, 0]                 # settled
    tr[:, 7] = [400, 400, 400, 0]            # widths (the last record has none)
    tr[:, 8] = [4.0, 1.0, 1.0, 0.5]          # ms
    X, y = md.delta_samples(tr)
    assert X.shape == (3, 16) and y.dtype == np.float32
    want = np.log2([6.5e6 / 20, 2.5e6 / 18, 1.5e6 / 10])
    assert np.allclose(y, want, rtol=1e-6), (y, want)
    for j in range(3):
        assert np.array_equal(X[j], md.policy_features(md.POLICY_DELTA, tr[j], tr[j, 7]))


def test_delta_run_sample_and_candidates():
    """A run's single training row is row 0 of the per-phase rows (the whole run's cost per
    settled vertex); the candidate widths are mean x 2^k, k = -2..8, clamped below 2^32 - 1."""
    tr = np.zeros((3, md.TRACE_FIELDS))
    tr[:, 5], tr[:, 6] = 1000, 3000
    tr[:, 2] = 100
    tr[:, 3] = [0, 400, 800]
    tr[:, 4] = [5, 5, 10]
    tr[:, 7] = 400
    tr[:, 8] = [1.0, 2.0, 1.0]
    x, y = md.delta_run_sample(tr)
    assert np.array_equal(x, md.policy_features(md.POLICY_DELTA, tr[0], 400))
    assert np.isclose(y, np.log2(4e6 / 20))
    assert md.delta_candidates(100) == [25, 50, 100, 200, 400, 800, 1600, 3200, 6400, 12800, 25600]
    assert md.delta_candidates(1)[:3] == [1, 1, 1]
    assert max(md.delta_candidates(2 ** 28)) == 4294967294
