import pathlib
import sys

import numpy as np
import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C ABI on the GPU)")
    config.addinivalue_line("markers", "slow: full-size workload")


def csr_from_edges(n, edges, weights=None):
    """Stable CSR from a python edge list (test helper, independent of the library)."""
    edges = np.asarray(edges, dtype=np.int64).reshape(-1, 2)
    order = np.argsort(edges[:, 0], kind="stable")
    row_ptr = np.zeros(n + 1, np.uint32)
    np.add.at(row_ptr, edges[:, 0] + 1, 1)
    row_ptr = np.cumsum(row_ptr).astype(np.uint32)
    col = edges[order, 1].astype(np.uint32)
    w = None if weights is None else np.asarray(weights, np.uint32)[order]
    return row_ptr, col, w


def random_graph(n, m, seed, self_loops=True):
    rng = np.random.default_rng(seed)
    src = rng.integers(0, n, m)
    dst = rng.integers(0, n, m)
    if not self_loops:
        keep = src != dst
        src, dst = src[keep], dst[keep]
    return csr_from_edges(n, np.stack([src, dst], 1))


@pytest.fixture(scope="session")
def oracle():
    import oracle as orc
    orc.lib()
    return orc


# Trained policies of the learned scheduler, shared by the GPU tests (trained once per session
# on the GPU, from graphs other than the ones the tests run them on).
@pytest.fixture(scope="session")
def direction_policy():
    import policies
    pol = policies.train_direction_policy()
    yield pol
    pol.close()


@pytest.fixture(scope="session")
def mbfs_policy():
    import policies
    pol = policies.train_mbfs_policy()
    yield pol
    pol.close()


@pytest.fixture(scope="session")
def delta_policy():
    import policies
    pol = policies.train_delta_policy()
    yield pol
    pol.close()


@pytest.fixture(scope="session")
def variant_data():
    import policies
    return policies.variant_training_set()


@pytest.fixture(scope="session")
def variant_policy(variant_data):
    import policies
    pol = policies.train_variant_policy(*variant_data)
    yield pol
    pol.close()
