"""The graph arrays the library keeps in HBM, read back through medusa_graph_device_arrays:
CSR as given, CSC offsets and sources, and the CSR-to-CSC slot map (the reference's
msgDstID, built by EdgeArray::buildAA), each equal to the oracle's CSC build on RMAT-20 and
on a small graph with duplicate edges, self-loops and vertices without edges."""
import ctypes

import numpy as np
import pytest

import medusa_amd as md
from medusa_amd import workloads as wl
from conftest import csr_from_edges

pytestmark = pytest.mark.gpu
HIP_MEMCPY_DEVICE_TO_HOST = 2


@pytest.fixture(autouse=True)
def _announce(request):
    print(f"\n[case] {request.node.nodeid}", flush=True)


def _d2h(ptr, count):
    """Copies `count` uint32 from device address `ptr` with the HIP runtime the library uses."""
    memcpy = md.lib().hipMemcpy
    memcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    memcpy.restype = ctypes.c_int
    out = np.empty(count, np.uint32)
    if count:
        assert memcpy(out.ctypes.data, ptr, out.nbytes, HIP_MEMCPY_DEVICE_TO_HOST) == 0
    return out


def _check(oracle, rp, col):
    n, m = len(rp) - 1, len(col)
    cptr, csrc, slot = oracle.csc(rp, col)
    with md.Graph(rp, col) as G:
        a = G.device_arrays()
        assert np.array_equal(_d2h(a["row_ptr"], n + 1), rp)
        assert np.array_equal(_d2h(a["col"], m), col)
        assert np.array_equal(_d2h(a["csc_ptr"], n + 1), cptr)
        assert np.array_equal(_d2h(a["csc_src"], m), csrc)
        assert np.array_equal(_d2h(a["msg_slot"], m), slot)
    # the slot map sends CSR edge e (u -> v) to a CSC position of v whose source is u
    src_of_e = np.repeat(np.arange(n, dtype=np.uint32), np.diff(rp.astype(np.int64)))
    assert np.array_equal(csrc[slot], src_of_e)
    assert np.all((slot >= cptr[col]) & (slot < cptr[col + 1]))
    assert np.array_equal(np.sort(slot), np.arange(m, dtype=np.uint32))  # every CSC position once


def test_device_arrays_match_oracle_csc_small(oracle):
    edges = [(0, 1), (0, 1), (2, 2), (1, 0), (4, 1), (4, 3), (1, 4), (4, 1)]
    rp, col, _ = csr_from_edges(6, edges)
    _check(oracle, rp, col)


def test_device_arrays_match_oracle_csc_rmat20(oracle):
    rp, col = wl.rmat_graph()
    _check(oracle, rp, col)
