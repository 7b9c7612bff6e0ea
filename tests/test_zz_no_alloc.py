"""mh_expectation_noalloc (terms from a static device workspace, no device
allocation) against mh_expectation through the C ABI: the same terms bit for
bit, on two calls in a row with the second smaller, at both precisions; and
term sets beyond the static workspace refused with MH_ERR_INVALID."""
import ctypes

import numpy as np
import pytest

from circuits import random_state
from menos_amd import workloads as W

pytestmark = pytest.mark.gpu


def _terms(n, count, seed):
    terms = W.random_paulis(n, count, seed, max_weight=min(4, n))
    # X masks too wide for a tile take the per-mask kernel: cover that path too.
    terms += W.random_paulis(n, 3, seed + 1, max_weight=n, min_weight=min(n, 9))
    return terms + [(0, 0), (0, (1 << n) - 1)]


def _noalloc(hip, d, terms):
    xs = np.array([t[0] for t in terms], np.uint64)
    zs = np.array([t[1] for t in terms], np.uint64)
    out = np.zeros(2 * len(terms), np.float64)
    st = hip.lib().mh_expectation_noalloc(d.ptr, d.n, d.precision, xs.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                          zs.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), len(terms),
                                          out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return st, out.view(np.complex128)


@pytest.mark.parametrize("precision", [0, 1])
def test_noalloc_expectation_equals_the_pooled_entry_point(hip, precision):
    # The larger call first, then a smaller one: the second must not see the
    # first's descriptors or partials left in the shared workspace.
    for n, count, seed in ((20, 40, 11), (13, 6, 12)):
        s = random_state(n, 900 + n)
        d = hip.DeviceState(n, precision)
        d.upload(s.astype(np.complex64) if precision == hip.C64 else s)
        terms = _terms(n, count, seed)
        want = d.expectation(terms)
        st, got = _noalloc(hip, d, terms)
        d.free()
        assert st == hip.MH_OK, hip.lib().mh_last_error()
        assert got.tobytes() == want.tobytes(), (n, precision)


def test_noalloc_refuses_terms_beyond_the_workspace(hip):
    # 64 tile-reduced terms at 28 qubits take 16 MiB of partials, 40 terms at
    # 30 qubits 40 MiB: the second is over the 32 MiB workspace.
    n = 30
    d = hip.DeviceState(n, hip.C64)
    d.init()
    terms = [(1 << (q % 8), 0) for q in range(40)]
    st, _ = _noalloc(hip, d, terms)
    msg = hip.lib().mh_last_error().decode()
    d.free()
    assert st == hip.MH_ERR_INVALID and "static workspace" in msg, (st, msg)
