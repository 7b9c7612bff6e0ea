import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TESTS = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, TESTS):
    if p not in sys.path:
        sys.path.insert(0, p)

import oracle  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and the built HIP library")


@pytest.fixture(scope="session")
def orc():
    if not os.path.exists(oracle.LIB_PATH):
        oracle.build()
    return oracle.Oracle()


@pytest.fixture(scope="session")
def ref():
    if not oracle.Reference.available():
        pytest.skip("oracle/_ref is not built (the reference sources were not present at build time)")
    return oracle.Reference()


@pytest.fixture(scope="session")
def golden():
    with open(os.path.join(TESTS, "golden", "ref_vectors.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def golden_states():
    return dict(np.load(os.path.join(TESTS, "golden", "ref_states.npz"), allow_pickle=False))


@pytest.fixture(scope="session")
def hip():
    """The HIP library, on a machine with a GPU. Fails (does not skip) when the
    library is missing or the device is unusable: there is no fallback path."""
    from menos_amd import hip as H
    info = H.device_info()
    assert info.arch.decode().startswith("gfx950"), info.arch
    return H


def circuit_from_spec(spec):
    from menos_amd import workloads as W
    c = W.Circuit(spec["num_qubits"], spec.get("num_clbits", 0), spec.get("global_phase", 0.0))
    c.ops = [(int(op), tuple(int(x) for x in q), tuple(float(x) for x in p)) for op, q, p in spec["ops"]]
    return c


def cpx(pairs):
    """[(re, im), ...] as complex128, keeping the sign of every zero."""
    a = np.ascontiguousarray(np.asarray(pairs, np.float64).reshape(-1, 2))
    return a.view(np.complex128).reshape(-1)
