"""Batched execution (mh_batch_execute) against the oracle and the
single-circuit engine: complex128 states bit for bit, expectation terms equal
to the single engine's bit for bit, complex64 within the stated error."""
import numpy as np
import pytest

from circuits import diagonal_heavy_circuit, random_mixed_circuit, random_unitary
from menos_amd import workloads as W

pytestmark = pytest.mark.gpu


def _mixed_batch(count, seed):
    rng = np.random.default_rng(seed)
    circuits, terms = [], []
    for i in range(count):
        n = int(rng.integers(10, 15))
        # Blocks of at most two qubits (wider ones: test_batch_wide_and_empty_circuits).
        if i % 4 == 0:
            c = W.random_circuit(n, int(rng.integers(1, 5)), seed + i)
        elif i % 4 == 1:
            c = W.ansatz_circuit(n, int(rng.integers(1, 4)), rng.uniform(-np.pi, np.pi, 2 * n * 3).tolist())
        elif i % 4 == 2:
            c = diagonal_heavy_circuit(n, 80, seed + i)
        else:
            c = random_mixed_circuit(n, 60, seed + i, max_unitary=2, max_arity=2)
        circuits.append(c)
        terms.append(W.random_paulis(n, int(rng.integers(0, 10)), seed + i, max_weight=4))
    return circuits, terms


def _unitary_only(c):
    """The circuit without measurements (the batch path reduces terms only)."""
    out = W.Circuit(c.num_qubits, 0, c.global_phase)
    out.ops = [op for op in c.ops if op[0] != W.MEASURE]
    return out


@pytest.mark.parametrize("fusion", [0, 1])
def test_batch_states_and_terms(orc, hip, fusion):
    circuits, terms = _mixed_batch(12, 300 + fusion)
    circuits = [_unitary_only(c) for c in circuits]
    r = hip.batch_execute(circuits, terms, fusion=fusion, statevectors=True)
    for i, c in enumerate(circuits):
        want, _, e = orc.execute(c, fuse_gates=bool(fusion), terms=terms[i])
        assert r.statevectors[i].tobytes() == want.tobytes(), i
        single = hip.HipDenseEngine().execute(c, fusion=fusion, terms=terms[i])
        if terms[i]:
            assert r.expectation[i].tobytes() == single.expectation.tobytes(), i
            assert np.max(np.abs(r.expectation[i] - e)) <= 1e-12, i
    assert r.rounds == 1


def test_batch_rounds_give_the_same_bits(hip):
    circuits, terms = W.batch_workload(40, 11, qubits=(10, 13), depth=(2, 5))
    one = hip.batch_execute(circuits, terms)
    many = hip.batch_execute(circuits, terms, round_bytes=1 << 20)
    assert many.rounds > 1
    for a, b in zip(one.expectation, many.expectation):
        assert a.tobytes() == b.tobytes()
    assert np.all(np.diff(many.completion_seconds) >= 0)


def test_batch_groups_give_the_single_engine_bits(orc, hip):
    """A round of more than 64 circuits goes to the device in growing groups
    (host planning of one group while the device runs the one before), which
    alternate between two kernel streams: every circuit's terms equal the
    single engine's bit for bit, in every group."""
    circuits, terms = W.batch_workload(90, 17, qubits=(10, 12), depth=(1, 4))
    r = hip.batch_execute(circuits, terms, statevectors=True)
    assert r.rounds == 1
    eng = hip.HipDenseEngine()
    for i in range(90):
        single = eng.execute(circuits[i], terms=terms[i], statevector=False, targets=[], width=0)
        assert r.expectation[i].tobytes() == single.expectation.tobytes(), i
        want, _, _ = orc.execute(circuits[i], fuse_gates=True)
        assert r.statevectors[i].tobytes() == want.tobytes(), i


def test_batch_complex64_error(hip):
    circuits, terms = W.batch_workload(48, 12, qubits=(11, 15), depth=(4, 12))
    ref = hip.batch_execute(circuits, terms, precision=hip.C128)
    low = hip.batch_execute(circuits, terms, precision=hip.C64)
    err = max(float(np.max(np.abs(a - b))) for a, b in zip(ref.expectation, low.expectation))
    assert err <= 1e-5, err


def test_batch_refuses_small_states(hip):
    c = W.random_circuit(8, 2, 1)
    with pytest.raises(hip.HipError):
        hip.batch_execute([c], [[(0, 1)]])


def test_batch_wide_and_empty_circuits(orc, hip):
    """States whose first pass is not a window pass (a first block of three
    or more qubits, on the per-group kernel, or no gate at all) start from an
    initialisation kernel; the others generate their initial state in their
    first window pass. Every state and term is checked, with global phases."""
    circuits, terms = [], []
    for i in range(10):
        n = 10 + i % 4
        if i % 3 == 0:
            c = random_mixed_circuit(n, 50, 900 + i, max_unitary=4, max_arity=3)
            u = random_unitary(3, np.random.default_rng(i))
            c.ops.insert(0, (W.UNITARY, (0, 4, 2), tuple(np.stack([u.real, u.imag], -1).reshape(-1).tolist())))
        elif i % 3 == 1:
            c = W.Circuit(n, 0, 0.37 * i)
        else:
            c = W.random_circuit(n, 3, 900 + i)
        c = _unitary_only(c)
        c.global_phase = 0.21 * i
        circuits.append(c)
        terms.append(W.random_paulis(n, 6, 900 + i, max_weight=3))
    for fusion in (0, 1):
        r = hip.batch_execute(circuits, terms, fusion=fusion, statevectors=True)
        for i, c in enumerate(circuits):
            want, _, e = orc.execute(c, fuse_gates=bool(fusion), terms=terms[i])
            assert r.statevectors[i].tobytes() == want.tobytes(), (fusion, i)
            assert np.max(np.abs(r.expectation[i] - e)) <= 1e-12, (fusion, i)
