"""Host-side pieces of the serving path, on the CPU: the SplitMix64 seed
stream against the published vector, the library's mh_derive_seeds against
it, the observable combination rule, and the served workload's shape."""
import math

import numpy as np

import oracle
from menos_amd import workloads as W

# Vigna's splitmix64.c, seed 1234567: its first five outputs.
PUBLISHED_1234567 = [6457827717110365317, 3203168211198807973, 9817491932198370423, 4593380528125082431,
                     16408922859458223821]


def test_splitmix64_published_vector():
    assert oracle.splitmix64(1234567, 5) == PUBLISHED_1234567
    assert W.splitmix64_stream(1234567, 5) == PUBLISHED_1234567
    assert oracle.splitmix64(2026, 100) == W.splitmix64_stream(2026, 100)


def test_library_derive_seeds_matches_the_stream():
    from menos_amd import hip
    assert hip.derive_seeds(1234567, 5) == PUBLISHED_1234567
    for seed in (0, 1, 2**63 + 5, 2**64 - 1):
        assert hip.derive_seeds(seed, 17) == oracle.splitmix64(seed, 17)
    assert hip.derive_seeds(9, 0) == []


def test_combine_expectation_exact_and_ordered():
    # Dyadic values: every product and sum is exact, so a dropped or
    # mis-signed cross term changes the result.
    terms = [complex(0.5, -0.25), complex(-0.75, 0.125)]
    coefs = [complex(2.0, 1.0), complex(-1.0, 0.5)]
    want = (2.0 * 0.5 - 1.0 * -0.25) + (-1.0 * -0.75 - 0.5 * 0.125), \
        (2.0 * -0.25 + 1.0 * 0.5) + (-1.0 * 0.125 + 0.5 * -0.75)
    assert oracle.combine_expectation(terms, coefs) == want
    # Summation in term order from (0, 0): (1e16 + 1) rounds back to 1e16.
    assert oracle.combine_expectation([1e16, 1.0, -1e16], [1.0, 1.0, 1.0])[0] == 0.0
    assert oracle.combine_expectation([1e16, -1e16, 1.0], [1.0, 1.0, 1.0])[0] == 1.0
    assert oracle.combine_expectation([], []) == (0.0, 0.0)


def test_flat_brick_is_the_random_circuit():
    n, layers, seed = 7, 4, 99
    ref = W.random_circuit(n, layers, seed).measure_all()
    angles = np.array([p for op, _, ps in ref.ops if op == W.U for p in ps])
    fb = W.FlatBrick(n, layers, angles)
    for a, b in zip(fb.flat(), ref.flat()):
        assert np.array_equal(a, b)
    assert fb.measure_targets() == ref.measure_targets()


def test_serve_workload_shape():
    jobs = W.serve_workload(400, 50.0, 7)
    arr = np.array([j.arrival for j in jobs])
    assert arr[0] == 0.0 and np.all(np.diff(arr) >= 0)
    assert 400 / 50.0 * 0.7 < arr[-1] < 400 / 50.0 * 1.3
    ns = np.array([j.circuit.num_qubits for j in jobs])
    assert ns.min() >= 12 and ns.max() <= 30
    # log-uniform: about ln(15/12)/ln(31/12) of the structures below 15 qubits
    assert 0.12 < np.mean(ns < 15) < 0.35
    upd = np.mean([j.update for j in jobs])
    assert 0.3 < upd < 0.5
    est = np.mean([j.kind == "estimator" for j in jobs])
    assert 0.6 < est < 0.8
    for j in jobs:
        if j.kind == "estimator":
            assert len(j.observables) == 16 and j.shots == 0
        else:
            assert j.shots == 1000 and j.circuit.measure_targets() == [(q, q) for q in range(j.circuit.num_qubits)]
    first = {}
    for j in jobs:
        if j.structure in first:
            f = first[j.structure]
            assert j.update and j.kind == f.kind and j.circuit.num_qubits == f.circuit.num_qubits
            assert np.array_equal(j.circuit.flat()[0], f.circuit.flat()[0])
            assert np.array_equal(j.circuit.flat()[2], f.circuit.flat()[2])
            assert j.observables == f.observables
        else:
            first[j.structure] = j
            assert not j.update
    assert math.isclose(sum(1 for j in jobs if j.update) / len(jobs), upd)
