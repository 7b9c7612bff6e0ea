"""The estimator and sampler primitives on the job server
(menos_amd/primitives.py), against the oracle: each observable's value is the
reference's combine_expectation of the oracle's raw terms (within 1e-12 at
complex128, 1e-5 at complex64), for every parameter binding; counts equal the
oracle's for the SplitMix64 seeds the reference's sampler derives (pub j: the
j-th output of the run seed; binding b: the b-th output of the pub's seed)."""
import math

import numpy as np
import pytest

import oracle
from menos_amd import primitives as P
from menos_amd import workloads as W

pytestmark = pytest.mark.gpu


def _observables(n, seed, count=3):
    rng = np.random.default_rng(seed)
    out = []
    for o in range(count):
        obs = {}
        for _ in range(1 + o):
            letters = rng.choice(list("IXYZ"), size=n)
            obs["".join(letters)] = complex(rng.uniform(-1, 1), 0.0 if o % 2 == 0 else rng.uniform(-0.5, 0.5))
        out.append(obs)
    return out


def _want(orc, circuit, observable):
    terms = P.observable_terms(observable, circuit.num_qubits)
    _, _, raw = orc.execute(circuit, terms=[(x, z) for x, z, _ in terms], statevector=False)
    return oracle.combine_expectation(raw, [c for _, _, c in terms])[0]


@pytest.mark.parametrize("precision,tol", [(0, 1e-12), (1, 1e-5)])
def test_estimator_against_the_oracle(orc, hip, precision, tol):
    rng = np.random.default_rng(4)
    ansatz = W.FlatAnsatz(12, 3, rng.uniform(-math.pi, math.pi, 72))
    values = rng.uniform(-math.pi, math.pi, (3, 72))
    plain = W.random_circuit(11, 4, 9)
    obs_a, obs_p = _observables(12, 1), _observables(11, 2)
    est = P.HipEstimator(precision=precision, seed=5)
    res = est.run([(ansatz, obs_a, values), (plain, obs_p)])
    assert res[0].evs.shape == (3, 3) and res[1].evs.shape == (3,)
    assert not res[0].stds.any()
    for b in range(3):
        bound = ansatz.with_angles(values[b])
        for o, observable in enumerate(obs_a):
            assert abs(res[0].evs[b, o] - _want(orc, bound, observable)) <= tol, (b, o)
    for o, observable in enumerate(obs_p):
        assert abs(res[1].evs[o] - _want(orc, plain, observable)) <= tol, o


def test_sampler_seeds_and_counts(orc, hip):
    rng = np.random.default_rng(8)
    c1 = W.random_circuit(10, 3, 21).measure_all()
    c2 = W.random_circuit(12, 2, 22).measure_all()
    values = rng.uniform(0, 2 * math.pi, (2, 3 * 12 * 2))
    sam = P.HipSampler(seed=123, default_shots=700)
    res = sam.run([(c1,), (c2, values, 900)])
    pub_seeds = hip.derive_seeds(123, 2)
    assert res[0].metadata["binding_seeds"] == [pub_seeds[0]]
    _, idx, _ = orc.execute(c1, shots=700, seed=pub_seeds[0], statevector=False)
    assert res[0].counts[0] == oracle.counts_from_indices(idx, c1.measure_targets())
    seeds = hip.derive_seeds(pub_seeds[1], 2)
    for b in range(2):
        bound = P.bind(c2, values[b])
        _, idx, _ = orc.execute(bound, shots=900, seed=seeds[b], statevector=False)
        assert res[1].counts[b] == oracle.counts_from_indices(idx, c2.measure_targets())
        assert sum(res[1].counts[b].values()) == 900
