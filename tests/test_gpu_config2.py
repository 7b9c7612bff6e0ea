"""The 30-qubit complex128 random circuit (16 GiB; 24 layers of U3 on every qubit
then CZ on the brick pattern) through the engine, against golden values the
reference's own code produced in this build's container
(tests/golden/make_config2.py: oracle/_ref applying the same block list) for
two plans: the reference's containment fusion and width-capped fusion at width
5 (dense blocks of up to 5 qubits, on the per-group kernel). The statevector's
FNV-1a digest and the 4096 seeded counts must match bit for bit, the 64 Pauli
terms (weight 1..6) within 1e-12."""
import json
import os

import numpy as np
import pytest

from menos_amd import workloads as W

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "config2_30q.json")


@pytest.fixture(scope="module")
def config2():
    with open(GOLDEN) as f:
        return json.load(f)


@pytest.mark.parametrize("plan", ["containment", "width5"])
def test_30_qubits_against_the_reference(orc, hip, config2, plan):
    g = config2
    n = g["num_qubits"]
    c = W.random_circuit(n, g["layers"], g["seed"])
    terms = [tuple(t) for t in g["terms"]]
    want = g["plans"][plan]
    fusion = dict(fusion=hip.FUSE_CONTAINMENT) if plan == "containment" else dict(fusion=hip.FUSE_WIDTH,
                                                                                  fusion_width=5)
    r = hip.HipDenseEngine().execute(c, shots=g["shots"], seed=g["shot_seed"], targets=[(q, q) for q in range(n)],
                                     width=n, terms=terms, statevector=True, **fusion)
    assert r.num_blocks == want["blocks"]
    digest = "%016x" % orc.fnv1a(r.statevector.view(np.float64))
    del r.statevector
    print(f"{plan}: {r.num_blocks} blocks in {r.num_passes} passes, gate passes {r.gate_seconds:.3f} s, "
          f"digest {digest}", flush=True)
    assert digest == want["statevector_fnv1a"]
    assert sorted(r.counts.items()) == [tuple(kv) for kv in want["counts"]]
    ref_terms = np.array([complex(re, im) for re, im in want["expectation"]])
    assert np.max(np.abs(r.expectation - ref_terms)) <= 1e-12
