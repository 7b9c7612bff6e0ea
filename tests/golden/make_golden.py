"""Writes the golden vectors under tests/golden/ from the reference's own code.

Run in the container that holds the reference sources, after `make -C oracle`
has compiled them into oracle/_ref/libmenos_ref.so:

    python tests/golden/make_golden.py

Outputs (data only; no reference source is copied):
  ref_vectors.json  gate matrices, containment fusion plans, counts, expectation
                    terms and statevector digests produced by the compiled
                    reference (DenseEngine::execute, ir::plan_fusion,
                    ir::generate_matrix, engines::dense::sample, ...);
  ref_states.npz    a few small raw statevectors from DenseEngine::execute.
The reference's committed digests for its dense memory fixtures
(tests/memory/data/baseline_linux.json) are copied into ref_vectors.json
under "reference_baseline" as they stand.
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import oracle  # noqa: E402
from circuits import diagonal_heavy_circuit, random_mixed_circuit  # noqa: E402
from menos_amd import workloads as W  # noqa: E402

REFERENCE = "/root/reference"


def cpx(a) -> list:
    a = np.asarray(a, np.complex128).reshape(-1)
    return np.stack([a.real, a.imag], axis=1).tolist()


def circuit_spec(c: W.Circuit) -> dict:
    return {"num_qubits": c.num_qubits, "num_clbits": c.num_clbits, "global_phase": c.global_phase,
            "ops": [[op, list(q), list(p)] for op, q, p in c.ops]}


def main() -> None:
    ref = oracle.Reference()
    out: dict = {"generator": "tests/golden/make_golden.py", "source": "compiled reference (oracle/_ref)"}

    # Named-gate matrices.
    rng = np.random.default_rng(20260929)
    gates = []
    for op in range(21):
        for _ in range(3):
            params = rng.uniform(-4 * np.pi, 4 * np.pi, size=W.GATE_PARAMS.get(op, 0)).tolist()
            gates.append({"opcode": op, "params": params, "matrix": cpx(ref.gate_matrix(op, params))})
            if op not in W.GATE_PARAMS:
                break
    out["gate_matrices"] = gates

    # Containment fusion plans.
    plans = []
    corpus = [random_mixed_circuit(n, 40, 100 + n, measures=True) for n in (3, 5, 7, 9)]
    corpus += [diagonal_heavy_circuit(n, 60, 200 + n) for n in (4, 6, 8)]
    corpus += [W.layered(12, 3), W.random_circuit(10, 4, 7)]
    for c in corpus:
        p = ref.plan_fusion(c)
        plans.append({"circuit": circuit_spec(c), "order": p.order.tolist(), "block_begin": p.block_begin.tolist(),
                      "block_qubits": p.block_qubits.tolist(), "qubits_begin": p.qubits_begin.tolist(),
                      "fused_blocks": p.fused_blocks, "gates_absorbed": p.gates_absorbed})
    out["plans"] = plans

    # Whole-engine runs on small mixed circuits: raw states, counts, terms.
    runs, states = [], {}
    for i, n in enumerate((3, 4, 5, 6, 7, 8)):
        c = random_mixed_circuit(n, 12 * n, 300 + i)
        terms = W.random_paulis(n, 6, 400 + i, max_weight=min(4, n))
        for fuse in (False, True):
            sv, counts, e, _ = ref.execute(c, fuse_gates=fuse, shots=2000, seed=9000 + i, terms=terms,
                                           targets=[(q, q) for q in range(n)], width=n)
            key = f"mixed_{n}_{'fused' if fuse else 'unfused'}"
            states[key] = sv
            runs.append({"key": key, "circuit": circuit_spec(c), "fuse_gates": fuse, "shots": 2000,
                         "seed": 9000 + i, "terms": [list(t) for t in terms],
                         "counts": {str(k): v for k, v in counts.items()}, "expectation": cpx(e),
                         "statevector_digest": W.digest_complex(sv)})
    out["engine_runs"] = runs

    # The 20-qubit random circuit, 20 layers: digests, 1024-shot counts, <Z_i>.
    c = W.random_circuit(20, 20, 2025)
    terms = W.z_terms(20)
    sv, counts, e, secs = ref.execute(c, fuse_gates=True, shots=1024, seed=2025, terms=terms,
                                      targets=[(q, q) for q in range(20)], width=20)
    o = oracle.Oracle()
    out["random_circuit_20q"] = {"num_qubits": 20, "layers": 20, "circuit_seed": 2025, "shots": 1024,
                                 "sample_seed": 2025, "fuse_gates": True,
                                 "statevector_digest": "%016x" % o.fnv1a(sv),
                                 "counts_digest": W.digest_counts(counts),
                                 "counts": {str(k): v for k, v in counts.items()},
                                 "z_expectation": cpx(e), "reference_seconds_8_threads": secs}

    # The reference's known-answer tests.
    out["kat"] = {
        "mt19937_64_default_seed_10000th": 9981545732273789042,  # [rand.predef]
        "sample_plus_20_shots_seed_42": {"0": 9, "1": 11},     # tests/engines/dense/sample_test.cpp:85-92
    }
    plus = np.array([1, 1], np.complex128) / np.sqrt(2.0)
    assert ref.sample_counts(plus, 20, 42, [(0, 0)], 1) == {0: 9, 1: 11}

    # The reference's committed digests for its dense fixtures.
    base = json.load(open(os.path.join(REFERENCE, "tests/memory/data/baseline_linux.json")))
    out["reference_baseline"] = {k: v for k, v in base.items()
                                 if k.startswith("dense") and not k.startswith("dense_disk")}

    with open(os.path.join(HERE, "ref_vectors.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    np.savez_compressed(os.path.join(HERE, "ref_states.npz"), **states)
    print("wrote", os.path.join(HERE, "ref_vectors.json"))


if __name__ == "__main__":
    main()
