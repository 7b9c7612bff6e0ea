"""Writes tests/golden/config2_30q.json: the 30-qubit complex128 random circuit
(24 layers of U3 on every qubit + CZ brick) run by the reference's own code
(oracle/_ref: the reference's fused-block sweep, ref_run_plan, over the plan
given), for two plans: the reference's containment fusion and this engine's
width-capped fusion at width 5. For each: the statevector's FNV-1a digest
(the 16 GiB of (re, im) doubles in index order), 4096 seeded shots through the
reference sampler, and 64 random Pauli terms (weight 1..6) by the reference's
expectation_of_term.

Run in the container that holds the reference sources, after `make -C oracle`
(about 62 GB of RAM; roughly an hour on 8 cores):

    python tests/golden/make_config2.py
"""
from __future__ import annotations

import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from menos_amd import hip as H  # noqa: E402
from menos_amd import workloads as W  # noqa: E402

N, LAYERS, SEED = 30, 24, 2030
SHOTS, SHOT_SEED = 4096, 4242
TERMS, TERM_SEED = 64, 31


def main() -> None:
    ref = oracle.Reference()
    orc = oracle.Oracle()
    c = W.random_circuit(N, LAYERS, SEED)
    terms = W.random_paulis(N, TERMS, TERM_SEED, max_weight=6)
    out = {"num_qubits": N, "layers": LAYERS, "seed": SEED, "shots": SHOTS, "shot_seed": SHOT_SEED,
           "terms": [[int(x), int(z)] for x, z in terms], "plans": {}}
    threads = os.cpu_count() or 8
    plans = {"containment": ref.plan_fusion(c)}
    order, bb, bq, qb, fused, absorbed = H.plan_fusion(c, H.FUSE_WIDTH, 5)
    plans["width5"] = oracle.Plan(order, bb, bq, qb, fused, absorbed)
    for name, plan in plans.items():
        t0 = time.perf_counter()
        sv, secs = ref.run_plan(c, plan, threads=threads)
        digest = orc.fnv1a(sv.view(np.float64))
        counts = ref.sample_counts(sv, SHOTS, SHOT_SEED, [(q, q) for q in range(N)], N)
        values = [ref.expectation_of_term(sv, x, z) for x, z in terms]
        del sv
        out["plans"][name] = {
            "blocks": plan.num_blocks,
            "statevector_fnv1a": "%016x" % digest,
            "counts": [[int(k), int(v)] for k, v in sorted(counts.items())],
            "counts_digest": W.digest_counts(counts),
            "expectation": [[float(v.real), float(v.imag)] for v in values],
            "reference_seconds": round(secs, 2),
            "threads": threads,
        }
        print(f"{name}: {plan.num_blocks} blocks, run {secs:.1f} s on {threads} threads, "
              f"total {time.perf_counter() - t0:.1f} s", flush=True)
        with open(os.path.join(HERE, "config2_30q.json"), "w") as f:
            json.dump(out, f)


if __name__ == "__main__":
    main()
