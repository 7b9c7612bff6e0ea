"""The whole random circuit on the GPU and on the reference CPU engine, state
against state, at sizes no pytest run should take (33 qubits: a 128 GiB state on
each side, about ten minutes of CPU time on 16 cores).

The GPU applies the reference's containment plan with mh_apply_blocks (the
engine's pass schedule); the CPU applies the same blocks with the reference's
own kernels (oracle/_ref: kernels::cpu::apply_matrix with its AVX2 table and
thread pool, fused-block matrices by the reference's fused_block_matrix). The
states are then compared in chunks, byte for byte. Prints one progress line per
layer and a JSON summary line.

    python tests/full_state_check.py --qubits 33 --layers 16
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from menos_amd import hip as H  # noqa: E402
from menos_amd import workloads as W  # noqa: E402


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--qubits", type=int, default=33)
    ap.add_argument("--layers", type=int, default=16)
    ap.add_argument("--seed", type=int, default=2025)
    ap.add_argument("--chunk-log", type=int, default=24)
    args = ap.parse_args()
    n = args.qubits
    ref = oracle.Reference()
    threads = len(os.sched_getaffinity(0))
    c = W.random_circuit(n, args.layers, args.seed)
    plan = ref.plan_fusion(c)
    blocks = []
    for b, (members, qubits) in enumerate(plan.blocks()):
        if len(members) == 1:
            op, qs, params = c.ops[members[0]]
            blocks.append((ref.gate_matrix(op, params), list(qs), members[0]))
        else:
            blocks.append((ref.fused_block_matrix(c, plan, b), list(qubits), members[-1]))
    print(f"{n} qubits, {args.layers} layers: {len(blocks)} blocks; {threads} host threads", flush=True)

    t0 = time.perf_counter()
    d = H.DeviceState(n)
    d.init()
    passes = d.apply_blocks([(m, q) for m, q, _ in blocks])
    H.lib().mh_pool_trim()
    gpu_s = time.perf_counter() - t0
    print(f"GPU: {passes} passes, {gpu_s:.2f} s including initialisation", flush=True)

    state = np.zeros(1 << n, np.complex128)
    state[0] = 1.0
    gates_per_layer = len(c.ops) / args.layers
    t0 = time.perf_counter()
    layer_done = 0
    for m, q, last_op in blocks:
        ref.apply_matrix(state, m, q, threads=threads)
        layer = int((last_op + 1) // gates_per_layer)
        if layer > layer_done:
            layer_done = layer
            print(f"CPU: layer {layer} done at {time.perf_counter() - t0:.1f} s", flush=True)
    cpu_s = time.perf_counter() - t0
    print(f"CPU: {len(blocks)} blocks in {cpu_s:.1f} s", flush=True)

    chunk = 1 << args.chunk_log
    differ = 0
    first_diff = None
    t0 = time.perf_counter()
    for at in range(0, 1 << n, chunk):
        idx = np.arange(at, at + chunk, dtype=np.uint64)
        got = d.gather(idx)
        want = state[at:at + chunk]
        if got.tobytes() != want.tobytes():
            bad = np.flatnonzero(got.view(np.uint64).reshape(-1, 2) != want.view(np.uint64).reshape(-1, 2))
            differ += int(np.unique(bad // 2).size)
            if first_diff is None:
                first_diff = at + int(bad[0] // 2)
        if (at // chunk) % 64 == 63:
            print(f"compare: {(at + chunk) / (1 << n):.0%}", flush=True)
    d.free()
    out = {"qubits": n, "layers": args.layers, "blocks": len(blocks), "gpu_passes": passes,
           "gpu_seconds_with_init": round(gpu_s, 3), "cpu_seconds": round(cpu_s, 1), "cpu_threads": threads,
           "amplitudes_compared": 1 << n, "amplitudes_differing": differ, "first_difference": first_diff,
           "compare_seconds": round(time.perf_counter() - t0, 1)}
    print(json.dumps(out), flush=True)
    return 0 if differ == 0 else 1


if __name__ == "__main__":
    sys.exit(main())
