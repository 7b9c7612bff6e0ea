"""Sampled per-pass verification of states too large for the oracle.

A pass applies a run of consecutive blocks to the state. Its dense blocks act on
a set D of qubits; its diagonal blocks act amplitude by amplitude. So a *group*
(the 2^|D| amplitudes that share every bit outside D) is closed under the pass:
the pass's result on a group depends on that group alone. For a sample of
groups, the check gathers their amplitudes from the device before and after the
pass and applies the same blocks to the gathered values on the host with the
oracle's generic kernel (reference src/kernels/cpu/dense_kernels.cpp:465-524)
and the reference's streaming diagonal form ((+0) + d[local] * a,
dense_kernels.cpp:299-313). The groups sit side by side in one compact array of
2^(|D| + g) amplitudes: bit j < |D| is D's j-th qubit, the high g bits number the
group, so one oracle call applies a dense block to every sampled group.
"""
from __future__ import annotations

import numpy as np


def is_diagonal(m: np.ndarray) -> bool:
    return not np.any(m[~np.eye(m.shape[0], dtype=bool)])


def deposit(values: np.ndarray, positions) -> np.ndarray:
    """Bit j of each value moved to bit positions[j] (uint64)."""
    out = np.zeros(values.shape, np.uint64)
    for j, q in enumerate(positions):
        out |= ((values >> np.uint64(j)) & np.uint64(1)) << np.uint64(q)
    return out


def group_indices(n: int, dense_qubits, group_bits: int, rng) -> np.ndarray:
    """Global indices of 2^group_bits random groups over D = dense_qubits:
    shape (2^group_bits, 2^|D|), row r the group's members in D-local order."""
    others = [q for q in range(n) if q not in dense_qubits]
    outer = rng.integers(0, 1 << len(others), 1 << group_bits, dtype=np.uint64) if others else \
        np.zeros(1 << group_bits, np.uint64)
    base = deposit(outer, others)
    local = deposit(np.arange(1 << len(dense_qubits), dtype=np.uint64), dense_qubits)
    return base[:, None] | local[None, :]


def apply_diagonal(compact: np.ndarray, diag: np.ndarray, operands, idx: np.ndarray) -> np.ndarray:
    """(+0) + d[local] * a per amplitude, each product and sum rounded on its own;
    local takes bit j from the amplitude's global index at operands[j]."""
    loc = deposit_from(idx, operands)
    d = diag[loc.astype(np.int64)]
    ar, ai = compact.real, compact.imag
    pr = d.real * ar - d.imag * ai
    pi = d.real * ai + d.imag * ar
    out = np.empty_like(compact)
    out.real = 0.0 + pr
    out.imag = 0.0 + pi
    return out


def deposit_from(idx: np.ndarray, operands) -> np.ndarray:
    """Bit operands[j] of each global index gathered into bit j."""
    loc = np.zeros(idx.shape, np.uint64)
    for j, q in enumerate(operands):
        loc |= ((idx >> np.uint64(q)) & np.uint64(1)) << np.uint64(j)
    return loc


def expected_after(orc, before: np.ndarray, idx: np.ndarray, blocks, dense_qubits) -> np.ndarray:
    """The oracle's result of `blocks` on the sampled groups (flattened, in idx order)."""
    compact = np.ascontiguousarray(before, np.complex128).copy()
    flat = idx.reshape(-1)
    pos = {q: j for j, q in enumerate(dense_qubits)}
    for m, qs in blocks:
        if is_diagonal(m):
            compact = apply_diagonal(compact, np.diag(m).copy(), qs, flat)
        else:
            orc.apply_generic(compact, m, [pos[q] for q in qs])
    return compact


def dense_qubits_of(blocks) -> list[int]:
    return sorted({q for m, qs in blocks if not is_diagonal(m) for q in qs})


def check_passes(orc, d, blocks, passes, rng, max_amplitudes=1 << 21, on_pass=None):
    """Applies `blocks` to the device state pass by pass (passes: [(begin, end)]),
    checking each pass on sampled groups; returns the number of amplitudes
    checked. Raises AssertionError naming the first pass that differs."""
    checked = 0
    for p, (b0, b1) in enumerate(passes):
        part = blocks[b0:b1]
        dq = dense_qubits_of(part)
        gbits = max(0, min(12, int(np.log2(max_amplitudes)) - len(dq)))
        idx = group_indices(d.n, dq, gbits, rng)
        before = d.gather(idx.reshape(-1)).astype(np.complex128)
        d.apply_blocks(part)
        after = d.gather(idx.reshape(-1)).astype(np.complex128)
        want = expected_after(orc, before, idx, part, dq)
        if want.tobytes() != after.tobytes():
            bad = int(np.flatnonzero(want.view(np.uint64).reshape(-1, 2) != after.view(np.uint64).reshape(-1, 2))[0])
            raise AssertionError(f"pass {p} (blocks {b0}..{b1 - 1}): amplitude {int(idx.reshape(-1)[bad])} "
                                 f"device {after[bad]!r} oracle {want[bad]!r}")
        checked += idx.size
        if on_pass:
            on_pass(p, idx.size)
    return checked


def circuit_blocks(orc, circuit, plan):
    """The block list the engine runs for a plan: a single-gate block keeps its
    gate's matrix and operand order, a fused block its matrix over its ascending
    qubits (reference src/engines/dense/fused_block.h:59-101)."""
    blocks = []
    for b, (members, qubits) in enumerate(plan.blocks()):
        if len(members) == 1:
            op, qs, params = circuit.ops[members[0]]
            blocks.append((orc.gate_matrix(op, params), list(qs)))
        else:
            blocks.append((orc.fused_block_matrix(circuit, plan, b), list(qubits)))
    return blocks
