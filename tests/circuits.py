"""Seeded test circuits that exercise every gate kind, unitaries and barriers."""
from __future__ import annotations

import numpy as np

from menos_amd import workloads as W


def random_unitary(k: int, rng: np.random.Generator) -> np.ndarray:
    d = 1 << k
    z = rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d))
    q, r = np.linalg.qr(z)
    return q * (np.diag(r) / np.abs(np.diag(r)))


def random_mixed_circuit(n: int, num_ops: int, seed: int, *, unitaries: bool = True, measures: bool = False,
                         max_unitary: int = 5, max_arity: int = 3) -> W.Circuit:
    """Every named gate kind (arity permitting, at most max_arity operands),
    random angles and operand orders, k-qubit unitaries with shuffled operands,
    optional mid-circuit measurement barriers."""
    rng = np.random.default_rng(seed)
    c = W.Circuit(n, num_clbits=n)
    c.global_phase = float(rng.uniform(-np.pi, np.pi))
    kinds = [op for op, a in W.GATE_ARITY.items() if a <= min(n, max_arity)]
    for _ in range(num_ops):
        roll = rng.random()
        if unitaries and roll < 0.08:
            k = int(rng.integers(1, min(max_unitary, n) + 1))
            qs = rng.permutation(n)[:k].tolist()
            c.append_unitary(qs, random_unitary(k, rng))
        elif measures and roll < 0.12:
            q = int(rng.integers(0, n))
            c.append_measure(q, q)
        else:
            op = int(rng.choice(kinds))
            qs = rng.permutation(n)[:W.GATE_ARITY[op]].tolist()
            params = rng.uniform(-2 * np.pi, 2 * np.pi, size=W.GATE_PARAMS.get(op, 0)).tolist()
            c.append(op, qs, params)
    return c


def diagonal_heavy_circuit(n: int, num_ops: int, seed: int) -> W.Circuit:
    """Runs of diagonal gates between dense ones: the containment rule's
    diagonal cases and the diagonal kernel."""
    rng = np.random.default_rng(seed)
    c = W.Circuit(n)
    diag = [W.Z, W.S, W.SDG, W.T, W.TDG, W.RZ, W.P, W.CZ, W.CP, W.RZZ]
    dense = [W.H, W.RX, W.RY, W.U, W.CX, W.SWAP, W.SX]
    for _ in range(num_ops):
        op = int(rng.choice(diag if rng.random() < 0.6 else dense))
        if W.GATE_ARITY[op] > n:
            continue
        qs = rng.permutation(n)[:W.GATE_ARITY[op]].tolist()
        params = rng.uniform(-2 * np.pi, 2 * np.pi, size=W.GATE_PARAMS.get(op, 0)).tolist()
        c.append(op, qs, params)
    return c


def random_state(n: int, seed: int, dtype=np.complex128) -> np.ndarray:
    rng = np.random.default_rng(seed)
    s = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    s /= np.linalg.norm(s)
    return s.astype(dtype)
