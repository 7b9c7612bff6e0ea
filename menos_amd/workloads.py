"""Circuits and observables in the engine's flat layout, plus the workload generators.

The flat layout is the reference's batch form (``Circuit::append_batch``,
reference src/ir/circuit.cpp:34-67): ``opcodes``, ``qubits`` with
``qubit_offsets``, ``params`` with ``param_offsets``. Two operations beyond
named gates ride in it: a ``UNITARY`` op carries its row-major 2^k x 2^k complex
matrix as interleaved doubles in ``params``; a ``MEASURE`` op carries
``(qubit, clbit)`` as its two ``qubits`` entries.

Opcode numbers are the reference's ``enum class Opcode``
(reference include/menos/circuit.h).

The random-circuit and Pauli generators draw from ``std::mt19937_64``, which the
C++ standard specifies completely ([rand.predef]); ``Mt19937_64`` restates it so
a workload is the same bytes in Python, in the C++ oracle and in the engine.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

# reference include/menos/circuit.h: enum class Opcode
X, Y, Z, H, S, SDG, T, TDG, SX, SXDG, RX, RY, RZ, P, U, CX, CZ, SWAP, CP, RZZ, CCX, UNITARY, MEASURE = range(23)

GATE_ARITY = {X: 1, Y: 1, Z: 1, H: 1, S: 1, SDG: 1, T: 1, TDG: 1, SX: 1, SXDG: 1, RX: 1, RY: 1, RZ: 1,
              P: 1, U: 1, CX: 2, CZ: 2, SWAP: 2, CP: 2, RZZ: 2, CCX: 3}
GATE_PARAMS = {RX: 1, RY: 1, RZ: 1, P: 1, U: 3, CP: 1, RZZ: 1}
# The fusion rule's diagonal opcodes (reference src/ir/fuse_diagonal.cpp:16-32).
DIAGONAL_OPCODES = {Z, S, SDG, T, TDG, RZ, P, CZ, CP, RZZ}


class Mt19937_64:
    """``std::mt19937_64``: w=64, n=312, m=156, r=31 ([rand.predef])."""

    _N, _M = 312, 156
    _MATRIX_A = 0xB5026F5AA96619E9
    _UPPER, _LOWER = 0xFFFFFFFF80000000, 0x7FFFFFFF
    _MASK = (1 << 64) - 1

    def __init__(self, seed: int = 5489):
        mt = [0] * self._N
        mt[0] = seed & self._MASK
        for i in range(1, self._N):
            mt[i] = (6364136223846793005 * (mt[i - 1] ^ (mt[i - 1] >> 62)) + i) & self._MASK
        self._mt = mt
        self._i = self._N

    def _twist(self):
        mt, n, m = self._mt, self._N, self._M
        for i in range(n):
            x = (mt[i] & self._UPPER) | (mt[(i + 1) % n] & self._LOWER)
            xa = x >> 1
            if x & 1:
                xa ^= self._MATRIX_A
            mt[i] = mt[(i + m) % n] ^ xa
        self._i = 0

    def __call__(self) -> int:
        if self._i >= self._N:
            self._twist()
        x = self._mt[self._i]
        self._i += 1
        x ^= (x >> 29) & 0x5555555555555555
        x ^= (x << 17) & 0x71D67FFFEDA60000
        x ^= (x << 37) & 0xFFF7EEE000000000
        x ^= x >> 43
        return x & self._MASK


def word_to_uniform(word: int) -> float:
    """``(word >> 11) * 2^-53`` (reference src/common/rng.cpp:8-24)."""
    return float(word >> 11) * (1.0 / 9007199254740992.0)


def splitmix64_stream(seed: int, count: int) -> list[int]:
    """The reference's SplitMix64 sub-seeds (src/common/splitmix64.h:15-28)."""
    mask = (1 << 64) - 1
    state, out = seed & mask, []
    for _ in range(count):
        state = (state + 0x9E3779B97F4A7C15) & mask
        z = state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & mask
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & mask
        out.append(z ^ (z >> 31))
    return out


@dataclass
class Circuit:
    num_qubits: int
    num_clbits: int = 0
    global_phase: float = 0.0
    ops: list = field(default_factory=list)  # (opcode, qubits tuple, params tuple)

    def append(self, opcode: int, qubits, params=()):
        self.ops.append((int(opcode), tuple(int(q) for q in qubits), tuple(float(p) for p in params)))
        return self

    def append_unitary(self, qubits, matrix):
        m = np.asarray(matrix, dtype=np.complex128).reshape(-1)
        flat = np.empty(2 * m.size, dtype=np.float64)
        flat[0::2], flat[1::2] = m.real, m.imag
        self.ops.append((UNITARY, tuple(int(q) for q in qubits), tuple(flat.tolist())))
        return self

    def append_measure(self, qubit: int, clbit: int):
        self.ops.append((MEASURE, (int(qubit), int(clbit)), ()))
        return self

    def measure_all(self):
        self.num_clbits = max(self.num_clbits, self.num_qubits)
        for q in range(self.num_qubits):
            self.append_measure(q, q)
        return self

    def measure_targets(self):
        """(qubit, clbit) of the last measurement into each clbit, by clbit."""
        last = {}
        for opcode, qubits, _ in self.ops:
            if opcode == MEASURE:
                last[qubits[1]] = qubits[0]
        return [(last[c], c) for c in sorted(last)]

    def flat(self):
        """The flat arrays: opcodes, qubit_offsets, qubits, param_offsets, params."""
        opcodes = np.array([o[0] for o in self.ops], dtype=np.int32)
        qo, po, qs, ps = [0], [0], [], []
        for _, q, p in self.ops:
            qs.extend(q)
            ps.extend(p)
            qo.append(len(qs))
            po.append(len(ps))
        return (opcodes, np.array(qo, dtype=np.uint32), np.array(qs, dtype=np.uint32),
                np.array(po, dtype=np.uint32), np.array(ps, dtype=np.float64))

    def with_params(self, params_by_op):
        """A copy with new angles for the given op indices (shape unchanged)."""
        c = Circuit(self.num_qubits, self.num_clbits, self.global_phase, list(self.ops))
        for i, p in params_by_op.items():
            op = c.ops[i]
            c.ops[i] = (op[0], op[1], tuple(float(v) for v in p))
        return c


def random_circuit(num_qubits: int, layers: int, seed: int) -> Circuit:
    """The dense random circuit of the workloads.

    Per layer: ``U(theta, phi, lambda)`` on qubits 0..n-1, each angle
    ``word_to_uniform(g()) * 2*pi`` from ``g = mt19937_64(seed)`` in
    (layer, qubit, theta -> phi -> lambda) order; then ``CZ(i, i+1)`` for
    ``i = layer % 2, layer % 2 + 2, ... < n - 1``.
    """
    g = Mt19937_64(seed)
    c = Circuit(num_qubits)
    two_pi = 2.0 * math.pi
    for layer in range(layers):
        for q in range(num_qubits):
            theta = word_to_uniform(g()) * two_pi
            phi = word_to_uniform(g()) * two_pi
            lam = word_to_uniform(g()) * two_pi
            c.append(U, (q,), (theta, phi, lam))
        for q in range(layer % 2, num_qubits - 1, 2):
            c.append(CZ, (q, q + 1))
    return c


def layered(num_qubits: int, depth: int) -> Circuit:
    """The reference's memory fixture circuit (tests/memory/memory_fixtures.cpp:53-67)."""
    c = Circuit(num_qubits)
    for layer in range(depth):
        for q in range(num_qubits):
            c.append(RY, (q,), (0.13 + (0.071 * float(q)) + (0.37 * float(layer)),))
        c.append(T, (0,))
        for q in range(layer % 2, num_qubits - 1, 2):
            c.append(CX, (q, q + 1))
    return c


def ansatz_circuit(num_qubits: int, depth: int, angles) -> Circuit:
    """Hardware-efficient ansatz: per layer RY then RZ on every qubit, then a
    CNOT ladder 0->1->...->n-1. ``angles`` holds 2*n*depth values in
    (layer, RY of q0..q{n-1}, RZ of q0..q{n-1}) order."""
    c = Circuit(num_qubits)
    a = iter(angles)
    for _ in range(depth):
        for q in range(num_qubits):
            c.append(RY, (q,), (next(a),))
        for q in range(num_qubits):
            c.append(RZ, (q,), (next(a),))
        for q in range(num_qubits - 1):
            c.append(CX, (q, q + 1))
    return c


def pauli_masks(spec: dict[int, str]) -> tuple[int, int]:
    """(x_mask, z_mask) of a Pauli string given as {qubit: 'X'|'Y'|'Z'}."""
    x = z = 0
    for q, letter in spec.items():
        if letter in "XY":
            x |= 1 << q
        if letter in "ZY":
            z |= 1 << q
    return x, z


def random_paulis(num_qubits: int, count: int, seed: int, max_weight: int = 6,
                  min_weight: int = 1) -> list[tuple[int, int]]:
    """Random Pauli strings: weight uniform in [min_weight, max_weight] (capped
    at n), qubits drawn without replacement, letters uniform over X/Y/Z, all
    from ``mt19937_64(seed ^ 0x9E3779B97F4A7C15)``."""
    g = Mt19937_64(seed ^ 0x9E3779B97F4A7C15)
    out = []
    hi = min(max_weight, num_qubits)
    for _ in range(count):
        w = min_weight + g() % (hi - min_weight + 1)
        pool = list(range(num_qubits))
        spec = {}
        for _ in range(w):
            q = pool.pop(g() % len(pool))
            spec[q] = "XYZ"[g() % 3]
        out.append(pauli_masks(spec))
    return out


def z_terms(num_qubits: int) -> list[tuple[int, int]]:
    """<Z_i> for every qubit, as single-term observables."""
    return [(0, 1 << q) for q in range(num_qubits)]


def zz_neighbour_terms(num_qubits: int) -> list[tuple[int, int]]:
    """<Z_i Z_{i+1}> for i = 0 .. n-2."""
    return [(0, (1 << q) | (1 << (q + 1))) for q in range(num_qubits - 1)]


def fnv1a(data: bytes, h: int = 0xCBF29CE484222325) -> int:
    """FNV-1a 64 (reference tests/memory/memory_fixtures.cpp:599-645)."""
    for b in data:
        h ^= b
        h = (h * 0x100000001B3) & 0xFFFFFFFFFFFFFFFF
    return h


def digest_complex(values: np.ndarray) -> str:
    """The reference's statevector / expectation digest: FNV-1a over the
    (re, im) doubles in order. Byte-serial, so meant for small vectors."""
    buf = np.ascontiguousarray(values, dtype=np.complex128).tobytes()
    return "%016x" % fnv1a(buf)


def digest_counts(counts: dict[int, int]) -> str:
    """The reference's counts digest: per outcome in ascending order, the
    lowercase ``0x`` hex text of the outcome then its count as 8 raw bytes."""
    h = 0xCBF29CE484222325
    for key in sorted(counts):
        h = fnv1a(hex(key).encode(), h)
        h = fnv1a(int(counts[key]).to_bytes(8, "little"), h)
    return "%016x" % h


class FlatAnsatz:
    """``ansatz_circuit`` built directly in the flat layout (vectorised: a batch
    of thousands of circuits is built in well under a second)."""

    def __init__(self, num_qubits: int, depth: int, angles: np.ndarray):
        n, d = int(num_qubits), int(depth)
        self.num_qubits = n
        self.num_clbits = 0
        self.global_phase = 0.0
        per_layer = 2 * n + (n - 1)
        layer_ops = np.concatenate([np.full(n, RY), np.full(n, RZ), np.full(n - 1, CX)]).astype(np.int32)
        self._opcodes = np.tile(layer_ops, d)
        layer_q = [np.arange(n), np.arange(n)]
        ladder = np.stack([np.arange(n - 1), np.arange(1, n)], axis=1).reshape(-1)
        q_layer = np.concatenate([layer_q[0], layer_q[1], ladder]).astype(np.uint32)
        self._qubits = np.tile(q_layer, d)
        arity = np.concatenate([np.ones(2 * n, np.uint32), np.full(n - 1, 2, np.uint32)])
        self._qo = np.concatenate([[0], np.cumsum(np.tile(arity, d))]).astype(np.uint32)
        nparams = np.concatenate([np.ones(2 * n, np.uint32), np.zeros(n - 1, np.uint32)])
        self._po = np.concatenate([[0], np.cumsum(np.tile(nparams, d))]).astype(np.uint32)
        self._params = np.ascontiguousarray(angles, np.float64)
        assert self._params.size == 2 * n * d and self._opcodes.size == per_layer * d
        self.ops = range(self._opcodes.size)  # len() only
        self.depth = d

    @property
    def num_angles(self) -> int:
        return self._params.size

    def with_angles(self, angles) -> "FlatAnsatz":
        """The same circuit with new angles (a parameter-only update)."""
        return FlatAnsatz(self.num_qubits, self.depth, angles)

    def flat(self):
        return self._opcodes, self._qo, self._qubits, self._po, self._params

    def measure_targets(self):
        return []


def batch_workload(count: int, seed: int, qubits=(10, 22), depth=(4, 12), terms: int = 8, max_weight: int = 4):
    """The batched estimator workload: ``count`` ansatz circuits with n uniform
    on ``qubits`` and depth uniform on ``depth`` (inclusive), angles uniform on
    [-pi, pi) (numpy PCG64 from ``seed``), each with ``terms`` Pauli strings of
    weight 1..``max_weight`` from ``random_paulis(n, terms, seed + i)``.
    Returns (circuits, terms per circuit)."""
    rng = np.random.default_rng(seed)
    ns = rng.integers(qubits[0], qubits[1] + 1, count)
    ds = rng.integers(depth[0], depth[1] + 1, count)
    circuits, paulis = [], []
    for i in range(count):
        n, d = int(ns[i]), int(ds[i])
        angles = rng.uniform(-math.pi, math.pi, 2 * n * d)
        circuits.append(FlatAnsatz(n, d, angles))
        paulis.append(random_paulis(n, terms, seed + i, max_weight=max_weight, min_weight=1))
    return circuits, paulis


class FlatBrick:
    """``random_circuit``'s shape (U3 on every qubit, then a CZ brick, per
    layer) built directly in the flat layout, with the angles given
    (``3 * n * layers`` values in (layer, qubit, theta -> phi -> lambda)
    order), then a measurement of qubit q into clbit q for every qubit."""

    def __init__(self, num_qubits: int, layers: int, angles: np.ndarray):
        n, d = int(num_qubits), int(layers)
        self.num_qubits = n
        self.num_clbits = n
        self.global_phase = 0.0
        ops, qs, arity, npar = [], [], [], []
        for layer in range(d):
            ops.append(np.full(n, U))
            qs.append(np.arange(n))
            arity.append(np.ones(n, np.uint32))
            npar.append(np.full(n, 3, np.uint32))
            lo = np.arange(layer % 2, n - 1, 2)
            ops.append(np.full(lo.size, CZ))
            qs.append(np.stack([lo, lo + 1], axis=1).reshape(-1))
            arity.append(np.full(lo.size, 2, np.uint32))
            npar.append(np.zeros(lo.size, np.uint32))
        ops.append(np.full(n, MEASURE))
        qs.append(np.stack([np.arange(n), np.arange(n)], axis=1).reshape(-1))
        arity.append(np.full(n, 2, np.uint32))
        npar.append(np.zeros(n, np.uint32))
        self._opcodes = np.concatenate(ops).astype(np.int32)
        self._qubits = np.concatenate(qs).astype(np.uint32)
        self._qo = np.concatenate([[0], np.cumsum(np.concatenate(arity))]).astype(np.uint32)
        self._po = np.concatenate([[0], np.cumsum(np.concatenate(npar))]).astype(np.uint32)
        self._params = np.ascontiguousarray(angles, np.float64)
        assert self._params.size == 3 * n * d
        self.ops = range(self._opcodes.size)  # len() only

    def flat(self):
        return self._opcodes, self._qo, self._qubits, self._po, self._params

    def measure_targets(self):
        return [(q, q) for q in range(self.num_qubits)]


@dataclass
class ServeJobSpec:
    kind: str                 # "estimator" or "sampler"
    circuit: object           # FlatAnsatz or FlatBrick
    observables: list         # estimator: [[(x, z, coefficient), ...], ...]
    shots: int
    seed: int
    arrival: float            # seconds from the start
    structure: int            # index of the structure it shares (parameter-only update when reused)
    update: bool              # a parameter-only update of an earlier structure


def serve_workload(count: int, rate: float, seed: int, *, qubits=(12, 30), estimator_share=0.7,
                   update_share=0.4, observables=16, max_terms=4, max_weight=4, ansatz_depth=(4, 12),
                   sampler_layers=20, shots=1000, job_seed=2026):
    """The served workload: ``count`` jobs arriving as a Poisson process of
    ``rate`` jobs/s (exponential gaps, numpy PCG64 from ``seed``).

    * n is log-uniform on ``qubits``: ``floor(exp(U[ln lo, ln(hi + 1))))``;
    * ``estimator_share`` of new structures are estimator jobs: an ansatz
      (RY, RZ on every qubit, CNOT ladder; depth uniform on ``ansatz_depth``)
      with ``observables`` observables of 1..``max_terms`` Pauli terms (weight
      1..``max_weight``, coefficients uniform on [-1, 1]); the rest are sampler
      jobs: ``sampler_layers`` layers of U3 + CZ brick, measure all, ``shots``
      shots, seeds the SplitMix64 stream of ``job_seed`` (a sampler's pub seeds);
    * with probability ``update_share`` a job (after the first) is a
      parameter-only update: an earlier job's structure with fresh angles (and,
      for an estimator, the same observables).
    """
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, count)
    arrivals = np.cumsum(gaps) - gaps[0]
    lo, hi = qubits
    seeds = splitmix64_stream(job_seed, count)
    structures = []  # (kind, n, depth, observables)
    jobs = []
    for i in range(count):
        if structures and rng.random() < update_share:
            s = int(rng.integers(0, len(structures)))
            kind, n, depth, obs = structures[s]
            update = True
        else:
            n = int(math.floor(math.exp(rng.uniform(math.log(lo), math.log(hi + 1)))))
            n = min(max(n, lo), hi)
            if rng.random() < estimator_share:
                kind = "estimator"
                depth = int(rng.integers(ansatz_depth[0], ansatz_depth[1] + 1))
                obs = []
                for _ in range(observables):
                    terms = []
                    for _ in range(int(rng.integers(1, max_terms + 1))):
                        w = int(rng.integers(1, min(max_weight, n) + 1))
                        qs = rng.choice(n, w, replace=False)
                        spec = {int(q): "XYZ"[int(rng.integers(0, 3))] for q in qs}
                        x, z = pauli_masks(spec)
                        terms.append((x, z, float(rng.uniform(-1.0, 1.0))))
                    obs.append(terms)
            else:
                kind, depth, obs = "sampler", sampler_layers, []
            structures.append((kind, n, depth, obs))
            s = len(structures) - 1
            update = False
        if kind == "estimator":
            c = FlatAnsatz(n, depth, rng.uniform(-math.pi, math.pi, 2 * n * depth))
            jobs.append(ServeJobSpec(kind, c, obs, 0, 0, float(arrivals[i]), s, update))
        else:
            c = FlatBrick(n, depth, rng.uniform(0.0, 2.0 * math.pi, 3 * n * depth))
            jobs.append(ServeJobSpec(kind, c, [], shots, seeds[i], float(arrivals[i]), s, update))
    return jobs
