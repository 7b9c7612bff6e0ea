"""ctypes front of the oracle (``liboracle.so``) and of the reference build
(``_ref/libmenos_ref.so``).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
and by bench.py's cpu_baseline leg, always as the checker or the CPU baseline,
never as the thing measured. The product package (menos_amd) never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libmenos_ref.so")

_u32p = ctypes.POINTER(ctypes.c_uint32)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_i32p = ctypes.POINTER(ctypes.c_int32)
_f64p = ctypes.POINTER(ctypes.c_double)


class CCircuit(ctypes.Structure):
    _fields_ = [("num_qubits", ctypes.c_uint32), ("num_ops", ctypes.c_uint32),
                ("opcodes", _i32p), ("qubit_offsets", _u32p), ("qubits", _u32p),
                ("param_offsets", _u32p), ("params", _f64p)]


class CPlan(ctypes.Structure):
    _fields_ = [("num_blocks", ctypes.c_uint32), ("order", _u32p), ("block_begin", _u32p),
                ("block_qubits", _u32p), ("qubits_begin", _u32p)]


class CRequest(ctypes.Structure):
    _fields_ = [("circuit", ctypes.POINTER(CCircuit)), ("plan", ctypes.POINTER(CPlan)),
                ("fuse_gates", ctypes.c_int32), ("global_phase", ctypes.c_double),
                ("shots", ctypes.c_uint64), ("seed", ctypes.c_uint64), ("threads", ctypes.c_uint32),
                ("num_terms", ctypes.c_uint32), ("x_masks", _u64p), ("z_masks", _u64p)]


def _ptr(a, t):
    return a.ctypes.data_as(t) if a is not None and a.size else ctypes.cast(None, t)


def build() -> None:
    """Build liboracle.so (and _ref when the reference sources are present)."""
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)


def _load(path):
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} is not built (run `make -C oracle`)")
    return ctypes.CDLL(path)


class _Keep:
    """Holds the numpy arrays a ctypes struct points into."""

    def __init__(self, *arrays):
        self.arrays = arrays


def c_circuit(circuit):
    opcodes, qo, qs, po, ps = circuit.flat()
    keep = _Keep(opcodes, qo, qs, po, ps)
    cc = CCircuit(circuit.num_qubits, len(opcodes), _ptr(opcodes, _i32p), _ptr(qo, _u32p), _ptr(qs, _u32p),
                  _ptr(po, _u32p), _ptr(ps, _f64p))
    return cc, keep


class Plan:
    """A fusion plan in the reference's FusionPlan layout."""

    def __init__(self, order, block_begin, block_qubits, qubits_begin, fused_blocks=0, gates_absorbed=0):
        self.order = np.asarray(order, dtype=np.uint32)
        self.block_begin = np.asarray(block_begin, dtype=np.uint32)
        self.block_qubits = np.asarray(block_qubits, dtype=np.uint32)
        self.qubits_begin = np.asarray(qubits_begin, dtype=np.uint32)
        self.fused_blocks, self.gates_absorbed = fused_blocks, gates_absorbed

    @property
    def num_blocks(self):
        return len(self.block_begin) - 1

    def c(self):
        return CPlan(self.num_blocks, _ptr(self.order, _u32p), _ptr(self.block_begin, _u32p),
                     _ptr(self.block_qubits, _u32p), _ptr(self.qubits_begin, _u32p))

    def blocks(self):
        for b in range(self.num_blocks):
            yield (self.order[self.block_begin[b]:self.block_begin[b + 1]].tolist(),
                   self.block_qubits[self.qubits_begin[b]:self.qubits_begin[b + 1]].tolist())

    def __eq__(self, other):
        return all(np.array_equal(getattr(self, f), getattr(other, f))
                   for f in ("order", "block_begin", "block_qubits", "qubits_begin"))


class _Api:
    """Shared entry points of liboracle.so and libmenos_ref.so."""

    prefix = ""

    def __init__(self, path):
        self.lib = _load(path)

    def _fn(self, name):
        return getattr(self.lib, self.prefix + name)

    def plan_fusion(self, circuit) -> Plan:
        cc, keep = c_circuit(circuit)
        n_ops = cc.num_ops
        nq = int(sum(len(o[1]) for o in circuit.ops)) + 1
        order = np.zeros(max(n_ops, 1), np.uint32)
        bb = np.zeros(n_ops + 2, np.uint32)
        bq = np.zeros(max(nq, 1), np.uint32)
        qb = np.zeros(n_ops + 2, np.uint32)
        fb, ga = ctypes.c_uint64(), ctypes.c_uint64()
        fn = self._fn("plan_fusion")
        fn.restype = ctypes.c_uint32
        nb = fn(ctypes.byref(cc), _ptr(order, _u32p), _ptr(bb, _u32p), _ptr(bq, _u32p), _ptr(qb, _u32p),
                ctypes.byref(fb), ctypes.byref(ga))
        del keep
        return Plan(order[:n_ops], bb[:nb + 1], bq[:qb[nb]], qb[:nb + 1], fb.value, ga.value)

    def fused_block_matrix(self, circuit, plan: Plan, block: int) -> np.ndarray:
        cc, keep = c_circuit(circuit)
        cp = plan.c()
        k = int(plan.qubits_begin[block + 1] - plan.qubits_begin[block])
        out = np.zeros((1 << k) * (1 << k), np.complex128)
        self._fn("fused_block_matrix")(ctypes.byref(cc), ctypes.byref(cp), ctypes.c_uint32(block),
                                       out.ctypes.data_as(_f64p))
        del keep
        return out.reshape(1 << k, 1 << k)

    def expectation_of_term(self, state: np.ndarray, x: int, z: int) -> complex:
        n = int(np.log2(state.size))
        out = np.zeros(2, np.float64)
        self._fn("expectation_of_term")(state.ctypes.data_as(_f64p), ctypes.c_uint32(n), ctypes.c_uint64(x),
                                        ctypes.c_uint64(z), out.ctypes.data_as(_f64p))
        return complex(out[0], out[1])


class Oracle(_Api):
    prefix = "orc_"

    def __init__(self):
        super().__init__(LIB_PATH)

    def gate_matrix(self, opcode: int, params=()) -> np.ndarray:
        p = np.asarray(params, np.float64)
        out = np.zeros(64, np.complex128)
        k = self.lib.orc_gate_matrix(ctypes.c_int32(opcode), _ptr(p, _f64p), out.ctypes.data_as(_f64p))
        if k < 0:
            raise ValueError(f"opcode {opcode} has no gate-table entry")
        return out[:(1 << k) ** 2].reshape(1 << k, 1 << k)

    def initialize(self, num_qubits: int, phase: float = 0.0) -> np.ndarray:
        s = np.zeros(1 << num_qubits, np.complex128)
        self.lib.orc_initialize(s.ctypes.data_as(_f64p), ctypes.c_uint64(s.size), ctypes.c_double(phase))
        return s

    def apply_generic(self, state: np.ndarray, matrix: np.ndarray, operands, threads: int = 1) -> None:
        n = int(np.log2(state.size))
        m = np.ascontiguousarray(matrix, np.complex128)
        ops = np.asarray(operands, np.uint32)
        self.lib.orc_apply_generic(state.ctypes.data_as(_f64p), ctypes.c_uint32(n), m.ctypes.data_as(_f64p),
                                   _ptr(ops, _u32p), ctypes.c_uint32(len(ops)), ctypes.c_uint32(threads))

    def sample_indices(self, state: np.ndarray, shots: int, seed: int) -> np.ndarray:
        n = int(np.log2(state.size))
        out = np.zeros(max(shots, 1), np.uint64)
        self.lib.orc_sample_indices(state.ctypes.data_as(_f64p), ctypes.c_uint32(n), ctypes.c_uint64(shots),
                                    ctypes.c_uint64(seed), out.ctypes.data_as(_u64p))
        return out[:shots]

    def sample_indices_from_probabilities(self, probs: np.ndarray, shots: int, seed: int) -> np.ndarray:
        p = np.ascontiguousarray(probs, np.float64)
        out = np.zeros(max(shots, 1), np.uint64)
        self.lib.orc_sample_indices_from_probabilities(p.ctypes.data_as(_f64p), ctypes.c_uint64(p.size),
                                                       ctypes.c_uint64(shots), ctypes.c_uint64(seed),
                                                       out.ctypes.data_as(_u64p))
        return out[:shots]

    def cumulative(self, state: np.ndarray) -> np.ndarray:
        n = int(np.log2(state.size))
        out = np.zeros(state.size, np.float64)
        self.lib.orc_cumulative(state.ctypes.data_as(_f64p), ctypes.c_uint32(n), out.ctypes.data_as(_f64p))
        return out

    def region_draw_indices(self, state: np.ndarray, draws: int, seed: int) -> np.ndarray:
        n = int(np.log2(state.size))
        out = np.zeros(max(draws, 1), np.uint64)
        self.lib.orc_region_draw_indices(state.ctypes.data_as(_f64p), ctypes.c_uint32(n), ctypes.c_uint64(draws),
                                         ctypes.c_uint64(seed), out.ctypes.data_as(_u64p))
        return out[:draws]

    def mt19937_64(self, seed: int, count: int) -> np.ndarray:
        out = np.zeros(count, np.uint64)
        self.lib.orc_mt19937_64(ctypes.c_uint64(seed), ctypes.c_uint64(count), out.ctypes.data_as(_u64p))
        return out

    def fnv1a(self, data: np.ndarray, h: int = 0xCBF29CE484222325) -> int:
        buf = np.ascontiguousarray(data)
        self.lib.orc_fnv1a.restype = ctypes.c_uint64
        return int(self.lib.orc_fnv1a(buf.ctypes.data_as(ctypes.c_void_p), ctypes.c_uint64(buf.nbytes),
                                      ctypes.c_uint64(h)))

    def execute(self, circuit, *, fuse_gates=True, plan: Plan | None = None, shots=0, seed=0, terms=(),
                statevector=True, threads=8):
        """The whole engine: returns (statevector or None, basis index per sorted shot, expectation terms)."""
        cc, keep = c_circuit(circuit)
        xs = np.array([t[0] for t in terms], np.uint64)
        zs = np.array([t[1] for t in terms], np.uint64)
        cp = plan.c() if plan is not None else None
        req = CRequest(ctypes.pointer(cc), ctypes.pointer(cp) if cp is not None else None, int(bool(fuse_gates)),
                       circuit.global_phase, shots, seed, threads, len(terms), _ptr(xs, _u64p), _ptr(zs, _u64p))
        sv = np.zeros(1 << circuit.num_qubits, np.complex128) if statevector else None
        idx = np.zeros(max(shots, 1), np.uint64) if shots else None
        ex = np.zeros(2 * max(len(terms), 1), np.float64) if terms else None
        self.lib.orc_execute(ctypes.byref(req), sv.ctypes.data_as(_f64p) if sv is not None else None,
                             idx.ctypes.data_as(_u64p) if idx is not None else None,
                             ex.ctypes.data_as(_f64p) if ex is not None else None)
        del keep
        e = ex.view(np.complex128).copy() if ex is not None else np.zeros(0, np.complex128)
        return sv, (idx[:shots] if idx is not None else np.zeros(0, np.uint64)), e[:len(terms)]


class Reference(_Api):
    """The reference's own code, compiled by oracle/Makefile into _ref/."""

    prefix = "ref_"

    def __init__(self):
        super().__init__(REF_PATH)
        self.lib.ref_execute.restype = ctypes.c_double
        self.lib.ref_run_plan.restype = ctypes.c_double
        self.lib.ref_sample_counts.restype = ctypes.c_uint64

    @staticmethod
    def available() -> bool:
        return os.path.exists(REF_PATH)

    def gate_matrix(self, opcode: int, params=()) -> np.ndarray:
        p = np.asarray(params, np.float64)
        out = np.zeros(64, np.complex128)
        size = self.lib.ref_gate_matrix(ctypes.c_int32(opcode), _ptr(p, _f64p), ctypes.c_uint32(p.size),
                                        out.ctypes.data_as(_f64p))
        d = int(round(size ** 0.5))
        return out[:size].reshape(d, d)

    def apply_matrix(self, state, matrix, operands, threads=1, generic=False):
        n = int(np.log2(state.size))
        m = np.ascontiguousarray(matrix, np.complex128)
        ops = np.asarray(operands, np.uint32)
        fn = self.lib.ref_apply_matrix_generic if generic else self.lib.ref_apply_matrix
        fn(state.ctypes.data_as(_f64p), ctypes.c_uint32(n), m.ctypes.data_as(_f64p), _ptr(ops, _u32p),
           ctypes.c_uint32(len(ops)), ctypes.c_uint32(threads))

    def tier_pass_plan(self, circuit, chunk_qubits, window_qubits, fuse=False, choose=False, capacity=4096):
        """The reference disk tier's pass plan (build_pass_plan, or
        choose_pass_plan with choose): dict(chunk, passes=[(mask, first, end)],
        global_gates, max_global, too_small), or None when no chunk fits."""
        cc, keep = c_circuit(circuit)
        masks = np.zeros(capacity, np.uint64)
        first = np.zeros(capacity, np.uint32)
        end = np.zeros(capacity, np.uint32)
        chunk, gg, mg, small = ctypes.c_uint32(), ctypes.c_uint64(), ctypes.c_uint32(), ctypes.c_int32()
        fn = self.lib.ref_tier_pass_plan
        fn.restype = ctypes.c_uint32
        count = fn(ctypes.byref(cc), ctypes.c_uint32(chunk_qubits), ctypes.c_uint32(window_qubits),
                   ctypes.c_int32(int(fuse)), ctypes.c_int32(int(choose)), _ptr(masks, _u64p), _ptr(first, _u32p),
                   _ptr(end, _u32p), ctypes.c_uint32(capacity), ctypes.byref(chunk), ctypes.byref(gg),
                   ctypes.byref(mg), ctypes.byref(small))
        del keep
        if count == 0xFFFFFFFF:
            return None
        m = min(count, capacity)
        return dict(chunk=chunk.value, count=count,
                    passes=[(int(masks[i]), int(first[i]), int(end[i])) for i in range(m)],
                    global_gates=gg.value, max_global=mg.value, too_small=bool(small.value))

    def run_plan(self, circuit, plan: Plan, threads=8):
        cc, keep = c_circuit(circuit)
        cp = plan.c()
        out = np.zeros(1 << circuit.num_qubits, np.complex128)
        secs = self.lib.ref_run_plan(ctypes.byref(cc), ctypes.byref(cp), ctypes.c_double(circuit.global_phase),
                                     ctypes.c_uint32(threads), out.ctypes.data_as(_f64p))
        del keep
        return out, secs

    def sample_counts(self, state, shots, seed, targets, width):
        n = int(np.log2(state.size))
        tq = np.array([t[0] for t in targets], np.uint32)
        tc = np.array([t[1] for t in targets], np.uint32)
        outc = np.zeros(max(shots, 1), np.uint64)
        cnt = np.zeros(max(shots, 1), np.uint64)
        k = self.lib.ref_sample_counts(state.ctypes.data_as(_f64p), ctypes.c_uint32(n), ctypes.c_uint64(shots),
                                       ctypes.c_uint64(seed), _ptr(tq, _u32p), _ptr(tc, _u32p),
                                       ctypes.c_uint32(len(targets)), ctypes.c_uint32(width),
                                       outc.ctypes.data_as(_u64p), cnt.ctypes.data_as(_u64p))
        return {int(outc[i]): int(cnt[i]) for i in range(k)}

    def region_draws(self, state, draws, seed):
        n = int(np.log2(state.size))
        out = np.zeros(max(draws, 1), np.uint64)
        self.lib.ref_region_draws(state.ctypes.data_as(_f64p), ctypes.c_uint32(n), ctypes.c_uint64(draws),
                                  ctypes.c_uint64(seed), out.ctypes.data_as(_u64p))
        return out[:draws]

    def execute(self, circuit, *, fuse_gates=True, shots=0, seed=0, terms=(), statevector=True, threads=8,
                targets=None, width=None):
        """DenseEngine::execute. Returns (statevector, counts dict, expectation terms, seconds)."""
        cc, keep = c_circuit(circuit)
        xs = np.array([t[0] for t in terms], np.uint64)
        zs = np.array([t[1] for t in terms], np.uint64)
        req = CRequest(ctypes.pointer(cc), None, int(bool(fuse_gates)), circuit.global_phase, shots, seed,
                       threads, len(terms), _ptr(xs, _u64p), _ptr(zs, _u64p))
        targets = targets if targets is not None else circuit.measure_targets()
        width = width if width is not None else circuit.num_clbits
        tq = np.array([t[0] for t in targets], np.uint32)
        tc = np.array([t[1] for t in targets], np.uint32)
        sv = np.zeros(1 << circuit.num_qubits, np.complex128) if statevector else None
        outc = np.zeros(max(shots, 1), np.uint64) if shots else None
        cnt = np.zeros(max(shots, 1), np.uint64) if shots else None
        nout = ctypes.c_uint64(0)
        ex = np.zeros(2 * max(len(terms), 1), np.float64) if terms else None
        secs = self.lib.ref_execute(ctypes.byref(req), _ptr(tq, _u32p), _ptr(tc, _u32p), ctypes.c_uint32(len(targets)),
                                    ctypes.c_uint32(width), sv.ctypes.data_as(_f64p) if sv is not None else None,
                                    outc.ctypes.data_as(_u64p) if outc is not None else None,
                                    cnt.ctypes.data_as(_u64p) if cnt is not None else None, ctypes.byref(nout),
                                    ex.ctypes.data_as(_f64p) if ex is not None else None)
        del keep
        counts = {int(outc[i]): int(cnt[i]) for i in range(nout.value)} if shots else {}
        e = ex.view(np.complex128).copy() if ex is not None else np.zeros(0, np.complex128)
        return sv, counts, e[:len(terms)], secs


def counts_from_indices(indices, targets, width=None) -> dict[int, int]:
    """Outcome histogram from per-shot basis indices: bit `clbit` of an outcome
    is bit `qubit` of the basis index (reference sample.cpp:45-53)."""
    counts: dict[int, int] = {}
    for b in np.asarray(indices, np.uint64).tolist():
        key = 0
        for q, c in targets:
            if (b >> q) & 1:
                key |= 1 << c
        counts[key] = counts.get(key, 0) + 1
    return counts


def combine_expectation(raw_terms, coefficients) -> complex:
    """One observable's value from its raw, coefficient-free terms (reference
    src/runtime/combine_expectation.cpp:7-25, one region): total starts at
    (0, 0); per term in order, product = (1, 0) * term and
    total += coefficient * product, each complex product the naive
    (a*c - b*d, a*d + b*c) of std::complex on finite values. Returns
    (re, im) as Python floats (IEEE doubles, the same operations)."""
    tr, ti = 0.0, 0.0
    for term, coef in zip(raw_terms, coefficients):
        er, ei = float(complex(term).real), float(complex(term).imag)
        pr = 1.0 * er - 0.0 * ei
        pi = 1.0 * ei + 0.0 * er
        cr, ci = float(complex(coef).real), float(complex(coef).imag)
        tr = tr + (cr * pr - ci * pi)
        ti = ti + (cr * pi + ci * pr)
    return tr, ti


def splitmix64(seed: int, count: int) -> list[int]:
    """The first `count` outputs of SplitMix64 from `seed` (reference
    src/common/splitmix64.h:15-28; Vigna's published splitmix64.c)."""
    mask = (1 << 64) - 1
    state, out = seed & mask, []
    for _ in range(count):
        state = (state + 0x9E3779B97F4A7C15) & mask
        z = state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & mask
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & mask
        out.append(z ^ (z >> 31))
    return out
