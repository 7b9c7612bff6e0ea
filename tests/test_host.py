"""Host-side checks of the product library that need no GPU: it loads, it
exports every entry point include/menos_hip.h declares, and its host plan
compilation (gate table, fusion plans) matches the oracle and the reference."""
import os
import re

import numpy as np
import pytest

from circuits import diagonal_heavy_circuit, random_mixed_circuit
from conftest import ROOT, circuit_from_spec, cpx
from menos_amd import hip as H
from menos_amd import workloads as W


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "menos_hip.h")).read()
    return sorted(set(re.findall(r"^(?:int|const char\*)\s+(mh_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    lib = H.lib()
    names = declared_symbols()
    assert len(names) >= 20
    for name in names:
        assert hasattr(lib, name), name
    assert sorted(H.EXPORTS) == names


def test_library_is_built_for_gfx950():
    data = open(H.LIB_PATH, "rb").read()
    assert b"gfx950" in data


def test_gate_matrices_match_reference_bits(golden):
    for g in golden["gate_matrices"]:
        m = H.gate_matrix(g["opcode"], g["params"]).reshape(-1)
        assert m.tobytes() == cpx(g["matrix"]).tobytes(), g["opcode"]


@pytest.mark.parametrize("opcode", [W.U, W.RX, W.RY, W.RZ, W.P, W.CP, W.RZZ])
def test_gate_matrices_match_oracle_over_many_angles(orc, opcode):
    """The host gate table must be the reference's bits for every angle: a
    compiler that lowers std::exp(complex) differently from g++/glibc cexp
    differs in the last bit for a few angles in a thousand."""
    rng = np.random.default_rng(opcode)
    for _ in range(3000):
        params = rng.uniform(0, 2 * np.pi, size=W.GATE_PARAMS[opcode]).tolist()
        assert H.gate_matrix(opcode, params).tobytes() == orc.gate_matrix(opcode, params).tobytes(), params


def test_gate_matrix_rejects_bad_input():
    with pytest.raises(H.HipError):
        H.gate_matrix(W.UNITARY)
    with pytest.raises(H.HipError):
        H.gate_matrix(99)


def test_containment_plans_match_reference(golden):
    for p in golden["plans"]:
        c = circuit_from_spec(p["circuit"])
        order, bb, bq, qb, fb, ga = H.plan_fusion(c, H.FUSE_CONTAINMENT)
        assert order.tolist() == p["order"]
        assert bb.tolist() == p["block_begin"]
        assert bq.tolist() == p["block_qubits"]
        assert qb.tolist() == p["qubits_begin"]
        assert (fb, ga) == (p["fused_blocks"], p["gates_absorbed"])


def test_containment_plans_match_oracle(orc):
    for seed in range(10):
        for c in (random_mixed_circuit(9, 80, 700 + seed, measures=True), diagonal_heavy_circuit(7, 80, 800 + seed)):
            order, bb, bq, qb, fb, ga = H.plan_fusion(c, H.FUSE_CONTAINMENT)
            p = orc.plan_fusion(c)
            assert order.tolist() == p.order.tolist()
            assert bb.tolist() == p.block_begin.tolist()
            assert bq.tolist() == p.block_qubits.tolist()


@pytest.mark.parametrize("width", [1, 2, 3, 4, 5, 6])
def test_width_plans_are_contiguous_and_capped(width):
    c = random_mixed_circuit(10, 120, 900 + width, measures=True, max_unitary=min(width, 5))
    order, bb, bq, qb, _, _ = H.plan_fusion(c, H.FUSE_WIDTH, width)
    assert order.tolist() == list(range(len(c.ops)))  # order-preserving
    for b in range(len(bb) - 1):
        k = qb[b + 1] - qb[b]
        members = order[bb[b]:bb[b + 1]].tolist()
        if any(c.ops[i][0] == W.MEASURE for i in members):
            assert members and len(members) == 1 and k == 0
            continue
        wide = max(len(c.ops[i][1]) for i in members)
        assert k <= max(width, wide)
        union = set()
        for i in members:
            union |= set(c.ops[i][1])
        assert sorted(union) == bq[qb[b]:qb[b + 1]].tolist()


def test_unfused_plan_is_one_block_per_op():
    c = random_mixed_circuit(6, 30, 3, measures=True)
    order, bb, _, _, fb, ga = H.plan_fusion(c, H.FUSE_NONE)
    assert order.tolist() == list(range(30)) and bb.tolist() == list(range(31)) and fb == ga == 0


def test_plan_rejects_invalid_circuits():
    bad = W.Circuit(3).append(W.CX, (0, 0))
    with pytest.raises(H.HipError, match="repeated operand"):
        H.plan_fusion(bad)
    bad = W.Circuit(3).append(W.H, (3,))
    with pytest.raises(H.HipError, match="out of range"):
        H.plan_fusion(bad)
    bad = W.Circuit(3).append(W.RX, (0,), ())
    with pytest.raises(H.HipError, match="parameter"):
        H.plan_fusion(bad)


def test_workload_generators_are_deterministic():
    a, b = W.random_circuit(7, 3, 5), W.random_circuit(7, 3, 5)
    assert a.ops == b.ops and a.ops != W.random_circuit(7, 3, 6).ops
    assert len(a.ops) == 3 * 7 + 3 + 3 + 3  # U on 7 qubits per layer, CZ brick 3/3/3
    p = W.random_paulis(12, 64, 9)
    assert len(p) == 64 and all(1 <= bin(x | z).count("1") <= 6 for x, z in p)
    assert p == W.random_paulis(12, 64, 9)
