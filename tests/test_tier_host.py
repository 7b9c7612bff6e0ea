"""The host-offloaded tier's plan, on the CPU (mh_tier_estimate does host work
only): pass counts pinned by the reference's own pass-plan tests, the greedy
rule and the chunk choice equal to the reference's compiled pass_plan.cpp on
random circuits, the window derived from the budget, the link accounting, and
the infeasible cases."""
import numpy as np
import pytest

from menos_amd import hip as H
from menos_amd import workloads as W

ENGINE = H.HipHostTierEngine()


def crossover(n):
    """The reference's crossover family (benchmarks/crossover_circuit.h:17-31):
    6 layers of H on every qubit, T on qubit 0, CX on alternating pairs."""
    c = W.Circuit(n)
    for layer in range(6):
        for q in range(n):
            c.append(W.H, (q,))
        c.append(W.T, (0,))
        for q in range(layer % 2, n - 1, 2):
            c.append(W.CX, (q, q + 1))
    return c


def ladder(n):
    """The ladder family of pass_plan_test.cpp:22-33: 6 layers of H on every
    qubit then a CX ladder."""
    c = W.Circuit(n)
    for _ in range(6):
        for q in range(n):
            c.append(W.H, (q,))
        for q in range(n - 1):
            c.append(W.CX, (q, q + 1))
    return c


def passes(circuit, chunk, window, **kw):
    cost = ENGINE.estimate(circuit, tier_chunk_qubits=chunk, tier_window_qubits=window, fusion=H.FUSE_NONE, **kw)
    assert cost.feasible, cost.infeasible_reason
    return cost


# (n, chunk, window) -> passes: the figures of the reference's
# StepOnePassCounts tests (tests/engines/dense_tiered/pass_plan_test.cpp:148-173).
CROSSOVER = {(31, 27, 29): 24, (31, 27, 30): 13, (31, 27, 31): 1, (31, 28, 30): 15, (31, 25, 27): 36,
             (31, 25, 29): 18, (31, 24, 28): 21, (33, 28, 30): 30, (33, 27, 29): 36, (33, 24, 28): 27}
LADDER = {(31, 27, 29): 30, (31, 28, 30): 24, (31, 24, 28): 24, (31, 24, 26): 60}


@pytest.mark.parametrize("key", sorted(CROSSOVER))
def test_crossover_pass_counts_match_the_references_figures(key):
    n, chunk, window = key
    circuit = crossover(n)
    if n == 31:
        assert len(circuit.ops) == 282
    cost = passes(circuit, chunk, window)
    assert cost.passes == CROSSOVER[key]
    if key == (31, 27, 29):
        assert cost.global_blocks == 36 and cost.max_global_operands == 2


@pytest.mark.parametrize("key", sorted(LADDER))
def test_ladder_pass_counts_match_the_references_figures(key):
    n, chunk, window = key
    circuit = ladder(n)
    assert len(circuit.ops) == 366
    assert passes(circuit, chunk, window).passes == LADDER[key]


def test_the_chunk_with_the_fewest_passes_is_chosen():
    # pass_plan_test.cpp:178-186: in a 2^29 window the n = 31 crossover circuit
    # takes 18 passes at L = 25 against 24 at L = 27.
    cost = ENGINE.estimate(crossover(31), tier_window_qubits=29, fusion=H.FUSE_NONE)
    assert (cost.chunk_qubits, cost.window_qubits, cost.passes) == (25, 29, 18)
    # A window holding the whole state is one pass at the largest chunk.
    whole = ENGINE.estimate(crossover(12), tier_window_qubits=12, fusion=H.FUSE_NONE)
    assert (whole.chunk_qubits, whole.passes) == (12, 1)


def test_the_choice_agrees_with_the_reference_above_the_preferred_minimum(ref):
    # Windows above 2^24: the reference's choose_pass_plan itself.
    for n, window in ((31, 29), (31, 27), (33, 30), (30, 26)):
        c = crossover(n)
        want = ref.tier_pass_plan(c, 0, window, fuse=True, choose=True)
        got = ENGINE.estimate(c, tier_window_qubits=window)
        assert (got.chunk_qubits, got.passes) == (want["chunk"], want["count"])


def test_chunks_below_the_preferred_minimum_only_when_nothing_above_fits():
    # pass_plan_test.cpp:245-262.
    c = W.Circuit(30)
    c.append_unitary([24, 25, 26, 27, 28, 29], np.eye(64))
    cost = ENGINE.estimate(c, tier_window_qubits=27)
    assert (cost.chunk_qubits, cost.passes) == (21, 1)
    easy = W.Circuit(30).append(W.CX, (28, 29))
    assert ENGINE.estimate(easy, tier_window_qubits=27).chunk_qubits == 25


def test_the_greedy_rule_extends_a_set_while_it_fits():
    # pass_plan_test.cpp:79-106, shifted up by 10 qubits (the tier needs
    # 10 qubits): globals 13, 14, 15 over a chunk of 13, w = 2.
    c = W.Circuit(16, 16)
    c.append(W.CX, (0, 13)).append(W.H, (1,)).append(W.CX, (14, 1)).append_measure(0, 0)
    c.append(W.CX, (15, 0)).append(W.CX, (13, 15)).append_measure(1, 1)
    cost = passes(c, 13, 15)
    # Blocks are the gates (measurements carry no block): [0, 3) and [3, 5).
    assert cost.plan == [(0b011, 0, 3), (0b101, 3, 5)]
    assert cost.global_blocks == 4 and cost.max_global_operands == 2
    assert passes(c, 13, 16).plan == [(0b111, 0, 5)]


def test_a_block_wider_than_the_windows_globals_is_infeasible():
    c = W.Circuit(18)
    c.append_unitary([13, 14, 15, 16, 17], np.eye(32))
    small = ENGINE.estimate(c, tier_chunk_qubits=13, tier_window_qubits=17)
    assert not small.feasible and b"operands above the chunk" in small.infeasible_reason
    assert passes(c, 13, 18).plan == [(0b11111, 0, 1)]


def _random(n, layers, seed, measure=False):
    c = W.random_circuit(n, layers, seed)
    if not measure:
        c.ops = [op for op in c.ops if op[0] != W.MEASURE]
    return c


@pytest.mark.parametrize("seed", range(6))
def test_plans_equal_the_reference_pass_plan(ref, seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(12, 22))
    c = _random(n, int(rng.integers(2, 6)), seed)
    for _ in range(6):
        window = int(rng.integers(10, n + 1))
        chunk = int(rng.integers(max(0, window - 6), window + 1))
        # Over the gates: the same passes, index for index.
        want = ref.tier_pass_plan(c, chunk, window)
        got = ENGINE.estimate(c, tier_chunk_qubits=chunk, tier_window_qubits=window, fusion=H.FUSE_NONE)
        assert bool(got.feasible) == (not want["too_small"])
        if got.feasible:
            assert got.max_global_operands == want["max_global"]
            assert got.global_blocks == want["global_gates"]
            assert got.plan == want["passes"]
        # Over the containment fusion's blocks.
        want = ref.tier_pass_plan(c, chunk, window, fuse=True)
        got = ENGINE.estimate(c, tier_chunk_qubits=chunk, tier_window_qubits=window)
        if got.feasible:
            assert got.plan == want["passes"]
        # The chosen chunk: the reference's choice rule (fewest passes, the
        # larger chunk on a tie) over its own plans, for chunks down to
        # window - 6 (this tier's floor: at most 64 chunks to a window).
        best = None
        for chunk in range(window, max(0, window - 6) - 1, -1):
            plan = ref.tier_pass_plan(c, chunk, window, fuse=True)
            if not plan["too_small"] and (best is None or plan["count"] < best[1]):
                best = (chunk, plan["count"])
        got = ENGINE.estimate(c, tier_window_qubits=window)
        assert got.feasible and best is not None
        assert (got.chunk_qubits, got.passes) == best


def test_the_window_is_the_largest_the_budget_holds():
    c = _random(24, 3, 1)
    for precision, amp in ((H.C128, 16), (H.C64, 8)):
        for window in (12, 17, 20):
            need = ENGINE.estimate(c, tier_window_qubits=window, precision=precision, shots=100, terms=[(1, 2)])
            assert need.feasible
            assert need.device_bytes >= 2 * amp << window
            got = ENGINE.estimate(c, budget_bytes=need.device_bytes, precision=precision, shots=100, terms=[(1, 2)])
            assert got.window_qubits == window and got.device_bytes <= need.device_bytes
    none = ENGINE.estimate(c)
    assert none.window_qubits == 24  # no budget: min(n, 28)
    tight = ENGINE.estimate(c, budget_bytes=2 * 16 << 9)
    assert not tight.feasible and b"smallest windows" in tight.infeasible_reason


def test_link_accounting():
    c = _random(20, 4, 3)
    sb = 16 << 20
    base = ENGINE.estimate(c, tier_window_qubits=14)
    p = base.passes
    assert p > 1 and base.host_state_bytes == sb
    assert base.link_write_bytes == p * sb
    assert base.link_bytes == (2 * p - 1) * sb
    # Counts and in-window terms share one read; each value of x >> (W - 1) of
    # a term reaching above the window takes one more.
    x_hi = (1 << 17) | 1
    more = ENGINE.estimate(c, tier_window_qubits=14, shots=10, terms=[(3, 5), (x_hi, 0), (x_hi, 4), (1 << 19, 0)])
    assert more.link_bytes == (2 * p - 1 + 3) * sb
    assert more.seconds > base.seconds > 0
    empty = ENGINE.estimate(W.Circuit(12), tier_window_qubits=10)
    assert empty.passes == 0 and empty.link_bytes == empty.link_write_bytes == 16 << 12


def test_refusals():
    assert not ENGINE.estimate(W.Circuit(9)).feasible
    c = _random(14, 2, 0)
    bad = ENGINE.estimate(c, tier_window_qubits=15)
    assert not bad.feasible and b"window_qubits" in bad.infeasible_reason
    assert not ENGINE.estimate(c, tier_chunk_qubits=13, tier_window_qubits=12).feasible
    assert not ENGINE.estimate(c, terms=[(1 << 14, 0)]).feasible


def test_time_model_under_a_synthetic_profile():
    """With the window passes priced at nothing, a pass takes its copies: the
    first only writes the state back, later ones read it and write it back at
    once (the two directions overlap), and each output read takes one state's
    copy in."""
    old = H.get_calibration()
    profile = dict(hbm_bytes_per_second=6.0e12, pass_bytes_per_second=1.0e30, launch_seconds=0.0,
                   h2d_bytes_per_second=5.0e10, d2h_bytes_per_second=4.0e10, fp64_ops_per_second=3.0e13,
                   mfma_flops_per_second=1.0e14)
    H.set_calibration(H.Calibration(**profile))
    try:
        c = _random(22, 3, 4)
        sb = 16 << 22
        for shots, terms, reads in ((0, (), 0), (100, [(1, 2)], 1), (100, [(1, 2), ((1 << 21) | 1, 0)], 2)):
            e = ENGINE.estimate(c, tier_window_qubits=16, shots=shots, terms=terms)
            p = e.passes
            want = sb / 4.0e10 + (p - 1) * max(sb / 5.0e10, sb / 4.0e10) + reads * sb / 5.0e10
            assert e.seconds == pytest.approx(want, rel=1e-12)
        # Priced window passes only add time.
        H.set_calibration(H.Calibration(**dict(profile, pass_bytes_per_second=1.0e11, launch_seconds=1e-5)))
        assert ENGINE.estimate(c, tier_window_qubits=16).seconds > want - reads * sb / 5.0e10
    finally:
        H.set_calibration(old)
