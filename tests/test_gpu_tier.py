"""The host-offloaded tier on the device (mh_tier_execute): the state in pinned
host memory, the passes through two device windows. Its statevector and counts
are the in-memory engine's bit for bit at complex128 (the engine's are the
reference's: the golden and oracle tests); its raw terms agree with the
engine's and the oracle's within 1e-12 (FP64 sums in another order), for
terms inside a window, Z masks above it, and X masks reaching above it."""
import numpy as np
import pytest

from menos_amd import hip as H
from menos_amd import workloads as W

pytestmark = pytest.mark.gpu

TIER = H.HipHostTierEngine()
DENSE = H.HipDenseEngine()


def _random(n, layers, seed):
    c = W.random_circuit(n, layers, seed)
    c.ops = [op for op in c.ops if op[0] != W.MEASURE]
    return c


def _terms(n, window, seed):
    """Terms of every kind the tier reads: X inside the window, Z above it, X
    reaching above it (several values of x >> (window - 1)), Y pairs above it
    (the i^popcount(x & z) phase split between the window and the pair)."""
    rng = np.random.default_rng(seed)
    hi = list(range(window, n))
    terms = [(0, 0), (0b11, 0b10), (0, 1 << (n - 1)), (0b101, (1 << (n - 1)) | 1)]
    for k in (1, 2, 3):
        qs = rng.choice(hi, size=min(k, len(hi)), replace=False)
        x = sum(1 << int(q) for q in qs) | int(rng.integers(0, 1 << window))
        z = int(rng.integers(0, 1 << n))
        terms += [(x, z), (x, x), (x, x | 1), (x, 0)]
    terms.append(((1 << (window - 1)) | (1 << window), 1 << window))
    return terms


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint64 if a.dtype == np.complex128 else np.uint32)


@pytest.mark.parametrize("n,layers,chunk,window,seed", [
    (16, 4, 12, 14, 1),   # chunk and window given
    (18, 3, 0, 12, 2),    # chunk chosen
    (17, 5, 10, 12, 3),
    (20, 3, 0, 14, 4),
])
def test_statevector_counts_and_terms_equal_the_in_memory_engine(hip, orc, n, layers, chunk, window, seed):
    c = _random(n, layers, seed)
    terms = _terms(n, window, seed)
    targets = [(q, q) for q in range(n)]
    kw = dict(shots=20000, seed=seed + 100, targets=targets, width=n, terms=terms)
    got = TIER.execute(c, tier_chunk_qubits=chunk, tier_window_qubits=window, record_pass_times=True, **kw)
    want = DENSE.execute(c, **kw)
    assert got.cost.window_qubits == window
    assert got.num_passes == got.cost.passes > 1 and len(got.pass_ms) == got.num_passes
    assert np.array_equal(_bits(got.statevector), _bits(want.statevector))
    assert got.counts == want.counts
    assert np.max(np.abs(got.expectation - want.expectation)) <= 1e-12
    sv, _, ex = orc.execute(c, terms=terms)
    assert np.array_equal(_bits(got.statevector), _bits(sv))
    assert np.max(np.abs(got.expectation - ex)) <= 1e-12


def test_wide_blocks_across_global_qubits(hip):
    # Width-4 fusion: blocks of 3-4 qubits, many with operands above the chunk,
    # on the window kernels and the per-group kernel inside each window.
    c = _random(17, 4, 7)
    for fusion, width in ((H.FUSE_WIDTH, 4), (H.FUSE_CONTAINMENT, 0)):
        got = TIER.execute(c, tier_chunk_qubits=11, tier_window_qubits=15, fusion=fusion, fusion_width=width)
        want = DENSE.execute(c, fusion=fusion, fusion_width=width)
        assert got.cost.global_blocks > 0
        assert np.array_equal(_bits(got.statevector), _bits(want.statevector))


def test_autotuned_fusion(hip):
    # The autotuner's plan (containment or a width of 2-6, priced by the time
    # model) is the plan the tier runs, the same one the in-memory engine picks.
    c = _random(16, 5, 12)
    got = TIER.execute(c, tier_chunk_qubits=10, tier_window_qubits=13, fusion=H.FUSE_AUTO)
    want = DENSE.execute(c, fusion=H.FUSE_AUTO)
    assert (got.fusion_used, got.fusion_width_used) == (want.fusion_used, want.fusion_width_used)
    assert np.array_equal(_bits(got.statevector), _bits(want.statevector))


def test_complex64(hip, orc):
    c = _random(18, 4, 11)
    terms = _terms(18, 13, 11)
    kw = dict(precision=H.C64, terms=terms, shots=20000, seed=5, targets=[(q, q) for q in range(18)], width=18)
    got = TIER.execute(c, tier_window_qubits=13, **kw)
    want = DENSE.execute(c, **kw)
    assert got.counts == want.counts
    sv, _, ex = orc.execute(c, terms=terms)
    assert np.max(np.abs(got.statevector.astype(np.complex128) - sv)) <= 1e-5
    assert np.max(np.abs(got.statevector - want.statevector)) <= 1e-5
    assert np.max(np.abs(got.expectation - ex)) <= 1e-5
    # Each block's complex64 arithmetic is the same wherever its operands sit
    # in a window, so the state is the in-memory engine's bit for bit here too.
    assert np.array_equal(_bits(got.statevector), _bits(want.statevector))


def test_budget_derived_window_and_ledger(hip):
    c = _random(19, 3, 5)
    need = TIER.estimate(c, tier_window_qubits=13, shots=1000, terms=[(1, 3)])
    got = TIER.execute(c, budget_bytes=need.device_bytes, shots=1000, seed=9, terms=[(1, 3)],
                       targets=[(q, q) for q in range(19)], width=19)
    assert got.cost.window_qubits == 13
    assert got.peak_device_bytes <= need.device_bytes
    want = DENSE.execute(c, shots=1000, seed=9, terms=[(1, 3)], targets=[(q, q) for q in range(19)], width=19)
    assert got.counts == want.counts
    with pytest.raises(H.ResourceExceeded):
        TIER.execute(c, budget_bytes=1 << 10)


def test_a_circuit_without_gates_writes_the_initial_state(hip):
    c = W.Circuit(13, global_phase=0.3)
    got = TIER.execute(c, tier_window_qubits=11, shots=50, seed=1, targets=[(0, 0)], width=1)
    want = DENSE.execute(c, shots=50, seed=1, targets=[(0, 0)], width=1)
    assert got.num_passes == 0
    assert np.array_equal(_bits(got.statevector), _bits(want.statevector))
    assert got.counts == want.counts == {0: 50}


def test_the_monitor_abandons_at_a_pass_boundary(hip):
    c = _random(16, 6, 8)
    seen = []

    class Stop(Exception):
        pass

    def monitor(done):
        seen.append(done)
        if done == 2:
            raise Stop()

    cost = TIER.estimate(c, tier_chunk_qubits=10, tier_window_qubits=12)
    assert cost.passes > 2
    with pytest.raises(Stop):
        TIER.execute(c, tier_chunk_qubits=10, tier_window_qubits=12, monitor=monitor)
    assert seen == [1, 2]
    # The library is usable afterwards, and an uninterrupted run sees every pass.
    seen.clear()
    got = TIER.execute(c, tier_chunk_qubits=10, tier_window_qubits=12, monitor=seen.append)
    assert seen == list(range(1, cost.passes + 1))
    assert np.array_equal(_bits(got.statevector), _bits(DENSE.execute(c).statevector))


@pytest.mark.parametrize("name", ["uniform", "one_window", "sparse_windows"])
def test_counts_across_windows_of_adversarial_states(hip, name):
    """States whose running sums stress the carried window boundaries: the
    uniform superposition (every probability equal, exact ties at every
    boundary), all weight in one window of 32 (the others sum to exactly
    zero), and weight in a run of eight windows in the upper half only."""
    n, window = 16, 11
    c = W.Circuit(n)
    if name == "uniform":
        for q in range(n):
            c.append(W.H, (q,))
    elif name == "one_window":
        for q in range(window):
            c.append(W.H, (q,))
        c.append(W.X, (n - 2,))
    else:
        for q in range(window):
            c.append(W.H, (q,))
        for q in range(window, n - 2):
            c.append(W.H, (q,))
        c.append(W.X, (n - 1,))
    targets = [(q, q) for q in range(n)]
    for seed in (3, 4):
        kw = dict(shots=30000, seed=seed, targets=targets, width=n)
        got = TIER.execute(c, tier_window_qubits=window, statevector=False, **kw)
        want = DENSE.execute(c, statevector=False, **kw)
        assert got.counts == want.counts

