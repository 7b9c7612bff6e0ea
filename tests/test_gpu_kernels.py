"""Device primitives against the oracle: gate application bit for bit,
expectation terms within 1e-12, sampled indices identical."""
import numpy as np
import pytest

from circuits import random_state, random_unitary
from menos_amd import workloads as W

pytestmark = pytest.mark.gpu


def _apply_both(orc, hip, state, matrix, ops, generic=False, precision=0):
    n = int(np.log2(state.size))
    d = hip.DeviceState(n, precision)
    d.upload(state)
    d.apply_matrix(matrix, ops, generic=generic)
    got = d.download()
    d.free()
    want = state.copy()
    orc.apply_generic(want, matrix, ops)
    return got, want


def test_every_one_qubit_operand(orc, hip):
    """Below 10 qubits the generic kernel; from 10 on the wavefront windows
    (every operand on each lane bit and register bit of the layout)."""
    rng = np.random.default_rng(2421)
    for n in range(1, 14):
        s = random_state(n, n)
        for q in range(n):
            m = random_unitary(1, rng)
            got, want = _apply_both(orc, hip, s, m, [q])
            assert got.tobytes() == want.tobytes(), (n, q)


def test_every_ordered_two_qubit_pair(orc, hip):
    rng = np.random.default_rng(2422)
    for n in (2, 3, 4, 5, 6, 7, 8, 10, 11):
        s = random_state(n, 100 + n)
        for a in range(n):
            for b in range(n):
                if a == b:
                    continue
                m = random_unitary(2, rng)
                got, want = _apply_both(orc, hip, s, m, [a, b])
                assert got.tobytes() == want.tobytes(), (n, a, b)


@pytest.mark.parametrize("k", [3, 4, 5, 6])
def test_wide_blocks_window_and_generic(orc, hip, k):
    rng = np.random.default_rng(2423 + k)
    for trial in range(4):
        n = int(rng.integers(k, 17))
        ops = rng.permutation(n)[:k].tolist()
        m = random_unitary(k, rng)
        s = random_state(n, trial + 10 * k)
        got, want = _apply_both(orc, hip, s, m, ops)
        assert got.tobytes() == want.tobytes(), (k, n, ops)
        got, want = _apply_both(orc, hip, s, m, ops, generic=True)
        assert got.tobytes() == want.tobytes(), ("generic", k, n, ops)


def test_diagonal_blocks(orc, hip):
    rng = np.random.default_rng(7)
    for k in range(1, 6):
        for trial in range(3):
            n = int(rng.integers(k, 15))
            ops = rng.permutation(n)[:k].tolist()
            m = np.diag(np.exp(1j * rng.uniform(-np.pi, np.pi, 1 << k)))
            s = random_state(n, trial)
            got, want = _apply_both(orc, hip, s, m, ops)
            assert got.tobytes() == want.tobytes(), (k, n, ops)


@pytest.mark.parametrize("n", [6, 11])
def test_signed_zero_matrices_and_states(orc, hip, n):
    """-0 entries and exact zeros in the state: outputs start at +0 (n = 11
    runs the window passes, which compute up to the sign of zeros and add +0
    at the end)."""
    s = random_state(n, 3)
    s[::3] = -0.0 - 0.0j
    s[1::5] = complex(0.0, -0.0)
    for m in (np.array([[-0.0, 1.0], [1.0, -0.0]], np.complex128),
              np.array([[complex(-0.0, -0.0), 1j], [-1j, complex(0.0, -0.0)]])):
        for q in range(n):
            got, want = _apply_both(orc, hip, s, m, [q])
            assert got.tobytes() == want.tobytes(), q
    rz = np.diag([complex(1.0, -0.0), complex(-1.0, 0.0)])
    got, want = _apply_both(orc, hip, s, rz, [2])
    assert got.tobytes() == want.tobytes()


def test_runs_of_sign_diagonals(orc, hip):
    """Consecutive sign diagonals (entries +-1) of 1-4 qubits on lane, register
    and outside-window qubits, mixed with 1-qubit and phase diagonals: a pass
    applies each run as one XOR mask."""
    rng = np.random.default_rng(11)
    for n in (10, 12, 13):
        for trial in range(4):
            blocks = []
            for _ in range(40):
                r = rng.random()
                k = int(rng.integers(1, 5))
                ops = rng.permutation(n)[:k].tolist()
                if r < 0.6:
                    m = np.diag(rng.choice([1.0, -1.0], 1 << k).astype(np.complex128))
                elif r < 0.8:
                    ops = ops[:1]
                    m = random_unitary(1, rng)
                else:
                    m = np.diag(np.exp(1j * rng.uniform(-np.pi, np.pi, 1 << k)))
                blocks.append((m, ops))
            s = random_state(n, trial)
            d = hip.DeviceState(n)
            d.upload(s)
            d.apply_blocks(blocks)
            got = d.download()
            d.free()
            want = s.copy()
            for m, ops in blocks:
                orc.apply_generic(want, m, ops)
            assert got.tobytes() == want.tobytes(), (n, trial)


def _controlled(rng, u, cnot):
    """A 2-qubit block that never mixes operand u's values: a random 2x2 (or
    the identity then X, a controlled NOT) on the other operand per value of
    operand u, with signed-zero entries elsewhere."""
    m = np.zeros((4, 4), np.complex128)
    for cb in range(2):
        blk = (np.eye(2) if cb == 0 else np.array([[0, 1], [1, 0]])) if cnot else random_unitary(1, rng)
        for to in range(2):
            for ti in range(2):
                m[(cb << u) | (to << (1 - u)), (cb << u) | (ti << (1 - u))] = blk[to, ti]
    if not cnot:
        m[m == 0] = -0.0
    return m


def test_controlled_pairs(orc, hip):
    """Controlled pairs (2-qubit blocks that are a 2x2 block on one operand
    chosen by the other's bit, controlled NOTs among them) whose control is a
    register bit, a lane bit or a qubit outside the window, between dense
    1- and 2-qubit blocks: bit for bit the generic kernel's result."""
    rng = np.random.default_rng(23)
    for n in (10, 12, 14):
        for trial in range(3):
            blocks = []
            for _ in range(50):
                r = rng.random()
                ops = rng.permutation(n)[:2].tolist()
                if r < 0.5:
                    blocks.append((_controlled(rng, int(rng.integers(0, 2)), rng.random() < 0.4), ops))
                elif r < 0.8:
                    blocks.append((random_unitary(1, rng), ops[:1]))
                else:
                    blocks.append((random_unitary(2, rng), ops))
            s = random_state(n, 40 + trial)
            d = hip.DeviceState(n)
            d.upload(s)
            d.apply_blocks(blocks)
            got = d.download()
            d.free()
            want = s.copy()
            for m, ops in blocks:
                orc.apply_generic(want, m, ops)
            assert got.tobytes() == want.tobytes(), (n, trial)


def test_block_list_equals_gate_by_gate(orc, hip):
    """Many blocks per LDS-window pass give the bits of one launch per block."""
    rng = np.random.default_rng(99)
    for n in (5, 12, 16, 20):
        blocks = []
        for _ in range(40):
            k = int(rng.choice([1, 1, 1, 2, 2]))
            ops = rng.permutation(n)[:k].tolist()
            m = random_unitary(k, rng)
            if rng.random() < 0.25:
                m = np.diag(np.diag(m))
            blocks.append((m, ops))
        s = random_state(n, n)
        want = s.copy()
        for m, ops in blocks:
            orc.apply_generic(want, m, ops, threads=8)
        d = hip.DeviceState(n)
        d.upload(s)
        passes = d.apply_blocks(blocks)
        got = d.download()
        d.free()
        assert got.tobytes() == want.tobytes(), n
        assert 1 <= passes <= len(blocks)
        if n >= 16:
            d = hip.DeviceState(n)
            d.upload(s)
            d.apply_blocks(blocks, window_qubits=9)
            assert d.download().tobytes() == want.tobytes()
            d.free()


def test_init(hip):
    d = hip.DeviceState(10)
    d.init(np.exp(1j * 0.7))
    s = d.download()
    assert s[0] == np.exp(1j * 0.7) and not np.any(s[1:])
    d.free()


def test_expectation_terms(orc, hip):
    for n in (1, 3, 9, 16):
        s = random_state(n, 40 + n)
        terms = W.random_paulis(n, 24, n, max_weight=min(6, n)) + W.z_terms(n)
        d = hip.DeviceState(n)
        d.upload(s)
        got = d.expectation(terms)
        d.free()
        for (x, z), g in zip(terms, got):
            want = orc.expectation_of_term(s, x, z)
            assert abs(g - want) <= 1e-12, (n, x, z, g, want)


def test_expectation_tiles_and_wide_masks(orc, hip):
    """Terms from LDS tiles (sets of up to 16 terms whose X bits fit an
    11-qubit tile), sets split by their X bits and by their size, X masks too
    wide for a tile (the per-mask kernel), identity and all-qubit strings, at
    both precisions: within 1e-12 (complex64: 1e-5) of the sequential sum, and
    the same bits alone and in a batch."""
    rng = np.random.default_rng(77)
    for n in (4, 11, 12, 15):
        s = random_state(n, 60 + n)
        terms = W.random_paulis(n, 40, n + 5, max_weight=min(4, n))
        terms += W.random_paulis(n, 6, n + 6, max_weight=n, min_weight=min(n, 9))   # wide X masks
        terms += [(0, 0), ((1 << n) - 1, 0), ((1 << n) - 1, (1 << n) - 1), (0, (1 << n) - 1)]
        terms += [(int(x), int(rng.integers(0, 1 << n))) for x in rng.integers(0, 1 << n, 4)]
        for prec, tol in ((hip.C128, 1e-12), (hip.C64, 1e-5)):
            d = hip.DeviceState(n, prec)
            d.upload(s.astype(np.complex64) if prec == hip.C64 else s)
            got = d.expectation(terms)
            d.free()
            for (x, z), g in zip(terms, got):
                want = orc.expectation_of_term(s, x, z)
                assert abs(g - want) <= tol, (n, prec, x, z, g, want)


def test_expectation_is_reproducible(hip):
    s = random_state(18, 5)
    terms = W.random_paulis(18, 8, 3)
    d = hip.DeviceState(18)
    d.upload(s)
    a, b = d.expectation(terms), d.expectation(terms)
    d.free()
    assert a.tobytes() == b.tobytes()


def _uniforms(orc, shots, seed):
    return np.sort(np.array([W.word_to_uniform(int(w)) for w in orc.mt19937_64(seed, shots)]))


@pytest.mark.parametrize("n", [1, 4, 12, 13, 14, 20, 22])
def test_sampler_random_states(orc, hip, n):
    s = random_state(n, 500 + n)
    d = hip.DeviceState(n)
    d.upload(s)
    for shots, seed in ((1, 3), (1024, 2025), (5000, 7)):
        got = d.sample(_uniforms(orc, shots, seed))
        want = orc.sample_indices(s, shots, seed)
        assert got.tolist() == want.tolist(), (n, shots, seed)
    d.free()


@pytest.mark.parametrize("case", ["basis", "zeros", "powers_of_two", "ties", "tiny_then_large", "unnormalised"])
def test_sampler_adversarial(orc, hip, case):
    n = 15
    rng = np.random.default_rng(17)
    size = 1 << n
    if case == "basis":
        s = np.zeros(size, np.complex128)
        s[size - 3] = 1.0
    elif case == "zeros":
        s = random_state(n, 1) * (rng.random(size) < 0.05)
    elif case == "powers_of_two":
        s = np.sqrt(2.0 ** -rng.integers(14, 40, size).astype(np.float64)).astype(np.complex128)
    elif case == "ties":
        s = np.full(size, 2.0 ** -7.5, np.complex128)
        s[::7] *= 1 + 2.0 ** -40
    elif case == "tiny_then_large":
        s = random_state(n, 2) * 1e-160
        s[size // 3:] = random_state(n, 3)[size // 3:]
    else:
        s = random_state(n, 4) * 3.0
    d = hip.DeviceState(n)
    d.upload(s)
    got = d.sample(_uniforms(orc, 4096, 11))
    want = orc.sample_indices(s, 4096, 11)
    assert got.tolist() == want.tolist(), case
    u = np.array([W.word_to_uniform(int(w)) for w in orc.mt19937_64(5, 3000)])
    got = d.sample_unsorted(u)
    cum = orc.cumulative(s)
    want = np.minimum(np.searchsorted(cum, u, side="right"), size - 1)
    assert got.tolist() == want.tolist(), case
    d.free()


def test_complex64_apply_and_error(orc, hip):
    """complex64 state, FP32 gates: within single-precision error of the c128 oracle."""
    rng = np.random.default_rng(64)
    n = 14
    s = random_state(n, 64)
    blocks = [(random_unitary(int(k), rng), rng.permutation(n)[:int(k)].tolist()) for k in rng.integers(1, 3, 30)]
    want = s.copy()
    for m, ops in blocks:
        orc.apply_generic(want, m, ops)
    d = hip.DeviceState(n, hip.C64)
    d.upload(s.astype(np.complex64))
    d.apply_blocks(blocks)
    got = d.download().astype(np.complex128)
    assert np.max(np.abs(got - want)) < 2e-6
    ez = d.expectation(W.z_terms(n))
    for q, e in enumerate(ez):
        assert abs(e - orc.expectation_of_term(want, 0, 1 << q)) < 1e-5
    d.free()


def test_pool_reuse(hip):
    import ctypes
    a = hip.DeviceState(20)
    pa = a.ptr.value
    a.free()
    b = hip.DeviceState(20)
    assert b.ptr.value == pa  # the freed block is handed back
    b.free()
    r = ctypes.c_uint64()
    hip.check(hip.lib().mh_pool_reusable_bytes(ctypes.byref(r)))
    assert r.value >= 16 << 20
    hip.check(hip.lib().mh_pool_trim())
    hip.check(hip.lib().mh_pool_reusable_bytes(ctypes.byref(r)))
    assert r.value == 0


def test_invalid_operands_are_refused(hip):
    d = hip.DeviceState(4)
    with pytest.raises(hip.HipError, match="out of range"):
        d.apply_matrix(np.eye(2), [4])
    with pytest.raises(hip.HipError, match="repeated"):
        d.apply_matrix(np.eye(4), [1, 1])
    d.free()


def test_state_max_abs_diff(hip):
    """mh_state_max_abs_diff: zero against itself, the planted difference
    otherwise, across precisions."""
    n = 12
    s = random_state(n, 4)
    a = hip.DeviceState(n)
    a.upload(s)
    assert a.max_abs_diff(a) == 0.0
    t = s.copy()
    t[777] += 3e-3 + 4e-3j
    b = hip.DeviceState(n)
    b.upload(t)
    assert abs(a.max_abs_diff(b) - abs(t[777] - s[777])) < 1e-15
    c = hip.DeviceState(n, hip.C64)
    c.upload(s.astype(np.complex64))
    want = float(np.max(np.abs(s - s.astype(np.complex64).astype(np.complex128))))
    assert abs(a.max_abs_diff(c) - want) < 1e-15
    for d in (a, b, c):
        d.free()
