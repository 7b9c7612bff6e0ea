"""complex64 blocks of 5 and 6 qubits on the matrix cores (v_mfma_f32_32x32x2_f32)
against the FP64 oracle and against the one-thread-per-group kernel at
complex64: every operand order, alone and inside block lists beside window
passes; errors within 1e-5 of the complex128 oracle (the complex64 bar)."""
import numpy as np
import pytest

from circuits import random_state, random_unitary

pytestmark = pytest.mark.gpu

TOL = 1e-5


@pytest.mark.parametrize("k", [5, 6])
def test_matrix_core_block_against_the_oracle(orc, hip, k):
    rng = np.random.default_rng(70 + k)
    for n in (k + 5, 14, 17):
        s = random_state(n, 80 + n)
        for _ in range(3):
            ops = rng.permutation(n)[:k].tolist()
            m = random_unitary(k, rng)
            d = hip.DeviceState(n, hip.C64)
            d.upload(s.astype(np.complex64))
            d.apply_matrix(m, ops)
            got = d.download()
            d.upload(s.astype(np.complex64))
            d.apply_matrix(m, ops, generic=True)
            valu = d.download()
            d.free()
            want = s.copy()
            orc.apply_generic(want, m, ops)
            assert np.max(np.abs(got - want)) <= TOL, (n, ops)
            assert np.max(np.abs(got - valu)) <= TOL, (n, ops)


def test_matrix_core_blocks_in_a_block_list(orc, hip):
    rng = np.random.default_rng(91)
    n = 15
    s = random_state(n, 92)
    blocks = []
    for i in range(24):
        k = [1, 2, 5, 6][i % 4]
        blocks.append((random_unitary(k, rng), rng.permutation(n)[:k].tolist()))
    d = hip.DeviceState(n, hip.C64)
    d.upload(s.astype(np.complex64))
    d.apply_blocks(blocks)
    got = d.download()
    d.free()
    want = s.copy()
    for m, ops in blocks:
        orc.apply_generic(want, m, ops)
    assert np.max(np.abs(got - want)) <= TOL


def test_autotuned_plan_runs_the_plan_it_reports(orc, hip):
    """The engine with MH_FUSE_AUTO on a deep 5-qubit circuit of a 16-qubit
    state: it reports width-5 fusion, complex128 is bitwise the oracle on that
    plan, complex64 (the block on the matrix cores) within 1e-5 of it."""
    import oracle
    from test_time_model import deep_circuit
    c = deep_circuit(16, 10, 5, 30, 5)
    eng = hip.HipDenseEngine()
    r128 = eng.execute(c, fusion=hip.FUSE_AUTO, statevector=True)
    assert (r128.fusion_used, r128.fusion_width_used) == (hip.FUSE_WIDTH, 5)
    order, bb, bq, qb, fb, ga = hip.plan_fusion(c, hip.FUSE_WIDTH, 5)
    want, _, _ = orc.execute(c, plan=oracle.Plan(order, bb, bq, qb, fb, ga))
    assert r128.statevector.tobytes() == want.tobytes()
    r64 = eng.execute(c, fusion=hip.FUSE_AUTO, precision=hip.C64, statevector=True)
    assert r64.fusion_used == hip.FUSE_WIDTH and r64.num_passes == 1
    assert np.max(np.abs(r64.statevector - want)) <= TOL
