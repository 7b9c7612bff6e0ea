"""The engine end to end against the oracle and the reference's digests."""
import numpy as np
import pytest

import oracle
from circuits import diagonal_heavy_circuit, random_mixed_circuit
from conftest import circuit_from_spec, cpx
from menos_amd import workloads as W

pytestmark = pytest.mark.gpu


def test_reference_memory_fixture_digests(orc, hip, golden):
    """layered(18, 3): the reference's committed statevector / counts digests."""
    base = golden["reference_baseline"]
    eng = hip.HipDenseEngine()
    c = W.layered(18, 3).measure_all()
    for fusion, sv_key, cnt_key in ((hip.FUSE_NONE, "dense_statevector", "dense_counts"),
                                    (hip.FUSE_CONTAINMENT, "dense_statevector_fused", "dense_counts_fused")):
        r = eng.execute(c, fusion=fusion, shots=20000, seed=2025)
        assert "%016x" % orc.fnv1a(r.statevector) == base[sv_key]["statevector"]
        assert W.digest_counts(r.counts) == base[cnt_key]["counts"]


def test_reference_expectation_within_tolerance(orc, hip):
    terms = [W.pauli_masks({0: "Z"}), W.pauli_masks({0: "X", 1: "X"}), W.pauli_masks({0: "Z", 17: "Y"})]
    eng = hip.HipDenseEngine()
    r = eng.execute(W.layered(18, 3), terms=terms, statevector=False)
    _, _, want = orc.execute(W.layered(18, 3), terms=terms, statevector=False)
    assert np.max(np.abs(r.expectation - want)) <= 1e-12


def test_golden_engine_runs(hip, golden, golden_states):
    eng = hip.HipDenseEngine()
    for run in golden["engine_runs"]:
        c = circuit_from_spec(run["circuit"])
        n = c.num_qubits
        r = eng.execute(c, fusion=hip.FUSE_CONTAINMENT if run["fuse_gates"] else hip.FUSE_NONE, shots=run["shots"],
                        seed=run["seed"], targets=[(q, q) for q in range(n)], width=n,
                        terms=[tuple(t) for t in run["terms"]])
        assert r.statevector.tobytes() == golden_states[run["key"]].tobytes(), run["key"]
        assert {str(k): v for k, v in r.counts.items()} == run["counts"], run["key"]
        assert np.max(np.abs(r.expectation - cpx(run["expectation"]))) <= 1e-12, run["key"]


def test_random_circuit_20q_20_layers(orc, hip, golden):
    """The 20-qubit random circuit: statevector and 1024-shot counts bit-exact,
    <Z_i> within 1e-12 of the sequential sums."""
    g = golden["random_circuit_20q"]
    c = W.random_circuit(20, 20, g["circuit_seed"])
    eng = hip.HipDenseEngine()
    r = eng.execute(c, shots=1024, seed=g["sample_seed"], targets=[(q, q) for q in range(20)], width=20,
                    terms=W.z_terms(20))
    assert "%016x" % orc.fnv1a(r.statevector) == g["statevector_digest"]
    assert W.digest_counts(r.counts) == g["counts_digest"]
    assert np.max(np.abs(r.expectation - cpx(g["z_expectation"]))) <= 1e-12


@pytest.mark.parametrize("width", [2, 3, 4, 5, 6])
def test_width_fusion_matches_oracle_on_same_plan(orc, hip, width):
    c = W.random_circuit(14, 6, 40 + width)
    order, bb, bq, qb, _, _ = hip.plan_fusion(c, hip.FUSE_WIDTH, width)
    plan = oracle.Plan(order, bb, bq, qb)
    want, idx, e = orc.execute(c, plan=plan, shots=2048, seed=5, terms=W.random_paulis(14, 10, width))
    r = hip.HipDenseEngine().execute(c, fusion=hip.FUSE_WIDTH, fusion_width=width, shots=2048, seed=5,
                                     targets=[(q, q) for q in range(14)], width=14,
                                     terms=W.random_paulis(14, 10, width))
    assert r.statevector.tobytes() == want.tobytes()
    assert r.counts == oracle.counts_from_indices(idx, [(q, q) for q in range(14)])
    assert np.max(np.abs(r.expectation - e)) <= 1e-12


@pytest.mark.parametrize("kind", ["ansatz", "layered", "random"])
def test_pair_fusion_matches_oracle_on_same_plan(orc, hip, kind):
    """Containment fusion that keeps controlled-NOT blocks as controlled pairs:
    complex128 bit for bit the oracle applying the same block list, complex64
    within 1e-5; an ansatz fuses into one 1-qubit block and n - 1 pairs per layer."""
    n = 14
    if kind == "ansatz":
        c = W.ansatz_circuit(n, 4, np.random.default_rng(3).uniform(-np.pi, np.pi, 2 * n * 4).tolist())
    elif kind == "layered":
        c = W.layered(n, 5)
    else:
        c = W.random_circuit(n, 4, 9)
    order, bb, bq, qb, _, _ = hip.plan_fusion(c, hip.FUSE_PAIRS)
    plan = oracle.Plan(order, bb, bq, qb)
    terms = W.random_paulis(n, 12, 4)
    want, _, e = orc.execute(c, plan=plan, terms=terms)
    r = hip.HipDenseEngine().execute(c, fusion=hip.FUSE_PAIRS, terms=terms, targets=[], width=0)
    assert r.statevector.tobytes() == want.tobytes()
    assert np.max(np.abs(r.expectation - e)) <= 1e-12
    r64 = hip.HipDenseEngine().execute(c, precision=hip.C64, fusion=hip.FUSE_PAIRS, terms=terms, targets=[], width=0)
    assert np.max(np.abs(r64.statevector.astype(np.complex128) - want)) <= 1e-5
    if kind == "ansatz":
        assert plan.num_blocks == 4 * n


@pytest.mark.parametrize("seed", range(4))
def test_mixed_circuits(orc, hip, seed):
    for c in (random_mixed_circuit(11, 90, 1000 + seed, measures=True), diagonal_heavy_circuit(13, 120, 1100 + seed)):
        for fusion in (0, 1):
            want, idx, _ = orc.execute(c, fuse_gates=bool(fusion), shots=500, seed=seed)
            r = hip.HipDenseEngine().execute(c, fusion=fusion, shots=500, seed=seed,
                                             targets=[(q, q) for q in range(c.num_qubits)], width=c.num_qubits)
            assert r.statevector.tobytes() == want.tobytes()
            assert r.counts == oracle.counts_from_indices(idx, [(q, q) for q in range(c.num_qubits)])


def test_counts_follow_measure_targets(orc, hip):
    c = W.random_circuit(6, 3, 9)
    targets = [(5, 0), (0, 1), (3, 2)]
    want_sv, idx, _ = orc.execute(c, shots=3000, seed=1)
    r = hip.HipDenseEngine().execute(c, shots=3000, seed=1, targets=targets, width=3)
    assert r.counts == oracle.counts_from_indices(idx, targets)
    assert sum(r.counts.values()) == 3000


def test_complex64_engine_reports_error(orc, hip):
    c = W.random_circuit(16, 8, 3)
    want, _, e = orc.execute(c, terms=W.zz_neighbour_terms(16))
    r = hip.HipDenseEngine().execute(c, precision=hip.C64, terms=W.zz_neighbour_terms(16))
    amp_err = np.max(np.abs(r.statevector.astype(np.complex128) - want))
    exp_err = np.max(np.abs(r.expectation - e))
    assert amp_err < 1e-5 and exp_err < 1e-5, (amp_err, exp_err)


def test_budget_refused_before_allocation(hip):
    c = W.random_circuit(20, 1, 1)
    eng = hip.HipDenseEngine()
    with pytest.raises(hip.ResourceExceeded, match="budget"):
        eng.execute(c, budget_bytes=1 << 20)
    est = eng.estimate(c, budget_bytes=1 << 20)
    assert est.feasible == 0 and b"budget" in est.infeasible_reason
    est = eng.estimate(c)
    assert est.feasible == 1 and est.device_bytes >= 16 << 20 and est.passes >= 1


def test_peak_device_bytes_reported(hip):
    r = hip.HipDenseEngine().execute(W.random_circuit(18, 2, 1), shots=10, seed=1, targets=[(0, 0)], width=1)
    assert r.actual_bytes == 16 << 18
    assert r.peak_device_bytes >= r.actual_bytes


def test_monitor_sees_every_pass_and_changes_nothing(orc, hip):
    """The safe-boundary monitor runs once per pass, counting from 1, and a
    monitored run is bitwise the unmonitored one."""
    c = W.random_circuit(18, 8, 77)
    seen = []
    eng = hip.HipDenseEngine()
    r = eng.execute(c, monitor=seen.append, targets=[], width=0)
    plain = eng.execute(c, targets=[], width=0)
    assert seen == list(range(1, r.num_passes + 1))
    assert r.statevector.tobytes() == plain.statevector.tobytes()
    want, _, _ = orc.execute(c)
    assert r.statevector.tobytes() == want.tobytes()


def test_monitor_abandon_frees_the_state(hip):
    """A monitor that raises at its third boundary abandons the run: the call
    re-raises it, the state goes back to the pool (available device bytes as
    before) and the next run is unaffected (reference
    tests/api/cuda_execution_monitor_test.cpp)."""
    class Abandon(Exception):
        pass

    def stop_at_three(crossed):
        if crossed == 3:
            raise Abandon(crossed)

    c = W.random_circuit(24, 8, 5)                 # 256 MiB of state
    eng = hip.HipDenseEngine()
    eng.execute(c, statevector=False, targets=[], width=0)   # the pool holds a state of this size
    before = hip.available_bytes()
    with pytest.raises(Abandon):
        eng.execute(c, monitor=stop_at_three, statevector=False, targets=[], width=0)
    assert hip.available_bytes() == before
    small = W.random_circuit(12, 4, 6)
    assert eng.execute(small, targets=[], width=0).statevector.tobytes() == \
        eng.execute(small, targets=[], width=0).statevector.tobytes()


def test_calibration_and_time_model(hip):
    """mh_calibrate measures finite, positive rates of the right magnitude and
    installs them; the time model then predicts a 26-qubit run's device work
    within a factor of two of the measured time."""
    cal = hip.calibrate()
    for f in ("hbm_bytes_per_second", "pass_bytes_per_second", "h2d_bytes_per_second", "d2h_bytes_per_second",
              "fp64_ops_per_second"):
        v = getattr(cal, f)
        assert np.isfinite(v) and v > 0, f
    assert 1e12 < cal.hbm_bytes_per_second < 1.2e13
    assert 1e12 < cal.pass_bytes_per_second < 1.2e13
    assert 0 < cal.launch_seconds < 1e-3
    got = hip.get_calibration()
    assert got.hbm_bytes_per_second == cal.hbm_bytes_per_second
    c = W.random_circuit(26, 8, 11)
    eng = hip.HipDenseEngine()
    est = eng.estimate(c)
    eng.execute(c, statevector=False, targets=[], width=0)
    r = eng.execute(c, statevector=False, targets=[], width=0)
    assert est.passes == r.num_passes
    assert 0.5 < est.seconds / r.execute_seconds < 2.0, (est.seconds, r.execute_seconds)
