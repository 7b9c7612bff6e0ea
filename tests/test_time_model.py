"""The engine's calibrated time model on the CPU (mh_engine_estimate needs no
device): the prediction is the stated formula over the estimate's own counts
under a synthetic profile, it grows with the qubit count and with the passes,
an infeasible request is reported and never raised, and a profile with a
non-positive rate is refused. In the manner of the reference's synthetic-profile
backend-selection tests (tests/plan/cuda_backend_selection_test.cpp:81-651)."""
import math

import pytest

from menos_amd import hip as H
from menos_amd import workloads as W

PROFILE = dict(hbm_bytes_per_second=6.0e12, pass_bytes_per_second=5.0e12, launch_seconds=1.0e-5,
               h2d_bytes_per_second=5.0e10, d2h_bytes_per_second=4.0e10, fp64_ops_per_second=3.0e13,
               mfma_flops_per_second=1.0e14)


@pytest.fixture
def profile():
    old = H.get_calibration()
    H.set_calibration(H.Calibration(**PROFILE))
    yield PROFILE
    H.set_calibration(old)


def _formula(e, n, precision, shots, terms):
    sb = (1 << n) * (16 if precision == H.C128 else 8)
    p = PROFILE
    s = e.launches * p["launch_seconds"]
    s += e.passes * 2.0 * sb / p["pass_bytes_per_second"]
    s += sb / p["hbm_bytes_per_second"]
    s += (3.0 if shots else 0.0) * sb / p["hbm_bytes_per_second"]
    s += terms * sb / p["hbm_bytes_per_second"]
    s += e.h2d_bytes / p["h2d_bytes_per_second"]
    s += e.d2h_bytes / p["d2h_bytes_per_second"]
    return s


def test_prediction_is_the_formula(profile):
    eng = H.HipDenseEngine()
    for n, layers, shots, nterms, prec in ((12, 3, 0, 0, H.C128), (20, 6, 1024, 4, H.C128), (24, 4, 0, 8, H.C64)):
        c = W.random_circuit(n, layers, 7)
        terms = W.random_paulis(n, nterms, 3) if nterms else []
        e = eng.estimate(c, precision=prec, shots=shots, terms=terms, statevector=True)
        assert e.feasible == 1, e.infeasible_reason
        assert e.passes >= 1 and e.launches >= 1 + e.passes
        assert math.isclose(e.seconds, _formula(e, n, prec, shots, len(terms)), rel_tol=1e-12)


def test_prediction_grows_with_qubits_and_passes(profile):
    eng = H.HipDenseEngine()
    by_n = [eng.estimate(W.random_circuit(n, 4, 1)).seconds for n in (14, 18, 22, 26, 30)]
    assert all(a < b for a, b in zip(by_n, by_n[1:]))
    by_layers = [eng.estimate(W.random_circuit(24, d, 1)) for d in (2, 4, 8, 16)]
    assert all(a.passes < b.passes and a.seconds < b.seconds for a, b in zip(by_layers, by_layers[1:]))
    # At HBM scale the passes dominate: 33 qubits, 16 layers.
    e = eng.estimate(W.random_circuit(33, 16, 1))
    pass_part = e.passes * 2.0 * (16 << 33) / PROFILE["pass_bytes_per_second"]
    assert 0.9 < pass_part / e.seconds <= 1.0


def test_infeasible_is_reported_not_raised(profile):
    eng = H.HipDenseEngine()
    e = eng.estimate(W.random_circuit(20, 2, 1), budget_bytes=1 << 20)
    assert e.feasible == 0 and b"budget" in e.infeasible_reason
    e = eng.estimate(W.random_circuit(20, 2, 1), precision=7)
    assert e.feasible == 0 and e.infeasible_reason


def test_profile_rates_must_be_positive():
    bad = dict(PROFILE, pass_bytes_per_second=0.0)
    with pytest.raises(H.HipError):
        H.set_calibration(H.Calibration(**bad))
    bad = dict(PROFILE, launch_seconds=float("nan"))
    with pytest.raises(H.HipError):
        H.set_calibration(H.Calibration(**bad))


def deep_circuit(n, first, k, layers, seed):
    """`layers` layers of U3 on qubits first..first+k-1 and a CX brick among them:
    all the work on k qubits of an n-qubit state."""
    import numpy as np
    rng = np.random.default_rng(seed)
    c = W.Circuit(n)
    for layer in range(layers):
        for q in range(first, first + k):
            c.append(W.U, (q,), tuple(rng.uniform(0.0, 6.28, 3)))
        for q in range(first + layer % 2, first + k - 1, 2):
            c.append(W.CX, (q, q + 1))
    return c


def test_autotuner_keeps_window_passes_for_the_random_circuit(profile):
    eng = H.HipDenseEngine()
    for n in (24, 30, 33):
        c = W.random_circuit(n, 16, 2025)
        for prec in (H.C128, H.C64):
            auto = eng.estimate(c, precision=prec, fusion=H.FUSE_AUTO)
            plain = eng.estimate(c, precision=prec)
            assert (auto.fusion_used, auto.fusion_width_used) == (H.FUSE_CONTAINMENT, 0), (n, prec)
            assert auto.passes == plain.passes and auto.seconds == plain.seconds


@pytest.mark.parametrize("k", [4, 5, 6])
def test_autotuner_fuses_a_deep_narrow_circuit(profile, k):
    eng = H.HipDenseEngine()
    c = deep_circuit(24, 18, k, 40, k)
    for prec in (H.C128, H.C64):
        auto = eng.estimate(c, precision=prec, fusion=H.FUSE_AUTO)
        plain = eng.estimate(c, precision=prec)
        assert (auto.fusion_used, auto.fusion_width_used) == (H.FUSE_WIDTH, k), prec
        assert auto.passes == 1 and auto.seconds < plain.seconds


def test_autotuner_follows_the_profile(profile):
    """With arithmetic a thousand times slower, one wide block costs more than
    the window passes it replaces: the containment plan is kept."""
    H.set_calibration(H.Calibration(**dict(PROFILE, fp64_ops_per_second=3.0e10, mfma_flops_per_second=1.0e11)))
    eng = H.HipDenseEngine()
    c = deep_circuit(24, 18, 6, 40, 6)
    for prec in (H.C128, H.C64):
        auto = eng.estimate(c, precision=prec, fusion=H.FUSE_AUTO)
        assert auto.fusion_used == H.FUSE_CONTAINMENT, prec
