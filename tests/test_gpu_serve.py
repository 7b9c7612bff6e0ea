"""The job server (mh_server_*) against the single-circuit engine and the
oracle: raw terms equal to mh_engine_execute's bit for bit, observable values
equal to the reference's combine_expectation applied to them, sampler counts
equal to the engine's and the oracle's for the same seed, plan reuse for
parameter-only updates, admission against the budget, and arrival replay."""
import math

import numpy as np
import pytest

import oracle
from menos_amd import workloads as W

pytestmark = pytest.mark.gpu


def _estimator(hip, n, depth, seed, precision=0, observables=4):
    rng = np.random.default_rng(seed)
    c = W.FlatAnsatz(n, depth, rng.uniform(-math.pi, math.pi, 2 * n * depth))
    obs = []
    for o in range(observables):
        ps = W.random_paulis(n, 1 + o % 3, seed * 31 + o, max_weight=4)
        obs.append([(x, z, complex(rng.uniform(-1, 1), rng.uniform(-1, 1) if o % 2 else 0.0)) for x, z in ps])
    return c, obs, hip.Job(hip.JOB_ESTIMATOR, c, precision=precision, observables=obs)


def _sampler(hip, n, layers, seed, shots=1000):
    rng = np.random.default_rng(seed)
    c = W.FlatBrick(n, layers, rng.uniform(0, 2 * math.pi, 3 * n * layers))
    return c, hip.Job(hip.JOB_SAMPLER, c, shots=shots, seed=seed * 7919 + 1)


def _check_estimator(hip, c, obs, r, precision=0):
    assert r.status == 0, r.error
    terms = [(x, z) for o in obs for x, z, _ in o]
    single = hip.HipDenseEngine().execute(c, precision=precision, terms=terms, statevector=False, targets=[], width=0)
    assert r.raw_terms.tobytes() == single.expectation.tobytes()
    at = 0
    for k, o in enumerate(obs):
        want = oracle.combine_expectation(r.raw_terms[at:at + len(o)], [t[2] for t in o])
        got = (float(r.values[k].real), float(r.values[k].imag))
        assert np.array(got).tobytes() == np.array(want).tobytes(), k
        at += len(o)


def test_server_matches_engine_and_oracle(orc, hip):
    jobs = []
    with hip.Server() as s:
        for i in range(6):
            c, obs, j = _estimator(hip, 10 + i % 4, 2 + i % 3, 40 + i)
            jobs.append(("e", c, obs, j, s.submit(j)))
        for i in range(3):
            c, j = _sampler(hip, 10 + i, 3, 60 + i)
            jobs.append(("s", c, None, j, s.submit(j)))
        for kind, c, obs, j, t in jobs:
            r = s.wait(t)
            if kind == "e":
                _check_estimator(hip, c, obs, r)
            else:
                assert r.status == 0, r.error
                tg = c.measure_targets()
                single = hip.HipDenseEngine().execute(c, shots=j.shots, seed=j.c.seed, targets=tg,
                                                      width=c.num_qubits, statevector=False)
                assert r.counts == single.counts
                _, idx, _ = orc.execute(c, shots=j.shots, seed=j.c.seed, statevector=False)
                assert r.counts == oracle.counts_from_indices(idx, tg)
                assert sum(r.counts.values()) == j.shots
        st = s.stats()
        assert st.jobs_completed == len(jobs) and st.jobs_failed == 0


def test_large_jobs_one_at_a_time(hip):
    """large_job_bytes: jobs whose states reach it run in rounds of their own
    and never at once, smaller ones pass them, and every result is still the
    single engine's."""
    large = 16 << 16  # complex128 states of 16+ qubits
    jobs = []
    with hip.Server(large_job_bytes=large, lanes=3) as s:
        for i in range(9):
            if i % 3 == 0:
                c, j = _sampler(hip, 16 + i % 2, 4, 80 + i)
                jobs.append(("s", c, None, j, s.submit(j)))
            else:
                c, obs, j = _estimator(hip, 10 + i % 3, 2, 90 + i)
                jobs.append(("e", c, obs, j, s.submit(j)))
        results = [s.wait(t) for *_, t in jobs]
        st = s.stats()
    assert st.jobs_completed == len(jobs) and st.jobs_failed == 0
    big = [r for (kind, *_), r in zip(jobs, results) if kind == "s"]
    spans = sorted((r.started, r.completed) for r in big)
    assert all(b[0] >= a[1] for a, b in zip(spans, spans[1:]))
    big_rounds = {r.round for r in big}
    assert len(big_rounds) == len(big)
    assert not any(r.round in big_rounds for (kind, *_), r in zip(jobs, results) if kind == "e")
    for (kind, c, obs, j, _), r in zip(jobs, results):
        if kind == "e":
            _check_estimator(hip, c, obs, r)
        else:
            single = hip.HipDenseEngine().execute(c, shots=j.shots, seed=j.c.seed, targets=c.measure_targets(),
                                                  width=c.num_qubits, statevector=False)
            assert r.counts == single.counts


def test_parameter_updates_reuse_the_plan(hip):
    c, obs, j = _estimator(hip, 12, 3, 5)
    c2, _, _ = _estimator(hip, 12, 3, 6)  # same structure, fresh angles
    j2 = hip.Job(hip.JOB_ESTIMATOR, c2, observables=obs)
    with hip.Server() as s:
        r1 = s.wait(s.submit(j))
        r2 = s.wait(s.submit(j2))
        assert not r1.plan_cache_hit and r2.plan_cache_hit
        st = s.stats()
        assert st.plan_cache_hits == 1 and st.plan_cache_misses == 1
    _check_estimator(hip, c, obs, r1)
    _check_estimator(hip, c2, obs, r2)
    assert r1.raw_terms.tobytes() != r2.raw_terms.tobytes()


def test_budget_admission(hip):
    big = W.FlatAnsatz(23, 1, np.zeros(2 * 23))          # 128 MiB at complex128
    small_c, small_obs, small = _estimator(hip, 11, 2, 9)
    with hip.Server(budget_bytes=64 << 20) as s:
        assert s.stats().budget_bytes == 64 << 20
        t_big = s.submit(hip.Job(hip.JOB_ESTIMATOR, big, observables=[[(0, 1, 1.0)]]))
        t_small = s.submit(small)
        r_big, r_small = s.wait(t_big), s.wait(t_small)
    assert r_big.status == hip.MH_ERR_RESOURCE and "budget" in r_big.error
    _check_estimator(hip, small_c, small_obs, r_small)


def test_rounds_follow_the_budget_and_cap(hip):
    specs = [_estimator(hip, 14, 2, 100 + i) for i in range(10)]   # 256 KiB states
    with hip.Server(budget_bytes=3 * (16 << 14)) as s:             # three states a round
        tickets = [s.submit(j) for _, _, j in specs]
        results = [s.wait(t) for t in tickets]
        assert s.stats().rounds >= 4
    for (c, obs, _), r in zip(specs, results):
        _check_estimator(hip, c, obs, r)
    with hip.Server(max_round_jobs=2) as s:
        results2 = [s.wait(t) for t in [s.submit(j) for _, _, j in specs]]
        assert s.stats().rounds >= 5
    for a, b in zip(results, results2):
        assert a.values.tobytes() == b.values.tobytes()


def test_replay_gives_the_closed_loop_results(hip):
    jobs = W.serve_workload(30, 200.0, 3, qubits=(10, 14), sampler_layers=3, ansatz_depth=(2, 4), observables=4)
    hjobs = [hip.Job(hip.JOB_ESTIMATOR if j.kind == "estimator" else hip.JOB_SAMPLER, j.circuit,
                     observables=j.observables, shots=j.shots, seed=j.seed) for j in jobs]
    with hip.Server() as s:
        closed = [s.wait(t) for t in [s.submit(j) for j in hjobs]]
    with hip.Server() as s:
        timed = s.replay(hjobs, [j.arrival for j in jobs])
        st = s.stats()
    assert st.jobs_completed == len(jobs)
    # Jobs of one structure planned in the same round may all miss; every job looks once.
    assert st.plan_cache_hits + st.plan_cache_misses == len(jobs)
    assert st.plan_cache_hits >= 1
    for a, b, spec in zip(closed, timed, jobs):
        assert a.status == 0 and b.status == 0
        assert b.completed >= b.submitted >= spec.arrival - 1e-3
        if spec.kind == "estimator":
            assert a.values.tobytes() == b.values.tobytes()
        else:
            assert a.counts == b.counts


def test_small_jobs_waited_first_beside_a_large_one(hip):
    """One large job and several at most an eighth of its size in one round:
    the small ones are waited for first, while the large one may still run;
    every job completes once, with the single engine's results."""
    big_c, big_obs, big = _estimator(hip, 20, 3, 500)
    small = [_estimator(hip, 10 + i % 3, 2, 510 + i) for i in range(6)]
    with hip.Server(lanes=1) as s:
        t_big = s.submit(big)
        t_small = [s.submit(j) for _, _, j in small]
        r_small = [s.wait(t) for t in t_small]
        r_big = s.wait(t_big)
        st = s.stats()
    assert st.jobs_completed == 7 and st.jobs_failed == 0
    for (c, obs, _), r in zip(small, r_small):
        _check_estimator(hip, c, obs, r)
    _check_estimator(hip, big_c, big_obs, r_big)


def test_arrivals_overlap_a_running_round(hip):
    """Continuous batching: jobs that arrive while a long round runs are
    admitted by another lane, run beside it and complete first; each result is
    the one the job gets alone."""
    big_c, big_j = _sampler(hip, 28, 12, 600)     # 4 GiB, tens of milliseconds of passes
    small = [_estimator(hip, 12, 3, 610 + i) for i in range(8)]
    jobs = [big_j] + [j for _, _, j in small]
    arrivals = [0.0] + [0.005 + 0.003 * i for i in range(len(small))]
    with hip.Server(lanes=3) as s:
        res = s.replay(jobs, arrivals)
        st = s.stats()
    assert st.jobs_completed == len(jobs) and st.jobs_failed == 0
    assert st.lanes == 3 and st.peak_concurrent_rounds >= 2
    r_big = res[0]
    assert r_big.status == 0, r_big.error
    beside = [r for r in res[1:] if r.lane != r_big.lane]
    assert beside and any(r.completed < r_big.completed for r in beside)
    for (c, obs, _), r in zip(small, res[1:]):
        _check_estimator(hip, c, obs, r)
    single = hip.HipDenseEngine().execute(big_c, shots=big_j.shots, seed=big_j.c.seed, targets=big_c.measure_targets(),
                                          width=big_c.num_qubits, statevector=False)
    assert r_big.counts == single.counts


@pytest.mark.parametrize("pairs", [False, True])
def test_complex64_estimator_error(hip, pairs):
    """complex64 estimator jobs (under the containment plan or, with
    complex64_pairs, the pair-keeping one) within 1e-5 of complex128 jobs, whose
    raw terms stay the single engine's under the containment plan."""
    specs = [_estimator(hip, 12 + i, 4, 200 + i, precision=1) for i in range(3)]
    with hip.Server(complex64_pairs=pairs) as s:
        r64 = [s.wait(s.submit(j)) for _, _, j in specs]
        r128 = [s.wait(s.submit(hip.Job(hip.JOB_ESTIMATOR, c, observables=obs))) for c, obs, _ in specs]
    for (c, obs, _), a, b in zip(specs, r64, r128):
        assert a.status == 0 and b.status == 0
        assert np.max(np.abs(a.values - b.values)) <= 1e-5
        _check_estimator(hip, c, obs, b)


def test_invalid_jobs_are_refused_at_submission(hip):
    c = W.FlatAnsatz(6, 1, np.zeros(12))
    with hip.Server() as s:
        with pytest.raises(hip.HipError):
            s.submit(hip.Job(hip.JOB_ESTIMATOR, c, observables=[[(0, 1, 1.0)]]))
        c12 = W.FlatBrick(12, 1, np.zeros(36))
        with pytest.raises(hip.HipError):
            s.submit(hip.Job(hip.JOB_SAMPLER, c12, precision=hip.C64, shots=10))
        assert s.stats().jobs_completed == 0


def test_ticket_rules_and_shutdown(hip):
    """A ticket is waited for once; unknown tickets and decreasing arrival
    times are refused; a server closes with jobs nobody waited for."""
    c, obs, j = _estimator(hip, 11, 2, 300)
    s = hip.Server()
    t = s.submit(j)
    r = s.wait(t)
    assert r.status == 0
    s._pending[t] = j                       # ask again for the collected ticket
    with pytest.raises(hip.HipError):
        s.wait(t)
    s._pending[12345] = j
    with pytest.raises(hip.HipError):
        s.wait(12345)
    with pytest.raises(hip.HipError):
        s.replay([j, j], [0.2, 0.1])
    for _ in range(3):
        s.submit(j)                          # left for the shutdown to complete
    s.close()
