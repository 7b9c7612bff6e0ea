"""Benchmark: the dense random circuit at 33 qubits, complex128, on one MI355X.

A step is one execution of the whole circuit (16 layers; each layer U3 on every
qubit then CZ on the brick pattern) on the 128 GiB state already resident in
HBM: initialisation plus every gate pass. `value` is the achieved HBM bandwidth
of the gate passes: the algorithmic bytes they move (one read and one write of
the state per pass, 2 * 16 * 2^33 B) divided by their device time, measured
with HIP events on the engine's stream around every pass.

Prints one JSON line. Progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0  # MI355X datasheet
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "window_kernel_33q_c128_traffic.md")


def measured_traffic(qubits: int, precision: str):
    """HBM bytes per window-kernel launch from the committed PMC summary
    (FETCH_SIZE + WRITE_SIZE scaled by the library's copy kernel), for the
    33-qubit complex128 pass it was measured on; None for other shapes."""
    if qubits != 33 or precision != "c128" or not os.path.exists(TRAFFIC_PROFILE):
        return None
    with open(TRAFFIC_PROFILE) as f:
        for line in f:
            if line.startswith("traffic_bytes_per_launch:"):
                return int(line.split(":")[1])
    return None


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_cores() -> int:
    """Cores this process may run on (what `nproc` reports)."""
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_baseline(qubits: int, layers: int, threads: int, gpu_passes: int, full_qubits: int,
                 full_layers: int, second_qubits: int = 30, second_layers: int = 2) -> dict:
    """The reference CPU engine (oracle/_ref, DenseEngine::execute with the
    reference's containment plan) on bounded samples of the same workload: the
    random circuit at `qubits` qubits and `layers` layers, then at
    `second_qubits` qubits and `second_layers` layers. `value` is in the unit of
    the GPU figure: the GPU pass plan's algorithmic bytes for the first sample
    circuit (gpu_passes x 2 x 16 x 2^qubits) over the CPU's execute time, so
    value / cpu value is the speed-up on the same circuit. Seconds per layer are
    reported at both sizes; the full workload's CPU time is extrapolated from
    them (the per-qubit growth measured between the two samples, applied from
    the larger one) and labelled as such."""
    import oracle
    from menos_amd import workloads as W

    def run(n, depth):
        c = W.random_circuit(n, depth, 2025)
        if oracle.Reference.available():
            _, _, _, secs = oracle.Reference().execute(c, fuse_gates=True, statevector=False, threads=threads)
            return secs, "reference"
        t0 = time.perf_counter()
        oracle.Oracle().execute(c, fuse_gates=True, statevector=False, threads=threads)
        return time.perf_counter() - t0, "port"

    secs, kind = run(qubits, layers)
    secs2, _ = run(second_qubits, second_layers)
    algo = gpu_passes * 2 * 16 * (1 << qubits)
    per_layer = secs / layers
    per_layer2 = secs2 / second_layers
    growth = (per_layer2 / per_layer) ** (1.0 / (second_qubits - qubits))
    full = per_layer2 * growth ** (full_qubits - second_qubits) * full_layers
    return {"value": round(algo / secs / 1e9, 3), "unit": "GB/s", "cores": threads, "kind": kind,
            "sample": f"random circuit {qubits} qubits complex128, {layers} layers, reference containment plan, "
                      f"{secs:.2f} s on {threads} threads (nproc {host_cores()}); GB/s = the GPU pass plan's "
                      f"{gpu_passes} passes x 2 x 16 x 2^{qubits} B over that time; then {second_qubits} qubits, "
                      f"{second_layers} layers, {secs2:.2f} s",
            "seconds_per_layer": {str(qubits): round(per_layer, 4), str(second_qubits): round(per_layer2, 4)},
            "growth_per_qubit": round(growth, 3),
            "extrapolated_full_seconds": round(full, 1),
            "extrapolated_full": f"{full_qubits} qubits, {full_layers} layers: {full:.0f} s (extrapolated from the "
                                 f"{qubits}- and {second_qubits}-qubit samples at x{growth:.2f} per qubit, "
                                 "not measured)"}


def batched(args, H, W) -> dict:
    """The batched estimator workload: `args.batch_circuits` ansatz circuits
    (n uniform on 10..22, depth 4..12, RY and RZ on every qubit then a CNOT
    ladder per layer, 8 Pauli strings of weight 1..4 each), all executed by
    mh_batch_execute at complex64 and complex128. The error is the largest
    |<P>_c64 - <P>_c128| over every circuit and term. The per-circuit GPU path
    (one mh_engine_execute per circuit) is timed on the first
    `args.per_circuit_sample` circuits. Each form runs once untimed first."""
    import numpy as np
    circuits, terms = W.batch_workload(args.batch_circuits, 2026)
    # complex64 (the tolerance form) keeps controlled-NOT blocks as controlled pairs; complex128 keeps the
    # reference's containment plan. The per-circuit path uses the same fusion per precision.
    fusion = {H.C64: H.FUSE_PAIRS, H.C128: H.FUSE_CONTAINMENT}
    for prec in (H.C64, H.C128):  # one untimed run of the whole batch per precision (steady state)
        H.batch_execute(circuits, terms, precision=prec, fusion=fusion[prec])
    out = {"workload": f"{len(circuits)} ansatz circuits, n 10..22, depth 4..12, 8 Paulis each",
           "circuits": len(circuits), "fusion": {"c64": "pairs", "c128": "containment"}}
    res = {}
    for name, prec in (("c64", H.C64), ("c128", H.C128)):
        r = H.batch_execute(circuits, terms, precision=prec, fusion=fusion[prec])
        res[name] = r
        lat = np.sort(r.completion_seconds) * 1e3
        out[name] = {"circuits_per_s": round(len(circuits) / r.total_seconds, 1),
                     "seconds": round(r.total_seconds, 4), "plan_seconds": round(r.plan_seconds, 4),
                     "device_seconds": round(r.device_seconds, 4), "rounds": r.rounds, "launches": r.launches,
                     "p50_ms": round(float(np.percentile(lat, 50)), 2),
                     "p99_ms": round(float(np.percentile(lat, 99)), 2)}
        log(f"batched {name}: {out[name]}")
    out["max_abs_err_c64_vs_c128"] = float(max(np.max(np.abs(a - b)) if a.size else 0.0
                                               for a, b in zip(res["c64"].expectation, res["c128"].expectation)))
    k = min(args.per_circuit_sample, len(circuits))
    if k == 0:  # a profile of the batch alone
        return out
    eng = H.HipDenseEngine()
    eng.execute(circuits[0], precision=H.C64, fusion=fusion[H.C64], terms=terms[0], statevector=False, targets=[],
                width=0)
    per = {}
    for name, prec in (("c64", H.C64), ("c128", H.C128)):
        t0 = time.perf_counter()
        for c, t in zip(circuits[:k], terms[:k]):
            eng.execute(c, precision=prec, fusion=fusion[prec], terms=t, statevector=False, targets=[], width=0)
        per[name] = k / (time.perf_counter() - t0)
    out["per_circuit_path"] = {"sample": k, "c64_circuits_per_s": round(per["c64"], 1),
                               "c128_circuits_per_s": round(per["c128"], 1)}
    out["speedup_c64"] = round(out["c64"]["circuits_per_s"] / per["c64"], 2)
    out["speedup_c128"] = round(out["c128"]["circuits_per_s"] / per["c128"], 2)
    log(f"per-circuit path on {k} circuits: {out['per_circuit_path']}; speedup c64 {out['speedup_c64']}x")
    if args.graph_iterations > 0:
        out["graph"] = batched_graph(args, H, circuits, terms, fusion, per)
    return out


def batched_graph(args, H, circuits, terms, fusion, per) -> dict:
    """The variational loop on the resident, graph-captured batch
    (mh_batch_graph_*): the same circuits with new angles every iteration.
    An iteration is set_params (one array of every circuit's angles; block
    matrices and pass programs rebuilt on host threads, the matrix table
    staged) plus run (one graph launch: matrix upload,
    initialisation, every pass, the term reductions, the terms back). Rates are
    over `args.graph_iterations` iterations after one untimed one; the last
    iteration's terms are checked against mh_batch_execute on the same angles
    (bit for bit) and complex64 against complex128 (largest |difference|)."""
    import numpy as np
    sched = dict(group_bytes=args.group_bytes, kernel_streams=args.kernel_streams, cached_states=args.cached_states)
    out = {"iterations": args.graph_iterations, "schedule": sched}
    last = {}
    for name, prec in (("c64", H.C64), ("c128", H.C128)):
        rng = np.random.default_rng(7)
        with H.BatchGraph(circuits, terms, precision=prec, fusion=fusion[prec], **sched) as g:
            g.run()
            upd_s = run_s = 0.0
            lat = []
            sizes = [c.num_angles for c in circuits]
            assert g.param_count() == sum(sizes)
            for it in range(args.graph_iterations + 1):
                params = rng.uniform(-np.pi, np.pi, sum(sizes))
                u0 = time.perf_counter()
                g.set_params(params)
                u1 = time.perf_counter()
                e, res = g.run()
                u2 = time.perf_counter()
                if it == 0:
                    continue
                upd_s += u1 - u0
                run_s += u2 - u1
                lat.append((u2 - u0) * 1e3)
            cuts = np.cumsum([0] + sizes)
            new = [c.with_angles(params[cuts[i]:cuts[i + 1]]) for i, c in enumerate(circuits)]
            check = H.batch_execute(new, terms, precision=prec, fusion=fusion[prec], **sched).expectation
            same = all(a.tobytes() == b.tobytes() for a, b in zip(e, check))
            last[name] = e
            it_n = args.graph_iterations
            rate = it_n * len(circuits) / (upd_s + run_s)
            out[name] = {"circuits_per_s": round(rate, 1),
                         "replay_circuits_per_s": round(it_n * len(circuits) / run_s, 1),
                         "update_ms": round(upd_s / it_n * 1e3, 2), "run_ms": round(run_s / it_n * 1e3, 2),
                         "host_plan_ms": round(res.plan_seconds * 1e3, 2),
                         "host_compile_ms": round(res.compile_seconds * 1e3, 2),
                         "graph_launches": res.launches,
                         "p50_ms": round(float(np.percentile(lat, 50)), 2),
                         "p99_ms": round(float(np.percentile(lat, 99)), 2),
                         "equals_batch_execute": same,
                         "speedup_vs_per_circuit": round(rate / per[name], 2)}
        log(f"graph {name}: {out[name]}")
    out["max_abs_err_c64_vs_c128"] = float(max(np.max(np.abs(a - b)) if a.size else 0.0
                                               for a, b in zip(last["c64"], last["c128"])))
    return out


def served(args, H, W) -> dict:
    """The served workload through the job server (mh_server_*): `args.serve_jobs`
    jobs (n log-uniform on 12..30; 70% estimator jobs, an ansatz of depth 4..12
    with 16 observables of 1..4 Pauli terms, at `args.serve_estimator_precision`;
    30% sampler jobs, 20 layers of U3 + CZ brick with 1000 shots, complex128;
    40% parameter-only updates of an earlier job's circuit), under a 256 GB
    ledger budget. First every job is submitted at once: capacity = jobs / wall
    time. Then the same jobs arrive as a Poisson process at 0.8 x that capacity
    (unless `args.serve_rate` is set): circuits/s, p50 and p99 of
    submit -> result latency."""
    import numpy as np
    prec = H.C64 if args.serve_estimator_precision == "c64" else H.C128
    t0 = time.perf_counter()
    specs = W.serve_workload(args.serve_jobs, 1.0, 2027)
    gaps = np.array([s.arrival for s in specs])  # arrivals at rate 1/s, rescaled below

    def jobs_of():
        return [H.Job(H.JOB_ESTIMATOR if s.kind == "estimator" else H.JOB_SAMPLER, s.circuit,
                      precision=prec if s.kind == "estimator" else H.C128, observables=s.observables,
                      shots=s.shots, seed=s.seed) for s in specs]

    jobs = jobs_of()
    log(f"served: {len(jobs)} jobs built in {time.perf_counter() - t0:.1f} s")
    out = {"workload": f"{len(jobs)} jobs, n log-uniform 12..30; 70% estimator (ansatz depth 4..12, 16 observables, "
                       f"{args.serve_estimator_precision}), 30% sampler (U3 + CZ brick, 20 layers, 1000 shots, "
                       "c128); 40% parameter-only updates; 256 GB ledger budget; complex64 jobs with pair-keeping "
                       "fusion",
           "large_job_bytes": args.serve_large_bytes,
           "jobs": len(jobs),
           "estimators": sum(s.kind == "estimator" for s in specs),
           "updates": sum(s.update for s in specs)}
    if args.serve_skip_capacity:
        if not args.serve_rate:
            raise SystemExit("--serve-skip-capacity needs --serve-rate")
        return served_poisson(args, H, jobs_of, gaps, out, args.serve_rate)
    with H.Server(budget_bytes=256_000_000_000, reserve_bytes=256_000_000_000, complex64_pairs=True,
                  large_job_bytes=args.serve_large_bytes) as srv:
        w0 = time.perf_counter()
        res = srv.replay(jobs, np.zeros(len(jobs)))
        wall = time.perf_counter() - w0
        st = srv.stats()
    failed = [r for r in res if r.status != 0]
    if failed:
        raise RuntimeError(f"{len(failed)} served jobs failed: {failed[0].error}")
    cap = len(jobs) / wall
    out["capacity"] = {"circuits_per_s": round(cap, 2), "seconds": round(wall, 3), "rounds": st.rounds,
                       "plan_cache_hits": st.plan_cache_hits, "plan_cache_misses": st.plan_cache_misses,
                       "peak_round_GB": round(st.peak_round_bytes / 1e9, 2),
                       "phase_seconds": {"plan": round(st.plan_seconds, 3), "arena": round(st.arena_seconds, 3),
                                         "compile": round(st.compile_seconds, 3),
                                         "passes": round(st.pass_seconds, 3),
                                         "expectation": round(st.expectation_seconds, 3),
                                         "sample": round(st.sample_seconds, 3)}}
    log(f"served capacity: {out['capacity']}")
    return served_poisson(args, H, jobs_of, gaps, out, args.serve_rate or 0.8 * cap)


def replay_with_progress(srv, jobs, arrivals, every=1000) -> list:
    """mh_server_replay's arrival process driven from Python, for long runs: a
    submitter thread submits job i at arrivals[i] seconds from the start while
    this thread waits for the tickets in order, logging every `every` results
    (mh_server_replay reports nothing until its last job)."""
    import threading
    n = len(jobs)
    tickets = [None] * n
    cv = threading.Condition()
    t0 = time.perf_counter()

    def submitter():
        for i in range(n):
            delay = t0 + float(arrivals[i]) - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            t = srv.submit(jobs[i])
            with cv:
                tickets[i] = t
                cv.notify_all()

    th = threading.Thread(target=submitter, daemon=True)
    th.start()
    results = []
    for i in range(n):
        with cv:
            while tickets[i] is None:
                cv.wait()
        results.append(srv.wait(tickets[i]))
        if (i + 1) % every == 0:
            log(f"served Poisson: {i + 1} of {n} jobs done at {time.perf_counter() - t0:.1f} s")
    th.join()
    return results


def served_poisson(args, H, jobs_of, gaps, out, rate) -> dict:
    """The served jobs arriving as a Poisson process at `rate` jobs/s."""
    import numpy as np
    arrivals = gaps / rate
    jobs = jobs_of()
    with H.Server(budget_bytes=256_000_000_000, reserve_bytes=256_000_000_000, complex64_pairs=True,
                  large_job_bytes=args.serve_large_bytes) as srv:
        w0 = time.perf_counter()
        res = srv.replay(jobs, arrivals) if len(jobs) <= 2000 else replay_with_progress(srv, jobs, arrivals)
        wall = time.perf_counter() - w0
        st = srv.stats()
    failed = [r for r in res if r.status != 0]
    if failed:
        raise RuntimeError(f"{len(failed)} served jobs failed: {failed[0].error}")
    lat = np.array([r.latency for r in res]) * 1e3
    out["poisson"] = {"rate_per_s": round(rate, 2), "circuits_per_s": round(len(jobs) / wall, 2),
                      "seconds": round(wall, 3), "p50_ms": round(float(np.percentile(lat, 50)), 2),
                      "p99_ms": round(float(np.percentile(lat, 99)), 2), "rounds": st.rounds,
                      "plan_cache_hits": st.plan_cache_hits, "busy_seconds": round(st.busy_seconds, 3)}
    log(f"served Poisson: {out['poisson']}")
    return out


def complex64_report(args, H, W) -> dict:
    """The complex64 form: per-pass HBM rate of the 16-layer random circuit at
    34 qubits (128 GiB), and at 33 qubits its error against the complex128
    run of the same gates: the largest amplitude difference and the largest
    <Z_i Z_i+1> difference."""
    import numpy as np
    out = {}
    n34 = args.c64_qubits
    c = W.random_circuit(n34, args.layers, 2025)
    eng = H.HipDenseEngine()
    r = eng.execute(c, precision=H.C64, statevector=False, record_pass_times=True)
    ms = r.pass_ms.astype(np.float64)
    out["c64_qubits"] = n34
    out["c64_passes"] = int(r.num_passes)
    out["c64_gate_seconds"] = round(float(r.gate_seconds), 4)
    out["c64_pass_GBps"] = round(2 * 8 * (1 << n34) * len(ms) / (ms.sum() * 1e-3) / 1e9, 1)
    log(f"complex64 {n34} qubits: {out['c64_passes']} passes in {r.gate_seconds:.3f} s, "
        f"{out['c64_pass_GBps']} GB/s per pass")
    n = args.qubits
    c = W.random_circuit(n, args.layers, 2025)
    blocks = [(H.gate_matrix(op, params), list(qs)) for op, qs, params in c.ops]
    states = {}
    for prec in (H.C128, H.C64):
        d = H.DeviceState(n, prec)
        d.init()
        d.apply_blocks(blocks)
        states[prec] = d
    terms = W.zz_neighbour_terms(n)
    e128 = states[H.C128].expectation(terms)
    e64 = states[H.C64].expectation(terms)
    worst = states[H.C128].max_abs_diff(states[H.C64])
    for d in states.values():
        d.free()
    out["error_qubits"] = n
    out["max_amplitude_error"] = worst
    out["max_zz_error"] = float(np.max(np.abs(e128 - e64)))
    log(f"complex64 vs complex128 at {n} qubits: max amplitude error {worst:.3g}, "
        f"max <ZZ> error {out['max_zz_error']:.3g}")
    return out


def sampler_report(args, H, W) -> dict:
    """Exact sampling of the 16-layer random circuit's final state at --qubits
    (complex128): the sampler's device time and its HBM rate. Its tables read
    the state twice (block sums, then the block transfers); each shot then
    walks one block of 2^13 amplitudes."""
    n = args.qubits
    c = W.random_circuit(n, args.layers, 2025)
    c.ops = [op for op in c.ops if op[0] != W.MEASURE]
    eng = H.HipDenseEngine()
    out = {"qubits": n, "precision": "c128", "shots": args.sampler_shots}
    targets = [(q, q) for q in range(n)]
    runs = []
    for i in range(2):
        r = eng.execute(c, shots=args.sampler_shots, seed=7 + i, targets=targets, width=n, statevector=False)
        runs.append(float(r.sample_seconds))
        log(f"sampler {n} qubits, {args.sampler_shots} shots: {r.sample_seconds * 1e3:.1f} ms "
            f"({len(r.counts)} distinct outcomes)")
    state = (1 << n) * 16
    best = min(runs)
    table_bytes = 2 * state
    out.update({"sample_seconds": [round(x, 4) for x in runs],
                "state_reads_GBps": round(table_bytes / best / 1e9, 1),
                "roofline": {"bound": "hbm", "achieved": round(table_bytes / best / 1e9, 1), "peak": PEAK_HBM_GBS,
                             "unit": "GB/s", "frac": round(table_bytes / best / 1e9 / PEAK_HBM_GBS, 4),
                             "algorithmic_bytes": table_bytes,
                             "note": "two reads of the state over the whole sampling call (host wall time)"}})
    return out


def tiered(args, H, W) -> dict:
    """The host-offloaded tier: the random circuit at --tier-qubits (complex128,
    16 GiB at 30 qubits) in pinned host memory, run through two device windows
    of 2^--tier-window amplitudes; the link bytes per second of its passes
    against the library's measured copy rates, the in-memory engine's time for
    the same circuit, and a full-size check (counts and terms equal to the
    in-memory engine's)."""
    import numpy as np
    n, window = args.tier_qubits, args.tier_window
    c = W.random_circuit(n, args.layers, 2025)
    c.ops = [op for op in c.ops if op[0] != W.MEASURE]
    cal = H.calibrate()
    tier, dense = H.HipHostTierEngine(), H.HipDenseEngine()
    terms = W.z_terms(n)[:4] + [((1 << (n - 1)) | 1, 1 << (n - 2)), ((1 << (n - 1)) | (1 << (n - 3)), 0)]
    kw = dict(shots=10000, seed=7, targets=[(q, q) for q in range(n)], width=n, terms=terms)
    est = tier.estimate(c, tier_window_qubits=window, **kw)
    log(f"tier: {n} qubits c128, window 2^{window}, chunk 2^{est.chunk_qubits}: {est.passes} passes, "
        f"{est.link_bytes / 2**30:.0f} GiB over the link, model {est.seconds:.1f} s")
    monitor = None
    if args.tier_progress:
        # A line per pass (the monitor waits for each pass's write-backs, so
        # passes no longer overlap at their boundaries: for long runs only).
        def monitor(done):
            log(f"tier pass {done}/{est.passes}")
    r = tier.execute(c, tier_window_qubits=window, statevector=False, record_pass_times=True, monitor=monitor, **kw)
    log(f"tier: execute {r.execute_seconds:.2f} s, passes {r.gate_seconds:.2f} s, sampling {r.sample_seconds:.2f} s, "
        f"terms {r.expectation_seconds:.2f} s")
    m = dense.execute(c, statevector=False, **kw)
    log(f"in memory: execute {m.execute_seconds:.2f} s, passes {m.gate_seconds:.3f} s")
    # The link's two directions are separate lanes: its ceiling for the passes
    # (half their bytes each way) is the sum of the measured one-way rates.
    link = {"h2d_GBps": round(cal.h2d_bytes_per_second / 1e9, 1), "d2h_GBps": round(cal.d2h_bytes_per_second / 1e9, 1)}
    peak = round(link["h2d_GBps"] + link["d2h_GBps"], 1)
    log(f"host link: {link['h2d_GBps']} GB/s in, {link['d2h_GBps']} GB/s out (pinned, one way at a time)")
    state = (1 << n) * 16
    passes = max(int(r.num_passes), 1)
    pass_link = state * passes + state * (passes - 1)
    ms = np.asarray(r.pass_ms, np.float64)
    per_pass = (2 * state / (ms[1:] * 1e-3)).tolist() if len(ms) > 1 else []
    return {
        "qubits": n, "precision": "c128", "layers": args.layers, "window_qubits": int(r.cost.window_qubits),
        "chunk_qubits": int(r.cost.chunk_qubits), "passes": int(r.num_passes),
        "host_state_GiB": state / 2**30, "device_bytes": int(r.cost.device_bytes),
        "link_bytes": int(r.cost.link_bytes), "model_seconds": round(float(r.cost.seconds), 3),
        "execute_seconds": round(float(r.execute_seconds), 3), "pass_seconds": round(float(r.gate_seconds), 3),
        "sample_seconds": round(float(r.sample_seconds), 3),
        "expectation_seconds": round(float(r.expectation_seconds), 3),
        "pass_link_GBps": round(pass_link / r.gate_seconds / 1e9, 1),
        "steady_pass_link_GBps_median": round(float(np.median(per_pass)) / 1e9, 1) if per_pass else None,
        "link_measured": link,
        "roofline": {"bound": "host link", "achieved": round(pass_link / r.gate_seconds / 1e9, 1),
                     "peak": peak, "unit": "GB/s", "frac": round(pass_link / r.gate_seconds / 1e9 / peak, 3),
                     "peak_source": "sum of the measured pinned host-to-device and device-to-host rates"},
        "in_memory_pass_seconds": round(float(m.gate_seconds), 4),
        "counts_equal_in_memory": r.counts == m.counts,
        "max_term_difference": float(np.max(np.abs(r.expectation - m.expectation))),
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--qubits", type=int, default=33)
    ap.add_argument("--layers", type=int, default=16)
    ap.add_argument("--precision", choices=["c128", "c64"], default="c128")
    ap.add_argument("--fusion", choices=["none", "containment", "width"], default="containment")
    ap.add_argument("--fusion-width", type=int, default=0)
    ap.add_argument("--window", type=int, default=0)
    ap.add_argument("--max-pass-cost", type=int, default=0)
    ap.add_argument("--cpu-qubits", type=int, default=28)
    ap.add_argument("--cpu-layers", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batch-circuits", type=int, default=4096)
    ap.add_argument("--per-circuit-sample", type=int, default=256)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--graph-iterations", type=int, default=5)
    ap.add_argument("--group-bytes", type=int, default=0, help="batch cache-group size in bytes (0: one launch per pass)")
    ap.add_argument("--kernel-streams", type=int, default=0)
    ap.add_argument("--cached-states", action="store_true")
    ap.add_argument("--only", choices=["batched", "served", "complex64", "tiered", "sampler"],
                    help="run one section alone and print its JSON object (for profiling)")
    ap.add_argument("--serve-jobs", type=int, default=1000)
    ap.add_argument("--serve-rate", type=float, default=0.0, help="arrivals per second; 0: 0.8 x measured capacity")
    ap.add_argument("--serve-estimator-precision", choices=["c64", "c128"], default="c64")
    ap.add_argument("--serve-large-bytes", type=int, default=1 << 32,
                    help="server large_job_bytes: jobs with states of this many bytes or more (4 GiB: 28 qubits "
                         "complex128, 29 complex64) run one at a time in rounds of their own; 0: arrival order only")
    ap.add_argument("--serve-skip-capacity", action="store_true",
                    help="only the Poisson replay, at --serve-rate (for the 20,000-job configuration)")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-c64-report", action="store_true",
                    help="skip the complex64 form at --c64-qubits and its error against complex128")
    ap.add_argument("--c64-qubits", type=int, default=34)
    ap.add_argument("--sampler-shots", type=int, default=10000)
    ap.add_argument("--tier-qubits", type=int, default=30)
    ap.add_argument("--no-tier", action="store_true", help="skip the host-offloaded tier section (about 25 s)")
    ap.add_argument("--tier-progress", action="store_true",
                    help="log each tier pass (waits at every pass boundary; for runs of minutes)")
    ap.add_argument("--tier-window", type=int, default=28)
    ap.add_argument("--copy-probe", action="store_true",
                    help="first run the library's calibration (its streaming copy of 2 GiB each way is the "
                         "reference for scaling HBM traffic counters)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); this engine has no multi-GPU path.",
              file=sys.stderr)
        return 2

    from menos_amd import hip as H
    from menos_amd import workloads as W

    if args.copy_probe:
        cal = H.calibrate()
        log(f"calibration: copy {cal.hbm_bytes_per_second / 1e9:.0f} GB/s")
    info = H.device_info()
    log(f"device {info.name.decode()} {info.arch.decode()} CUs={info.compute_units} "
        f"HBM={info.total_bytes / 2**30:.0f} GiB free={info.free_bytes / 2**30:.0f} GiB")
    if args.only:
        section = {"batched": batched, "served": served, "complex64": complex64_report, "tiered": tiered,
                   "sampler": sampler_report}[args.only]
        print(json.dumps({args.only: section(args, H, W)}), flush=True)
        return 0
    precision = H.C128 if args.precision == "c128" else H.C64
    amp = 16 if precision == H.C128 else 8
    fusion = {"none": H.FUSE_NONE, "containment": H.FUSE_CONTAINMENT, "width": H.FUSE_WIDTH}[args.fusion]
    n = args.qubits
    circuit = W.random_circuit(n, args.layers, 2025)
    eng = H.HipDenseEngine()
    est = eng.estimate(circuit, precision=precision, fusion=fusion, fusion_width=args.fusion_width,
                       window_qubits=args.window, max_pass_cost=args.max_pass_cost)
    if not est.feasible:
        log(f"infeasible: {est.infeasible_reason.decode()}")
        return 1
    log(f"{n} qubits {args.precision}, {len(circuit.ops)} gates -> {est.blocks} blocks, {est.passes} passes, "
        f"state {est.device_bytes / 2**30:.1f} GiB")

    def step():
        return eng.execute(circuit, precision=precision, fusion=fusion, fusion_width=args.fusion_width,
                           window_qubits=args.window, statevector=False, record_pass_times=True,
                           max_pass_cost=args.max_pass_cost)

    for i in range(args.warmup):
        r = step()
        log(f"warmup {i}: execute {r.execute_seconds:.3f} s, gate passes {r.gate_seconds:.3f} s")
    pass_ms, gate_s, exec_s = [], 0.0, 0.0
    t0 = time.perf_counter()
    for i in range(args.steps):
        r = step()
        pass_ms.extend(r.pass_ms.tolist())
        gate_s += r.gate_seconds
        exec_s += r.execute_seconds
        log(f"step {i}: execute {r.execute_seconds:.3f} s, {r.num_passes} passes in {r.gate_seconds:.3f} s")
    wall = time.perf_counter() - t0
    passes = r.num_passes
    pass_bytes = 2 * amp * (1 << n)
    total_pass_s = sum(pass_ms) * 1e-3
    value = pass_bytes * len(pass_ms) / total_pass_s / 1e9
    avg_ms = sum(pass_ms) / len(pass_ms)
    achieved = pass_bytes / (avg_ms * 1e-3) / 1e9

    out = {
        "metric": "HBM GB/s vs peak at 33q c128; batched circuits/sec + p99 latency at err<=1e-5",
        "value": round(value, 1),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if precision == H.C128 else "f32",
        "data": "synthetic",
        "config": {
            "workload": f"dense random circuit, {n} qubits {args.precision} ({(1 << n) * amp / 2**30:.0f} GiB), "
                        f"{args.layers} layers of U3 on every qubit + CZ brick, {args.fusion} fusion",
            "num_qubits": n, "layers": args.layers, "precision": args.precision, "fusion": args.fusion,
            "gates": len(circuit.ops), "blocks": r.num_blocks, "passes_per_step": passes,
            "gate_seconds_per_step": round(gate_s / args.steps, 4),
            "layers_per_second": round(args.layers * args.steps / gate_s, 3),
            "window_qubits": args.window or "default",
            "max_pass_cost": args.max_pass_cost or "default",
        },
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM_GBS, 4),
                     "traffic": measured_traffic(n, args.precision),
                     "traffic_source": "rocprofv3 FETCH_SIZE + WRITE_SIZE, profiles/window_kernel_33q_c128_traffic.md",
                     "kernel": "window_kernel", "avg_launch_ms": round(avg_ms, 4),
                     "algorithmic_bytes_per_launch": pass_bytes},
    }
    if not args.no_tier:
        # Before the sections that free device memory: for some tens of seconds
        # after a large free (the served section's arena) the host link is
        # shared, and the same run measured a third slower.
        out["tiered"] = tiered(args, H, W)
    if not args.no_batch:
        out["batched"] = batched(args, H, W)
    if not args.no_serve:
        out["served"] = served(args, H, W)
    if not args.no_c64_report:
        out["complex64"] = complex64_report(args, H, W)
    if not args.no_cpu:
        threads = host_cores()
        sample = W.random_circuit(args.cpu_qubits, args.cpu_layers, 2025)
        gpu_passes = eng.estimate(sample, precision=H.C128).passes
        log(f"cpu baseline: {args.cpu_qubits} qubits, {args.cpu_layers} layers, {threads} threads "
            f"(GPU plan: {gpu_passes} passes)")
        out["cpu_baseline"] = cpu_baseline(args.cpu_qubits, args.cpu_layers, threads, gpu_passes, n, args.layers)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
