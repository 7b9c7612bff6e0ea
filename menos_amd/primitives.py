"""Estimator and sampler primitives on the job server.

The reference's Qiskit primitives run one ``menos_run`` per observable
(``MENOSEstimator._run_pub``, python/menos/qiskit/estimator.py:91-146) and one
per pub (``MENOSSampler._run_pub``, sampler.py:93-129), and refuse
parameterised circuits. Here every pub becomes jobs of one resident server
(``hip.Server``, ``mh_server_*``):

* an estimator pub is one job per parameter binding, holding all of the pub's
  observables: the state is prepared once and every observable's raw Pauli
  terms are reduced on the device, then combined per observable exactly as the
  reference's ``combine_expectation`` does (so ``evs`` are the reference's bits
  for the same raw terms); ``stds`` are 0 (the dense engine is exact);
* a sampler pub is one job per binding, pub j running under the j-th
  SplitMix64 output of the run's seed (``derive_seeds``, the rule of
  sampler.py:105-111 and run_batch), binding b of a pub under the b-th output of
  the pub's seed;
* parameter bindings are parameter-only updates of the pub's circuit: the
  server reuses the circuit's fusion plan across them (its plan cache is keyed
  by structure), which the reference's estimator cannot do.

Qiskit is not a dependency of this package: circuits are this package's flat
circuits (``workloads.Circuit``, ``workloads.FlatAnsatz``) and observables are
lists of ``(label, coefficient)`` with Qiskit's label order (the last character
acts on qubit 0), as a ``SparsePauliOp`` lists them. All jobs of a ``run`` are
submitted before any is waited for, so the server batches them.
"""
from __future__ import annotations

import secrets
from dataclasses import dataclass, field

import numpy as np

from . import hip as H
from . import workloads as W


def pauli_label_masks(label: str) -> tuple[int, int]:
    """(x_mask, z_mask) of a Pauli label in Qiskit order (label[-1] is qubit 0):
    X sets x, Z sets z, Y sets both (the raw term's i^popcount(x & z) phase
    makes it Y = iXZ), I neither."""
    x = z = 0
    for q, letter in enumerate(reversed(label.upper())):
        if letter not in "IXYZ":
            raise ValueError(f"not a Pauli label: {label!r}")
        if letter in "XY":
            x |= 1 << q
        if letter in "YZ":
            z |= 1 << q
    return x, z


def observable_terms(observable, num_qubits: int) -> list[tuple[int, int, complex]]:
    """An observable given as [(label, coefficient), ...] or {label: coefficient}
    as (x_mask, z_mask, coefficient) terms; a label longer than the circuit
    (beyond identities) is refused."""
    items = observable.items() if isinstance(observable, dict) else observable
    terms = []
    for label, coeff in items:
        if len(label) > num_qubits and set(label[:len(label) - num_qubits].upper()) - {"I"}:
            raise ValueError(f"label {label!r} acts beyond the circuit's {num_qubits} qubits")
        x, z = pauli_label_masks(label)
        terms.append((x, z, complex(coeff)))
    return terms


def bind(circuit, values):
    """The circuit with new parameter values: a FlatAnsatz takes its angle
    array; a Circuit takes one value per parameter of its parameterised
    operations, in operation order."""
    values = np.asarray(values, np.float64).ravel()
    if isinstance(circuit, W.FlatAnsatz):
        if values.size != circuit.num_angles:
            raise ValueError(f"{values.size} values for {circuit.num_angles} angles")
        return circuit.with_angles(values)
    slots = [(i, len(op[2])) for i, op in enumerate(circuit.ops) if op[0] != W.MEASURE and op[2]]
    need = sum(k for _, k in slots)
    if values.size != need:
        raise ValueError(f"{values.size} values for {need} parameters")
    new, at = {}, 0
    for i, k in slots:
        new[i] = values[at:at + k]
        at += k
    return circuit.with_params(new)


@dataclass
class EstimatorPubResult:
    evs: np.ndarray                      # (bindings, observables), or (observables,) without bindings
    stds: np.ndarray
    metadata: dict = field(default_factory=dict)


@dataclass
class SamplerPubResult:
    counts: list                         # one {clbit string as int: count} per binding
    metadata: dict = field(default_factory=dict)


def _bindings(circuit, parameter_values):
    if parameter_values is None:
        return [circuit], False
    values = np.asarray(parameter_values, np.float64)
    if values.ndim == 1:
        values = values[None, :]
    return [bind(circuit, v) for v in values], True


class HipEstimator:
    """Expectation values of pubs ``(circuit, observables[, parameter_values])``."""

    def __init__(self, *, server: H.Server | None = None, precision: int = H.C128, seed: int | None = None,
                 budget_bytes: int = 0):
        self._server = server
        self._own = server is None
        self._precision = precision
        self._seed = seed
        self._budget = budget_bytes

    def run(self, pubs) -> list[EstimatorPubResult]:
        seed = self._seed if self._seed is not None else secrets.randbits(63)
        server = self._server or H.Server(budget_bytes=self._budget)
        try:
            pending = []
            for pub in pubs:
                circuit, observables = pub[0], pub[1]
                values = pub[2] if len(pub) > 2 else None
                obs = [observable_terms(o, circuit.num_qubits) for o in observables]
                circuits, bound = _bindings(circuit, values)
                tickets = [server.submit(H.Job(H.JOB_ESTIMATOR, c, precision=self._precision, observables=obs))
                           for c in circuits]
                pending.append((tickets, len(obs), bound))
            out = []
            for tickets, nobs, bound in pending:
                evs = np.zeros((len(tickets), nobs))
                for b, t in enumerate(tickets):
                    r = server.wait(t)
                    if r.status != H.MH_OK:
                        raise H.HipError(r.status, r.error)
                    evs[b] = r.values.real
                evs = evs if bound else evs[0]
                out.append(EstimatorPubResult(evs=evs, stds=np.zeros_like(evs),
                                              metadata={"seed": seed, "precision": self._precision}))
            return out
        finally:
            if self._own:
                server.close()


class HipSampler:
    """Counts of pubs ``(circuit[, parameter_values[, shots]])``."""

    def __init__(self, *, server: H.Server | None = None, default_shots: int = 1024, seed: int | None = None,
                 budget_bytes: int = 0):
        self._server = server
        self._own = server is None
        self._shots = default_shots
        self._seed = seed
        self._budget = budget_bytes

    def run(self, pubs, *, shots: int | None = None) -> list[SamplerPubResult]:
        shots = self._shots if shots is None else shots
        seed = self._seed if self._seed is not None else secrets.randbits(63)
        pub_seeds = H.derive_seeds(seed, len(pubs))
        server = self._server or H.Server(budget_bytes=self._budget)
        try:
            pending = []
            for pub, pub_seed in zip(pubs, pub_seeds):
                circuit = pub[0]
                values = pub[1] if len(pub) > 1 else None
                pub_shots = pub[2] if len(pub) > 2 and pub[2] is not None else shots
                circuits, _ = _bindings(circuit, values)
                seeds = [pub_seed] if len(circuits) == 1 and values is None else H.derive_seeds(pub_seed, len(circuits))
                tickets = [server.submit(H.Job(H.JOB_SAMPLER, c, shots=pub_shots, seed=s))
                           for c, s in zip(circuits, seeds)]
                pending.append((tickets, pub_shots, seeds))
            out = []
            for tickets, pub_shots, seeds in pending:
                counts = []
                for t in tickets:
                    r = server.wait(t)
                    if r.status != H.MH_OK:
                        raise H.HipError(r.status, r.error)
                    counts.append(r.counts)
                out.append(SamplerPubResult(counts=counts,
                                            metadata={"shots": pub_shots, "seed": seed, "binding_seeds": seeds}))
            return out
        finally:
            if self._own:
                server.close()
