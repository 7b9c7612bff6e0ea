"""The HBM-scale random circuit (33 qubits, complex128, 128 GiB; 16 layers of U3
on every qubit then CZ on the brick pattern, the reference's containment plan)
checked on the MI355X where no oracle can hold the state: pass by pass, on
sampled groups, bit for bit against the oracle's generic kernel (see
sampled_passes.py), then the norm within 1e-12. A 22-qubit run takes the same
path and is also compared whole with the oracle, which checks the sampling
itself."""
import numpy as np
import pytest

import sampled_passes as S
from menos_amd import workloads as W

pytestmark = pytest.mark.gpu


def _run(orc, hip, n, layers, seed, max_amplitudes, log=None):
    c = W.random_circuit(n, layers, seed)
    plan = orc.plan_fusion(c)
    blocks = S.circuit_blocks(orc, c, plan)
    passes = hip.plan_passes(n, blocks)
    est = hip.HipDenseEngine().estimate(c)
    assert est.passes == len(passes), (est.passes, len(passes))
    d = hip.DeviceState(n)
    d.init()
    rng = np.random.default_rng(seed)
    checked = S.check_passes(orc, d, blocks, passes, rng, max_amplitudes, on_pass=log)
    return c, d, passes, checked


def test_sampled_passes_match_the_whole_state_at_22_qubits(orc, hip):
    c, d, passes, checked = _run(orc, hip, 22, 6, 2025, 1 << 16)
    sv, _, _ = orc.execute(c)
    got = d.download()
    d.free()
    assert got.tobytes() == sv.tobytes()
    assert checked >= len(passes) * (1 << 12)


def test_33_qubits_complex128_every_pass_sampled(orc, hip):
    def log(p, count):
        if p % 10 == 0:
            print(f"33q pass {p}: {count} amplitudes checked", flush=True)

    c, d, passes, checked = _run(orc, hip, 33, 16, 2025, 1 << 21, log)
    norm = d.expectation([(0, 0)])[0]
    d.free()
    print(f"33q: {len(passes)} passes, {checked} amplitudes checked bit for bit, norm {norm!r}", flush=True)
    assert len(passes) >= 60
    assert abs(norm - 1.0) <= 1e-12, norm
