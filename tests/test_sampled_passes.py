"""The sampled per-pass check (sampled_passes.py) on the CPU: against a stand-in
device that applies each pass with the oracle on a whole state, it passes on
every pass; with a mistake planted in the stand-in (one block's matrix off by
one unit in the last place, or one block skipped) it names the pass."""
import numpy as np
import pytest

import sampled_passes as S
from menos_amd import hip as H
from menos_amd import workloads as W


class OracleDevice:
    """gather / apply_blocks over a host state, by the oracle."""

    def __init__(self, orc, n, plant=None):
        self.orc, self.n, self.plant = orc, n, plant
        self.state = orc.initialize(n)
        self.calls = 0

    def gather(self, idx):
        return self.state[np.asarray(idx, np.int64)]

    def apply_blocks(self, blocks):
        for i, (m, qs) in enumerate(blocks):
            if self.plant == ("skip", self.calls, i):
                continue
            if self.plant == ("ulp", self.calls, i):
                m = m.copy()
                m.real.view(np.uint64)[0, 0] ^= np.uint64(1)
            self.orc.apply_generic(self.state, m, qs)
        self.calls += 1


def _setup(orc, n=14, layers=5):
    c = W.random_circuit(n, layers, 77)
    blocks = S.circuit_blocks(orc, c, orc.plan_fusion(c))
    return c, blocks, H.plan_passes(n, blocks)


def test_check_passes_and_whole_state(orc):
    c, blocks, passes = _setup(orc)
    dev = OracleDevice(orc, 14)
    S.check_passes(orc, dev, blocks, passes, np.random.default_rng(1), 1 << 12)
    sv, _, _ = orc.execute(c)
    assert dev.state.tobytes() == sv.tobytes()
    assert len(passes) > 3 and passes[0][0] == 0 and passes[-1][1] == len(blocks)


@pytest.mark.parametrize("kind", ["ulp", "skip"])
def test_planted_mistakes_are_found(orc, kind):
    _, blocks, passes = _setup(orc)
    p = len(passes) // 2
    dev = OracleDevice(orc, 14, plant=(kind, p, 0))
    with pytest.raises(AssertionError, match=f"pass {p} "):
        # All groups of the state (2^14 amplitudes): a one-ulp change in one
        # matrix entry must show in some amplitude.
        S.check_passes(orc, dev, blocks, passes, np.random.default_rng(2), 1 << 14)


def test_group_indices_cover_distinct_groups():
    rng = np.random.default_rng(3)
    idx = S.group_indices(20, [1, 4, 9], 5, rng)
    assert idx.shape == (32, 8)
    # Within a row only the D bits vary; rows differ only outside D.
    d_mask = np.uint64((1 << 1) | (1 << 4) | (1 << 9))
    assert np.all((idx & ~d_mask) == (idx[:, :1] & ~d_mask))
    assert sorted(set((idx[0] & d_mask).tolist())) == sorted(set(S.deposit(np.arange(8, dtype=np.uint64), [1, 4, 9]).tolist()))
