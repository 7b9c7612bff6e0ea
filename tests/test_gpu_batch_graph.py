"""The resident, graph-captured batch (mh_batch_graph_*) against mh_batch_execute:
terms bit for bit on every replay (with other device work between replays), after
parameter updates, at both precisions and under the cache-group schedule; and the
refusals: a structure change, or new parameters that change a pass program."""
import numpy as np
import pytest

from menos_amd import workloads as W

pytestmark = pytest.mark.gpu


def _workload(count, seed, qubits=(10, 15)):
    return W.batch_workload(count, seed, qubits=qubits)


def _same(a, b):
    return all(x.tobytes() == y.tobytes() for x, y in zip(a, b))


@pytest.mark.parametrize("precision,fusion,schedule", [
    (1, 3, {}),                                                  # complex64, pair-keeping fusion
    (0, 1, {}),                                                  # complex128, the reference's containment plan
    (1, 3, {"group_bytes": 1 << 20, "kernel_streams": 4, "cached_states": True}),
    (0, 1, {"group_bytes": 2 << 20, "kernel_streams": 3, "cached_states": True}),
])
def test_replays_equal_the_batch_entry_point(hip, precision, fusion, schedule):
    circuits, terms = _workload(96, 11 + precision)
    want = hip.batch_execute(circuits, terms, precision=precision, fusion=fusion).expectation
    d = hip.DeviceState(12, precision)
    d.init()
    with hip.BatchGraph(circuits, terms, precision=precision, fusion=fusion, **schedule) as g:
        for _ in range(4):
            got, res = g.run()
            assert _same(got, want)
            assert res.launches > 2
            # Eager work between replays (its own kernels and synchronisations).
            d.apply_matrix(np.eye(2, dtype=np.complex128), [3])
            d.expectation([(1, 2), (0, 5)])
    d.free()


def test_parameter_updates(hip):
    circuits, terms = _workload(64, 21)
    rng = np.random.default_rng(5)
    with hip.BatchGraph(circuits, terms, precision=1, fusion=3) as g:
        first, _ = g.run()
        for step in range(3):
            updated = [c.with_angles(rng.uniform(-np.pi, np.pi, c.num_angles)) for c in circuits]
            g.update(updated)
            got, res = g.run()
            want = hip.batch_execute(updated, terms, precision=1, fusion=3).expectation
            assert _same(got, want), step
            assert not _same(got, first)
            assert res.compile_seconds > 0
        # The same through one parameter array.
        sizes = [c.num_angles for c in circuits]
        assert g.param_count() == sum(sizes)
        params = rng.uniform(-np.pi, np.pi, sum(sizes))
        g.set_params(params)
        got, _ = g.run()
        cuts = np.cumsum([0] + sizes)
        updated = [c.with_angles(params[cuts[i]:cuts[i + 1]]) for i, c in enumerate(circuits)]
        assert _same(got, hip.batch_execute(updated, terms, precision=1, fusion=3).expectation)


def test_refusals_keep_the_graph(hip):
    circuits, terms = _workload(16, 31)
    with hip.BatchGraph(circuits, terms, precision=0, fusion=1) as g:
        before, _ = g.run()
        # A structure change: one layer fewer.
        changed = list(circuits)
        c0 = changed[0]
        changed[0] = W.FlatAnsatz(c0.num_qubits, c0.depth - 1, np.zeros(2 * c0.num_qubits * (c0.depth - 1)))
        with pytest.raises(hip.HipError, match="structure"):
            g.update(changed)
        # Angles that make a rotation the identity: its block becomes diagonal,
        # so the pass program would change.
        zeroed = [c.with_angles(np.zeros(c.num_angles)) for c in circuits]
        with pytest.raises(hip.HipError, match="pass program"):
            g.update(zeroed)
        after, _ = g.run()
        assert _same(before, after)


@pytest.mark.parametrize("schedule", [{}, {"group_bytes": 1 << 20, "kernel_streams": 4}])
def test_closing_returns_the_device_memory(hip, schedule):
    """A replayed graph's states, tables and workspaces go back to the device
    when it is closed (the pool trimmed): free memory returns to within 64 MiB
    of what it was before the graph."""
    circuits, terms = _workload(256, 41, qubits=(16, 20))
    hip.lib().mh_pool_trim()
    before = hip.device_info().free_bytes
    with hip.BatchGraph(circuits, terms, precision=1, fusion=3, **schedule) as g:
        for _ in range(2):
            g.run()
    hip.lib().mh_pool_trim()
    assert hip.device_info().free_bytes >= before - (64 << 20)
