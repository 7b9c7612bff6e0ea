"""Host logic of the primitives (menos_amd/primitives.py): Pauli labels in
Qiskit order to (x, z) masks, observables to terms, parameter binding."""
import numpy as np
import pytest

from menos_amd import primitives as P
from menos_amd import workloads as W


def test_labels_in_qiskit_order():
    # label[-1] acts on qubit 0: "XYZI" is I on 0, Z on 1, Y on 2, X on 3.
    assert P.pauli_label_masks("XYZI") == ((1 << 3) | (1 << 2), (1 << 2) | (1 << 1))
    assert P.pauli_label_masks("I") == (0, 0)
    assert P.pauli_label_masks("Y") == (1, 1)
    assert P.pauli_label_masks("zx") == (1, 2)
    with pytest.raises(ValueError):
        P.pauli_label_masks("XQ")


def test_observables_to_terms():
    terms = P.observable_terms({"ZZ": 0.5, "XI": -1j}, 2)
    assert terms == [(0, 3, 0.5 + 0j), (2, 0, -1j)]
    assert P.observable_terms([("IIZ", 2.0)], 2) == [(0, 1, 2.0 + 0j)]
    with pytest.raises(ValueError):
        P.observable_terms([("XZZ", 1.0)], 2)


def test_binding():
    a = W.FlatAnsatz(5, 2, np.zeros(20))
    b = P.bind(a, np.arange(20.0))
    assert isinstance(b, W.FlatAnsatz) and np.array_equal(b.flat()[4], np.arange(20.0))
    with pytest.raises(ValueError):
        P.bind(a, np.zeros(19))
    c = W.random_circuit(3, 2, 1).measure_all()
    values = np.arange(18.0)
    bc = P.bind(c, values)
    got = [p for op, _, params in bc.ops if op != W.MEASURE for p in params]
    assert got == list(values)
    assert [op[:2] for op in bc.ops] == [op[:2] for op in c.ops]
    with pytest.raises(ValueError):
        P.bind(c, np.zeros(17))
