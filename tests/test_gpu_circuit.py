"""A gate-level circuit (named gates, reversed lanes, one- and two-lane gates)
run on a device MPS through the layer schedule equals the dense state vector."""
import numpy as np
import pytest

from oracle import chain as oc
from oracle import statevector as sv

pytestmark = pytest.mark.gpu


def test_random_circuit_matches_statevector(qh):
    from qrochet_hip import circuit as qc
    from qrochet_hip import serving

    rng = np.random.default_rng(51)
    n = 8
    names1 = ["H", "X", "S", "T", "Rx", "Ry", "Rz"]
    names2 = ["CX", "CZ", "iSwap", "Rzz", "Swap"]
    circ = []
    for _ in range(60):
        if rng.random() < 0.4:
            nm = names1[rng.integers(len(names1))]
            circ.append(qc.Op(nm, (int(rng.integers(1, n + 1)),), (float(rng.random()),) if nm[0] == "R" else ()))
        else:
            a = int(rng.integers(1, n))
            lanes = (a, a + 1) if rng.random() < 0.5 else (a + 1, a)
            nm = names2[rng.integers(len(names2))]
            circ.append(qc.Op(nm, lanes, (float(rng.random()),) if nm == "Rzz" else ()))
    psi = serving.product_state(n)
    qc.run_(psi, circ)
    v = np.zeros((2,) * n, dtype=np.complex128)
    v[(0,) * n] = 1.0
    for op in circ:
        if len(op.lanes) == 1:
            v = sv.apply1(v, op.array(), op.lanes[0])
        else:
            s, m = qc._oriented(op)
            v = sv.apply2(v, m.reshape((2, 2, 2, 2), order="F"), s)
    got = oc.statevector(oc.Chain(psi.internal_arrays(), psi.lambdas()))
    assert np.max(np.abs(got - v)) < 1e-10
