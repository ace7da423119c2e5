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


def test_yao_circuit_matches_wire_network(qh):
    """A Yao-style circuit (puts, swaps, nested chains; ext/QrochetYaoExt.jl:15-48) run
    on a device MPS through `from_yao` equals the reference's wire construction
    evaluated on the dense state (tests/test_circuit.py: _wire_network)."""
    from qrochet_hip import circuit as qc
    from qrochet_hip import serving
    from oracle import workloads as wl
    from tests.test_circuit import _wire_network

    rng = np.random.default_rng(52)
    n = 7
    blocks = []
    for _ in range(40):
        r = rng.random()
        if r < 0.2:
            a = int(rng.integers(1, n + 1))
            b = int(rng.integers(1, n + 1))
            if a != b:
                blocks.append(qc.Swap(a, b))
        elif r < 0.5:
            blocks.append(qc.Put((int(rng.integers(1, n + 1)),), wl.haar_unitary(rng, 2)))
        else:
            a = int(rng.integers(1, n))
            locs = (a, a + 1) if rng.random() < 0.5 else (a + 1, a)
            blocks.append([qc.Put(locs, wl.haar_unitary(rng, 4))])
    # a two-lane block whose wires do not end on neighbouring sites cannot be an MPS
    # update: make those blocks one-lane (the wire map depends on the swaps only)
    ops = qc.from_yao(blocks, n)
    k = 0
    flat = qc.flatten(blocks)
    for i, b in enumerate(flat):
        if isinstance(b, qc.Put):
            if len(b.locs) == 2 and abs(ops[k].lanes[0] - ops[k].lanes[1]) != 1:
                flat[i] = qc.Put((b.locs[0],), wl.haar_unitary(rng, 2))
            k += 1
    ops = qc.from_yao(flat, n)
    assert sum(isinstance(b, qc.Swap) for b in flat) >= 4
    assert sum(len(o.lanes) == 2 for o in ops) >= 6
    psi = serving.product_state(n)
    qc.run_(psi, ops)
    v0 = np.zeros((2,) * n, dtype=np.complex128)
    v0[(0,) * n] = 1.0
    want = _wire_network(flat, n, v0)
    got = oc.statevector(oc.Chain(psi.internal_arrays(), psi.lambdas()))
    assert np.max(np.abs(got - want)) < 1e-10
