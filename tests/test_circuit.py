"""Circuit ingestion (qrochet_hip.circuit) on the CPU: gate matrices in the
reference's lane convention, the layer schedule, and a circuit applied to a state
vector with the same conventions as the oracle's gates."""
import numpy as np

from oracle import statevector as sv
from qrochet_hip import circuit as qc


def _apply_dense(v, op):
    m = op.array()
    if len(op.lanes) == 1:
        return sv.apply1(v, m, op.lanes[0])
    s, mm = qc._oriented(op)
    return sv.apply2(v, mm.reshape((2, 2, 2, 2), order="F"), s)


def test_cx_convention_first_lane_controls():
    v = np.zeros((2, 2), dtype=np.complex128)
    v[1, 0] = 1.0  # |site1=1, site2=0>
    out = _apply_dense(v, qc.Op("CX", (1, 2)))
    assert out[1, 1] == 1.0 and abs(out).sum() == 1.0
    v2 = np.zeros((2, 2), dtype=np.complex128)
    v2[0, 1] = 1.0  # control (site 1) is 0: nothing happens
    assert np.array_equal(_apply_dense(v2, qc.Op("CX", (1, 2))), v2)
    # reversed lanes: control on site 2
    out3 = _apply_dense(v2, qc.Op("CX", (2, 1)))
    assert out3[1, 1] == 1.0


def test_gates_are_unitary():
    for name, f in qc.GATES_1.items():
        args = (0.3,) * f.__code__.co_argcount
        m = np.asarray(f(*args))
        assert np.allclose(m.conj().T @ m, np.eye(2)), name
    for name, f in qc.GATES_2.items():
        args = (0.7,) * f.__code__.co_argcount
        m = np.asarray(f(*args))
        assert np.allclose(m.conj().T @ m, np.eye(4)), name


def test_layer_schedule_preserves_lane_order():
    c = [qc.Op("H", (1,)), qc.Op("CX", (1, 2)), qc.Op("CX", (3, 4)), qc.Op("CX", (2, 3)), qc.Op("X", (1,)),
         qc.Op("Rz", (4,), (0.1,))]
    ls = qc.layers(c)
    pos = {}
    for k, l in enumerate(ls):
        lanes = [x for op in l for x in op.lanes]
        assert len(lanes) == len(set(lanes))
        for op in l:
            pos[id(op)] = k
    for i, a in enumerate(c):
        for b in c[i + 1:]:
            if set(a.lanes) & set(b.lanes):
                assert pos[id(a)] < pos[id(b)]


def test_ghz_circuit_statevector():
    n = 5
    c = [qc.Op("H", (1,))] + [qc.Op("CX", (i, i + 1)) for i in range(1, n)]
    v = np.zeros((2,) * n, dtype=np.complex128)
    v[(0,) * n] = 1.0
    for layer in qc.layers(c):
        for op in layer:
            v = _apply_dense(v, op)
    exp = np.zeros((2,) * n, dtype=np.complex128)
    exp[(0,) * n] = exp[(1,) * n] = 1 / np.sqrt(2)
    assert np.allclose(v, exp)
