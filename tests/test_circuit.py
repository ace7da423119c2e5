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


# ---------------------------------------------------------------------------
# Yao circuits (ext/QrochetYaoExt.jl:15-48)
# ---------------------------------------------------------------------------
def _apply_on(v, m, axes):
    """Dense k-site operator on arbitrary axes (first axis = least significant digit)."""
    k = len(axes)
    g = np.asarray(m).reshape((2,) * (2 * k), order="F")  # (out_1..out_k, in_1..in_k)
    v = np.moveaxis(v, axes, range(k))
    v = np.tensordot(g, v, axes=(list(range(k, 2 * k)), list(range(k))))
    return np.moveaxis(v, range(k), axes)


def _wire_network(blocks, n, psi_in):
    """The reference's construction restated on a state vector: a Swap exchanges the
    wires of two lanes, every other block is a tensor on the wires at its locations;
    input site l feeds the wire that ends at lane l, output site l is that wire's end."""
    flat = qc.flatten(blocks)
    cur = list(range(n))  # lane -> wire
    tensors = []
    for b in flat:
        if isinstance(b, qc.Swap):
            cur[b.a - 1], cur[b.b - 1] = cur[b.b - 1], cur[b.a - 1]
        else:
            tensors.append((b.matrix, [cur[l - 1] for l in b.locs]))
    # state indexed by wire: wire cur[l] starts at input site l
    x = np.moveaxis(psi_in, list(range(n)), [cur[l] for l in range(n)])
    for m, wires in tensors:
        x = _apply_on(x, m, wires)
    return np.moveaxis(x, [cur[l] for l in range(n)], list(range(n)))


def test_yao_swap_relabels_wires():
    """X on lane 1 followed by swap(1, 2): the X tensor sits on the wire that ends at
    lane 2, so it acts on site 2 (a relabeling, not a SWAP applied to the state)."""
    ops = qc.from_yao([qc.Put((1,), qc._X), qc.Swap(1, 2)], 2)
    assert len(ops) == 1 and ops[0].lanes == (2,)
    v = np.zeros((2, 2), dtype=np.complex128)
    v[0, 0] = 1.0
    got = _apply_on(v, ops[0].array(), [1])
    assert np.array_equal(got, _wire_network([qc.Put((1,), qc._X), qc.Swap(1, 2)], 2, v))
    assert got[0, 1] == 1.0


def test_yao_circuit_matches_wire_network():
    rng = np.random.default_rng(8)
    n = 5
    blocks = [qc.Put((1,), wl_haar(rng, 2)), [qc.Put((2, 3), wl_haar(rng, 4)), qc.Swap(2, 4)],
              qc.Put((4, 5), wl_haar(rng, 4)), [[qc.Swap(1, 5)], qc.Put((5, 1), wl_haar(rng, 4))],
              qc.Put((3,), wl_haar(rng, 2)), qc.Swap(3, 4), qc.Put((4, 2), wl_haar(rng, 4))]
    ops = qc.from_yao(blocks, n)
    psi = rng.standard_normal((2,) * n) + 1j * rng.standard_normal((2,) * n)
    v = psi.copy()
    for op in ops:
        v = _apply_on(v, op.array(), [l - 1 for l in op.lanes])
    want = _wire_network(blocks, n, psi)
    assert np.max(np.abs(v - want)) < 1e-12
    # the same gates without the swaps differ (the swaps are not ignored)
    plain = psi.copy()
    for b in qc.flatten(blocks):
        if isinstance(b, qc.Put):
            plain = _apply_on(plain, b.matrix, [l - 1 for l in b.locs])
    assert np.max(np.abs(plain - want)) > 1e-3


def test_yao_rejects_other_blocks():
    import pytest

    with pytest.raises(qc.ArgumentError):
        qc.from_yao([qc.Put((1,), np.eye(3))], 2)
    with pytest.raises(qc.ArgumentError):
        qc.from_yao([object()], 2)


def wl_haar(rng, k):
    from oracle import workloads as wl

    return wl.haar_unitary(rng, k)
