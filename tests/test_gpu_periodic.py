"""Periodic chains (`pMPS`, src/Ansatz/Chain.jl:19,36-62) on the device against the
oracle: upload/download, the GHZ closed form, evolve on every bond 1..n-1 (the wrap
index rides through θ), expect/overlap/norm with the trace over the wrap bond, and
the open-only methods refusing periodic chains.  Unnormalised random states are
compared relative to their largest amplitude (1e-10)."""
import numpy as np
import pytest

from oracle import chain as oc
from oracle import statevector as sv
from oracle import workloads as wl

pytestmark = pytest.mark.gpu


def _random_pmps(rng, n, w, chi):
    dims = [w] + [chi] * (n - 1) + [w]
    return [(rng.standard_normal((2, dims[i], dims[i + 1])) + 1j * rng.standard_normal((2, dims[i], dims[i + 1])))
            / np.sqrt(2.0) for i in range(n)]  # reference order (o, l, r)


def _vec(psi):
    return oc.statevector(oc.Chain(psi.internal_arrays(), psi.lambdas(), periodic=True))


def test_pmps_roundtrip(qh):
    rng = np.random.default_rng(41)
    arrs = _random_pmps(rng, 5, 3, 4)
    psi = qh.pMPS(arrs)
    assert psi.periodic and psi.wrap_dim() == 3
    assert psi.bond_dims() == [4, 4, 4, 4]
    for a, b in zip(arrs, psi.arrays()):
        assert np.array_equal(a, b)


def test_pmps_ghz_closed_form(qh):
    a = np.zeros((2, 2, 2), dtype=np.complex128)  # (o, l, r): diag(1,0) for o=0, diag(0,1) for o=1
    a[0, 0, 0] = 1.0
    a[1, 1, 1] = 1.0
    n = 6
    psi = qh.pMPS([a] * n)
    assert abs(qh.norm(psi) - np.sqrt(2.0)) < 1e-14
    vals = qh.expect_many(psi, [[qh.gate(wl.Z, 2)], [qh.gate(wl.Z, 1), qh.gate(wl.Z, 5)],
                                [qh.gate(np.kron(wl.Z, wl.Z), 3, 4)]])
    assert np.allclose(vals, [0.0, 2.0, 2.0], atol=1e-14)
    assert abs(qh.expect(psi, [qh.gate(wl.X, 1), qh.gate(wl.X, 2), qh.gate(wl.X, 3), qh.gate(wl.X, 4),
                               qh.gate(wl.X, 5), qh.gate(wl.X, 6)]) - 2.0) < 1e-14


@pytest.mark.parametrize("maxdim", [None, 6])
def test_pmps_evolve_matches_oracle(qh, maxdim):
    rng = np.random.default_rng(42)
    n, w = 6, 3
    arrs = _random_pmps(rng, n, w, 3)
    psi = qh.pMPS(arrs)
    ref = oc.pMPS(arrs)
    v = oc.statevector(ref)
    for layer in range(3):
        for b in wl.brickwork_bonds(n, layer):
            g = wl.haar_gate4(rng)
            qh.evolve_(psi, qh.gate(g, b, b + 1), maxdim=maxdim)
            oc.evolve(ref, oc.gate2(g, b), maxdim=maxdim)
            v = sv.apply2(v, g, b)
    assert psi.bond_dims() == ref.bond_dims()
    vr = oc.statevector(ref)
    scale = np.max(np.abs(vr))
    assert np.max(np.abs(_vec(psi) - vr)) < 1e-10 * scale
    if maxdim is None:
        assert np.max(np.abs(vr - v)) < 1e-10 * scale
    # expectation values and norms through the wrap-bond trace
    z = qh.expect_many(psi, [[qh.gate(wl.Z, i)] for i in range(1, n + 1)])
    zr = np.array([oc.expect(ref, [oc.gate1(wl.Z, i)]) for i in range(1, n + 1)])
    assert np.max(np.abs(z - zr)) < 1e-10 * np.max(np.abs(zr))
    nr = oc.norm(ref)
    assert abs(qh.norm(psi) - nr) < 1e-10 * nr


def test_pmps_overlap_and_batched_layer(qh):
    rng = np.random.default_rng(43)
    n = 6
    arrs_a = _random_pmps(rng, n, 2, 4)
    arrs_b = _random_pmps(rng, n, 3, 2)
    a, b = qh.pMPS(arrs_a), qh.pMPS(arrs_b)
    ref = oc.overlap(oc.pMPS(arrs_a), oc.pMPS(arrs_b))
    assert abs(qh.overlap(a, b) - ref) < 1e-10 * abs(ref)
    c = a.copy()
    gates = [(bd, wl.haar_gate4(rng)) for bd in wl.brickwork_bonds(n, 1)]
    qh.evolve_layer_([(a, qh.gate(g, bd, bd + 1)) for bd, g in gates])
    for bd, g in gates:
        qh.evolve_(c, qh.gate(g, bd, bd + 1))
    for x, y in zip(a.internal_arrays(), c.internal_arrays()):
        assert np.array_equal(x, y)


def test_pmps_open_only_methods(qh):
    rng = np.random.default_rng(44)
    psi = qh.pMPS(_random_pmps(rng, 5, 2, 2))
    for fn in (lambda: qh.canonize_(psi), lambda: qh.mixed_canonize_(psi, 2),
               lambda: qh.canonize_site_(psi, 2, direction="right"), lambda: qh.normalize_(psi, 2)):
        with pytest.raises(qh.ArgumentError):
            fn()
    op = qh.MPS([np.ones((2, 2)) / 2, np.ones((2, 2, 2)) / 2, np.ones((2, 2, 2)) / 2, np.ones((2, 2, 2)) / 2,
                 np.ones((2, 2)) / 2])
    with pytest.raises(qh.ArgumentError):
        qh.overlap(psi, op)


def test_pmpo_apply_to_pmps(qh):
    from oracle import mpo as om
    from qrochet_hip import mpo as qm

    rng = np.random.default_rng(45)
    n, D = 5, 3
    arrs = _random_pmps(rng, n, 2, 2)
    marr = [(rng.standard_normal((2, 2, D, D)) + 1j * rng.standard_normal((2, 2, D, D))) for _ in range(n)]
    out = qm.apply(qm.pMPO(marr), qh.pMPS(arrs))
    assert out.periodic and out.wrap_dim() == 2 * D
    exp = oc.statevector(om.apply_exact(om.from_arrays_periodic(marr), oc.pMPS(arrs)))
    assert np.max(np.abs(_vec(out) - exp)) < 1e-10 * np.max(np.abs(exp))
    with pytest.raises(qh.ArgumentError):
        qm.apply(qm.pMPO(marr), qh.MPS([np.ones((2, 2))] + [np.ones((2, 2, 2))] * (n - 2) + [np.ones((2, 2))]))
