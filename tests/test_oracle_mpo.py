"""Pinning of the MPO oracle (oracle/mpo.py) by closed forms: dense H·ψ, exact
fitting when φ's bonds can hold H|ψ⟩, the identity MPO, the variational
identity ‖Hψ − φ‖² = ‖Hψ‖² − ⟨φ|Hψ⟩, and monotone overlaps."""
import numpy as np
import pytest

from oracle import chain as oc
from oracle import mpo as om
from oracle import workloads as wl


def _psi(rng, n, chi, canon=False):
    arrays = wl.left_canonize_normalize(wl.random_mps_arrays(rng, n, 2, wl.max_profile(n, 2, chi)))
    psi = oc.Chain(arrays)
    return oc.canonize_copy(psi) if canon else psi


def test_layout_roundtrip_and_orders():
    rng = np.random.default_rng(1)
    arrs = om.random_hermitian_mpo(rng, 4, 2, 3)
    ws = om.from_arrays(arrs)
    assert [w.shape for w in ws] == [(1, 2, 2, 3), (3, 2, 2, 3), (3, 2, 2, 3), (3, 2, 2, 1)]
    back = om.to_arrays(ws)
    for a, b in zip(arrs, back):
        assert np.array_equal(a, b)
    order = ("l", "r", "i", "o")
    perm = om.to_arrays(ws, order)
    for a, b in zip(ws, om.from_arrays(perm, order)):
        assert np.array_equal(a, b)


def test_random_mpo_is_hermitian():
    rng = np.random.default_rng(2)
    n = 4
    ws = om.from_arrays(om.random_hermitian_mpo(rng, n, 2, 3))
    h = om.dense(ws).reshape(2**n, 2**n)
    assert np.allclose(h, h.conj().T, atol=1e-14)
    assert np.linalg.norm(h) > 0.1


def test_dense_of_product_mpo_is_kron():
    # D = 1: H = Z ⊗ X ⊗ I ⊗ Z (site 1 on the most significant axis of the tensor)
    ops = [wl.Z, wl.X, wl.I2, wl.Z]
    ws = [op.reshape(1, 2, 2, 1) for op in ops]
    h = om.dense(ws).reshape(16, 16)
    k = ops[0]
    for op in ops[1:]:
        k = np.kron(k, op)
    assert np.allclose(h, k)


@pytest.mark.parametrize("canon", [False, True])
def test_apply_exact_matches_dense(canon):
    rng = np.random.default_rng(3)
    n = 5
    psi = _psi(rng, n, 4, canon)
    ws = om.from_arrays(om.random_hermitian_mpo(rng, n, 2, 3))
    hpsi = om.apply_exact(ws, psi)
    assert hpsi.bond_dims() == [b * 3 for b in psi.bond_dims()]
    v = oc.statevector(psi)
    assert np.max(np.abs(oc.statevector(hpsi) - om.apply_dense(ws, v))) < 1e-12


def test_fit_is_exact_when_bonds_suffice():
    rng = np.random.default_rng(4)
    n = 6
    psi = _psi(rng, n, 2)
    ws = om.from_arrays(om.random_hermitian_mpo(rng, n, 2, 2))
    phi0 = oc.Chain(wl.random_mps_arrays(rng, n, 2, wl.max_profile(n, 2, 64)))
    phi, ov = om.fit(ws, psi, phi0, sweeps=1)
    hv = om.apply_dense(ws, oc.statevector(psi))
    assert np.max(np.abs(oc.statevector(phi) - hv)) < 1e-10
    assert abs(ov[-1] - np.vdot(hv, hv)) < 1e-10


def test_fit_identity_mpo_returns_psi():
    rng = np.random.default_rng(5)
    n = 7
    psi = _psi(rng, n, 4, canon=True)
    ws = [np.eye(2, dtype=np.complex128).reshape(1, 2, 2, 1) for _ in range(n)]
    phi, ov = om.fit(ws, psi, psi.copy(), sweeps=1)
    v = oc.statevector(psi)
    assert np.max(np.abs(oc.statevector(phi) - v)) < 1e-12
    assert abs(ov[-1] - np.vdot(v, v)) < 1e-12


def test_fit_is_variational():
    rng = np.random.default_rng(6)
    n = 8
    psi = _psi(rng, n, 4)
    ws = om.from_arrays(om.random_hermitian_mpo(rng, n, 2, 3))
    phi, ov = om.fit(ws, psi, psi.copy(), sweeps=3)
    hv = om.apply_dense(ws, oc.statevector(psi))
    f = oc.statevector(phi)
    # φ is the projection of Hψ on its own tangent space: ‖Hψ − φ‖² = ‖Hψ‖² − ⟨φ|Hψ⟩
    assert abs(np.vdot(hv - f, hv - f) - (np.vdot(hv, hv) - ov[-1])) < 1e-10
    assert abs(np.vdot(f, hv) - ov[-1]) < 1e-10
    assert all(abs(o.imag) < 1e-12 for o in ov)
    assert all(ov[k + 1].real >= ov[k].real - 1e-12 for k in range(len(ov) - 1))
    # compression loses something here (χ = 4 cannot hold H|ψ⟩ exactly)
    assert np.vdot(hv, hv).real - ov[-1].real > 1e-6


def test_local_z_matches_expect():
    rng = np.random.default_rng(7)
    n = 6
    psi = _psi(rng, n, 4, canon=True)
    z = om.local_z(psi)
    nrm = oc.norm(psi) ** 2
    ref = np.array([oc.expect(psi, [oc.gate1(wl.Z, i)]) for i in range(1, n + 1)]) / nrm
    assert np.max(np.abs(z - ref)) < 1e-12


def test_periodic_mpo_apply_matches_dense():
    rng = np.random.default_rng(8)
    n, w, D = 5, 2, 3
    arrs = [(rng.standard_normal((2, w if i == 0 else 2, 2 if i < n - 1 else w))
             + 1j * rng.standard_normal((2, w if i == 0 else 2, 2 if i < n - 1 else w))) for i in range(n)]
    psi = oc.pMPS(arrs)
    marr = [(rng.standard_normal((2, 2, D, D)) + 1j * rng.standard_normal((2, 2, D, D))) for _ in range(n)]
    ws = om.from_arrays_periodic(marr)
    hpsi = om.apply_exact(ws, psi)
    assert hpsi.periodic
    v = oc.statevector(psi)
    exp = om.apply_dense(ws, v)
    assert np.max(np.abs(oc.statevector(hpsi) - exp)) < 1e-12 * np.max(np.abs(exp))
    # the dense periodic MPO of a product of single-site operators with D = 1 is their kron
    ops = [wl.Z, wl.X, wl.Z, wl.I2, wl.X]
    h = om.dense([op.reshape(1, 2, 2, 1) for op in ops]).reshape(32, 32)
    k = ops[0]
    for op in ops[1:]:
        k = np.kron(k, op)
    assert np.allclose(h, k)
