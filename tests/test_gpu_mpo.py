"""MPO × MPS on the device (qh_mpo_apply, qh_mpo_fit) against the oracle
(oracle/mpo.py).  FP64 path: 1e-10 absolute on state vectors and overlaps.
Reduced-precision environment contractions (FP32, bf16x3 on MFMA): relative
errors stated and bounded by 1e-5 (FP32) and 1e-3 (bf16x3)."""
import numpy as np
import pytest

from oracle import chain as oc
from oracle import mpo as om
from oracle import workloads as wl

pytestmark = pytest.mark.gpu

BOUND = {"fp32": 1e-5, "bf16x3": 1e-3}


def _setup(rng, n, chi, D, canon=False):
    arrays = wl.left_canonize_normalize(wl.random_mps_arrays(rng, n, 2, wl.max_profile(n, 2, chi)))
    ref = oc.Chain(arrays)
    if canon:
        ref = oc.canonize_copy(ref)
    marr = om.random_hermitian_mpo(rng, n, 2, D)
    return ref, marr


def _gpu_state(qh, ref):
    return qh.from_internal([a.copy() for a in ref.arrays], lambdas=[None if l is None else l.copy() for l in ref.lam])


def _vec(psi):
    return oc.statevector(oc.Chain(psi.internal_arrays(), psi.lambdas()))


@pytest.mark.parametrize("canon", [False, True])
def test_mpo_apply_exact(qh, canon):
    from qrochet_hip import mpo as qm

    rng = np.random.default_rng(31)
    ref, marr = _setup(rng, 6, 4, 3, canon)
    ws = om.from_arrays(marr)
    h = qm.MPO(marr)
    out = qm.apply(h, _gpu_state(qh, ref))
    exp = om.apply_exact(ws, ref)
    assert out.bond_dims() == exp.bond_dims()
    assert np.max(np.abs(_vec(out) - oc.statevector(exp))) < 1e-10


def test_mpo_fit_fp64_matches_oracle(qh):
    from qrochet_hip import mpo as qm

    rng = np.random.default_rng(32)
    n = 8
    ref, marr = _setup(rng, n, 4, 3)
    ws = om.from_arrays(marr)
    phi_ref, ov_ref = om.fit(ws, ref, ref.copy(), sweeps=2)
    psi = _gpu_state(qh, ref)
    phi, ov = qm.compress(qm.MPO(marr), psi, sweeps=2)
    assert phi.bond_dims() == phi_ref.bond_dims()
    assert abs(ov - ov_ref[-1]) < 1e-10
    assert np.max(np.abs(_vec(phi) - oc.statevector(phi_ref))) < 1e-10
    assert all(l is None for l in phi.lambdas())


def test_mpo_fit_with_lambdas_and_other_order(qh):
    """φ initial guess with Λ vectors (absorbed) and MPO arrays given in another order."""
    from qrochet_hip import mpo as qm

    rng = np.random.default_rng(33)
    n = 6
    ref, marr = _setup(rng, n, 4, 2, canon=True)
    ws = om.from_arrays(marr)
    order = ("l", "i", "r", "o")
    phi_ref, ov_ref = om.fit(ws, ref, ref.copy(), sweeps=2)
    psi = _gpu_state(qh, ref)
    phi = psi.copy()
    ov = qm.fit_(qm.MPO(om.to_arrays(ws, order), order=order), psi, phi, sweeps=2)
    assert abs(ov - ov_ref[-1]) < 1e-10
    assert np.max(np.abs(_vec(phi) - oc.statevector(phi_ref))) < 1e-10


@pytest.mark.parametrize("precision", ["fp32", "bf16x3"])
def test_mpo_fit_reduced_precision(qh, precision):
    from qrochet_hip import mpo as qm

    rng = np.random.default_rng(34)
    n = 10
    ref, marr = _setup(rng, n, 8, 4)
    ws = om.from_arrays(marr)
    phi_ref, ov_ref = om.fit(ws, ref, ref.copy(), sweeps=2)
    phi, ov = qm.compress(qm.MPO(marr), _gpu_state(qh, ref), sweeps=2, precision=precision)
    vr = oc.statevector(phi_ref)
    rel_state = np.linalg.norm(_vec(phi) - vr) / np.linalg.norm(vr)
    rel_ov = abs(ov - ov_ref[-1]) / abs(ov_ref[-1])
    print(f"{precision}: relative state error {rel_state:.3e}, relative overlap error {rel_ov:.3e}")
    assert rel_state < BOUND[precision]
    assert rel_ov < BOUND[precision]
    assert rel_state > 0.0


def test_mpo_errors(qh):
    from qrochet_hip import mpo as qm

    rng = np.random.default_rng(35)
    ref, marr = _setup(rng, 5, 4, 2)
    h = qm.MPO(marr)
    psi = _gpu_state(qh, ref)
    with pytest.raises(qh.ArgumentError):
        qm.fit_(h, psi, psi)
    other, _ = _setup(rng, 6, 4, 2)
    with pytest.raises(qh.DimensionMismatch):
        qm.apply(h, _gpu_state(qh, other))
    with pytest.raises(qh.ArgumentError):
        qm.MPO(marr[:1])
