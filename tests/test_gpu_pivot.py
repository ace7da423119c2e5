"""Column pivoting of the SVD's first QR pass (`qh_ctx_set_pivoting`, on by default):
X P = Q1 R1 with the columns ordered by decreasing norm, so the right factor of X is
P times the one the passes and the Jacobi produce.  Checked against the oracle with
1, 2 and 3 QR passes (the permutation lands on either factor of the Jacobi as the
pass count is odd or even), on untruncated evolution (every θ of 32..128 columns goes
through the preconditioner; tall and wide θ both occur) and on a truncated batched
layer, with pivoting on and off; and the automatic block width 32 on a θ of 2048
columns against LAPACK."""
import numpy as np
import pytest

from oracle import chain as oc
from oracle import statevector as sv
from oracle import workloads as wl

pytestmark = pytest.mark.gpu


def _ctx(qh, pivot, passes):
    ctx = qh.Context()
    ctx.set_pivoting(pivot)
    ctx.set_qr_passes(passes)
    return ctx


@pytest.mark.parametrize("passes", [1, 2, 3])
@pytest.mark.parametrize("pivot", [True, False])
def test_untruncated_evolution_matches_statevector(qh, pivot, passes):
    rng = np.random.default_rng(90 + passes)
    n = 14
    arrays = wl.left_canonize_normalize(wl.random_mps_arrays(rng, n, 2, wl.max_profile(n, 2, 128)))
    v = oc.statevector(oc.Chain([a.copy() for a in arrays]))
    layers = [[(b, wl.haar_gate4(rng)) for b in wl.brickwork_bonds(n, k)] for k in range(3)]
    ctx = _ctx(qh, pivot, passes)
    psi = qh.from_internal(arrays, context=ctx)
    for layer in layers:
        qh.evolve_layer_([(psi, qh.gate(g, b, b + 1)) for b, g in layer], context=ctx)
        for b, g in layer:
            v = sv.apply2(v, g, b)
    got = oc.statevector(oc.Chain(psi.internal_arrays(), psi.lambdas()))
    err = np.max(np.abs(got - v))
    print(f"pivot={pivot} passes={passes}: max |psi - psi_exact| = {err:.2e}")
    assert err < 1e-12


@pytest.mark.parametrize("passes", [2, 3])
def test_truncated_layer_matches_oracle_with_and_without_pivoting(qh, passes):
    rng = np.random.default_rng(7 + passes)
    n = 16
    arrays = wl.left_canonize_normalize(wl.random_mps_arrays(rng, n, 2, wl.max_profile(n, 2, 96)))
    layers = [[(b, wl.haar_gate4(rng)) for b in wl.brickwork_bonds(n, k)] for k in range(2)]
    ref = oc.Chain([a.copy() for a in arrays])
    for layer in layers:
        for b, g in layer:
            oc.evolve(ref, oc.gate2(g, b), threshold=1e-12, maxdim=64)
    z_ref = np.array([oc.expect(ref, [oc.gate1(wl.Z, i)]) for i in range(1, n + 1)])
    for pivot in (True, False):
        ctx = _ctx(qh, pivot, passes)
        psi = qh.from_internal(arrays, context=ctx)
        for layer in layers:
            qh.evolve_layer_([(psi, qh.gate(g, b, b + 1)) for b, g in layer], threshold=1e-12, maxdim=64,
                             context=ctx)
        assert psi.bond_dims() == ref.bond_dims()
        z = qh.expect_many(psi, [[qh.gate(wl.Z, i)] for i in range(1, n + 1)])
        assert np.max(np.abs(z - z_ref)) < 1e-10


def test_width_32_on_a_2048_column_theta(qh):
    """A batch whose widest θ has 2048 or more columns runs the Jacobi at block width 32
    by default (64-column pairs, with the pivoted preconditioner): one 2048 × 2048 θ
    (a gate on two sites between bonds of 1024) against LAPACK's singular values, and
    both new site tensors isometric."""
    rng = np.random.default_rng(97)
    chi = 1024
    arrays = wl.random_mps_arrays(rng, 4, 2, [chi, chi, chi])
    g = wl.haar_gate4(rng)
    th = oc.theta(oc.Chain([a.copy() for a in arrays]), 2, g)
    cl, d, _, cr = th.shape
    s_ref = np.linalg.svd(th.reshape(cl * d, d * cr), compute_uv=False)[:chi]
    psi = qh.from_internal(arrays)
    info = {}
    qh.evolve_(psi, qh.gate(g, 2, 3), maxdim=chi, info=info)
    assert info["kept"] == chi
    lam = psi.lambda_(2)
    err = np.max(np.abs(lam - s_ref)) / s_ref[0]
    a2, a3 = psi.internal_arrays()[1:3]
    u = a2.reshape(-1, a2.shape[2])
    vh = a3.reshape(a3.shape[0], -1)
    iso_u = np.max(np.abs(u.conj().T @ u - np.eye(chi)))
    iso_v = np.max(np.abs(vh @ vh.conj().T - np.eye(chi)))
    print(f"2048 x 2048: max |sigma - sigma_lapack| / sigma_1 = {err:.2e}, isometry {iso_u:.1e} / {iso_v:.1e}")
    assert err < 1e-12
    assert iso_u < 1e-11 and iso_v < 1e-11
