"""The batched Jacobi SVD on adversarial spectra, through the C ABI.

A 512 x 256 site matrix (the shape of a chi = 256 site, theta's QR-preconditioned
path) is factored by `canonize_site!(ψ, 2; direction = :right, method = :svd)`
(src/Ansatz/Chain.jl:339-376), so the bond's new Λ is the matrix's spectrum and
the site keeps the left factor.  Spectra:
  * clusters of 16 equal singular values (repeated σ),
  * 100 exact zeros (rank 156),
  * a geometric spectrum from 1 down to 1e-14,
  * nearly parallel columns (one column plus 1e-10 perturbations).
Each must converge within the sweep cap, give σ within 1e-12 of LAPACK's zgesdd
(absolute, relative to σ₁) and leave an isometry on the site, in both directions:
‖F^H F - I‖_F <= 1e-12 (the reference's `isleftcanonical` / `isrightcanonical`
criterion, src/Ansatz/Chain.jl:424-458; LAPACK's factors sit near 3e-14 here); and a
truncated evolve! through the same matrix keeps the oracle's bond dimension.
"""
import numpy as np
import pytest

from oracle import chain as oc
from oracle import workloads as wl

pytestmark = pytest.mark.gpu

M, N = 512, 256


def _haar(rng, m, n):
    z = rng.standard_normal((m, n)) + 1j * rng.standard_normal((m, n))
    q, r = np.linalg.qr(z)
    return q * (np.diag(r) / np.abs(np.diag(r)))[None, :]


def _with_spectrum(rng, s):
    return (_haar(rng, M, N) * s[None, :]) @ _haar(rng, N, N).conj().T


def _spectrum_case(kind, rng):
    if kind == "repeated":
        s = np.repeat(np.linspace(1.0, 0.05, N // 16), 16)
        return _with_spectrum(rng, s)
    if kind == "zeros":
        s = np.concatenate([np.linspace(1.0, 1e-3, N - 100), np.zeros(100)])
        return _with_spectrum(rng, s)
    if kind == "graded":
        return _with_spectrum(rng, np.geomspace(1.0, 1e-14, N))
    if kind == "parallel":
        base = rng.standard_normal(M) + 1j * rng.standard_normal(M)
        pert = (rng.standard_normal((M, N)) + 1j * rng.standard_normal((M, N))) * 1e-10
        return base[:, None] + pert
    raise ValueError(kind)


def _chain_with_site(mat):
    """A 3-site chain whose middle site, as the (χl d) × χr matrix, is `mat`."""
    chil, chir = M // 2, N
    mid = mat.reshape(chil, 2, chir, order="F")  # rows (l, o), l fastest
    first = np.zeros((1, 2, chil), dtype=np.complex128)
    first[0, :, :2] = np.eye(2)
    last = np.zeros((chir, 2, 1), dtype=np.complex128)
    last[: 2 * 1, :, 0] = np.eye(2)
    return [first, mid, last]


@pytest.mark.parametrize("direction", ["right", "left"])
@pytest.mark.parametrize("kind", ["repeated", "zeros", "graded", "parallel"])
def test_adversarial_spectra(qh, kind, direction):
    rng = np.random.default_rng({"repeated": 1, "zeros": 2, "graded": 3, "parallel": 4}[kind])
    mat = _spectrum_case(kind, rng)
    arrays = _chain_with_site(mat)
    mid = arrays[1]
    # RIGHT factors the site as (χl d) × χr (= mat), LEFT as χl × (d χr)
    m = mid.reshape(-1, mid.shape[2]) if direction == "right" else mid.reshape(mid.shape[0], -1)
    s_ref = np.linalg.svd(m, compute_uv=False)
    psi = qh.from_internal(arrays)
    qh.canonize_site_(psi, 2, direction=direction, method="svd")
    lam = psi.lambda_((2, 3) if direction == "right" else (1, 2))
    err = np.max(np.abs(lam - s_ref))
    site = psi.internal_arrays()[1]
    if direction == "right":
        f = site.reshape(-1, site.shape[2])
        dev = np.linalg.norm(f.conj().T @ f - np.eye(f.shape[1]))
    else:
        f = site.reshape(site.shape[0], -1)
        dev = np.linalg.norm(f @ f.conj().T - np.eye(f.shape[0]))
    print(f"{kind} {direction}: max |sigma - sigma_zgesdd| = {err:.2e} (sigma_1 = {s_ref[0]:.3e}), "
          f"||F^H F - I||_F = {dev:.2e}")
    assert lam.shape == s_ref.shape
    assert err < 1e-12 * s_ref[0]
    assert dev < 1e-12
    pred = qh.isleftcanonical if direction == "right" else qh.isrightcanonical
    assert pred(psi, 2)


def test_graded_spectrum_threshold_keeps_oracle_dimension(qh):
    """evolve! with an identity gate on θ = A_2 · A_3 (A_3 the identity) truncates the
    graded spectrum at threshold 1e-10: the same kept dimension and Λ as the oracle."""
    rng = np.random.default_rng(5)
    mat = _with_spectrum(rng, np.geomspace(1.0, 1e-14, N))  # θ rows (l, o2); cols (o3, r)
    chil, chir = M // 2, N // 2
    a2 = mat.reshape(chil, 2, N, order="F")           # (l, o, m), m = N
    a3 = np.eye(N, dtype=np.complex128).reshape(N, 2, chir, order="F")  # (m, o, r)
    first = np.zeros((1, 2, chil), dtype=np.complex128)
    first[0, :, :2] = np.eye(2)
    last = np.zeros((chir, 2, 1), dtype=np.complex128)
    last[:2, :, 0] = np.eye(2)
    arrays = [first, a2, a3, last]
    psi = qh.from_internal(arrays)
    ref = oc.Chain([a.copy() for a in arrays])
    eye = wl.as_gate4(np.eye(4))
    info = {}
    qh.evolve_(psi, qh.gate(eye, 2, 3), threshold=1e-10, info=info)
    oc.evolve(ref, oc.gate2(eye, 2), threshold=1e-10)
    assert info["kept"] == ref.bond_dims()[1]
    assert np.max(np.abs(psi.lambda_(2) - ref.lam[1])) < 1e-12
