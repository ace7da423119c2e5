"""Parity of the HIP path (through the C ABI) against the CPU oracle.

Tolerances: plain FP64 path, 1e-10 absolute on gauge-invariant quantities
(state vectors of small chains, ⟨Z_i⟩, overlaps, Λ spectra) and identical
kept bond dimensions.
"""
import numpy as np
import pytest

from oracle import chain as oc
from oracle import statevector as sv
from oracle import workloads as wl

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _random_chain(rng, n, d, chi, canon=True):
    arrays = wl.random_mps_arrays(rng, n, d, wl.max_profile(n, d, chi))
    if canon:
        arrays = wl.left_canonize_normalize(arrays)
    return arrays


def _state(psi_gpu):
    return oc.statevector(oc.Chain(psi_gpu.internal_arrays(), _lams(psi_gpu)))


def _lams(psi_gpu):
    return psi_gpu.lambdas()


def test_roundtrip_upload_download(qh):
    rng = np.random.default_rng(1)
    arrays = _random_chain(rng, 5, 2, 4, canon=False)
    psi = qh.from_internal(arrays)
    got = psi.internal_arrays()
    for a, b in zip(arrays, got):
        assert a.shape == b.shape
        assert np.array_equal(a, b)
    assert psi.bond_dims() == [2, 4, 4, 2]


@pytest.mark.parametrize("n,chi,seed", [(4, 2, 0), (6, 4, 1), (8, 8, 2)])
def test_evolve2_no_truncation_matches_statevector(qh, n, chi, seed):
    rng = np.random.default_rng(seed)
    arrays = _random_chain(rng, n, 2, chi)
    psi = qh.from_internal(arrays)
    ref = oc.Chain([a.copy() for a in arrays])
    v = oc.statevector(ref)
    for layer in range(3):
        for b in wl.brickwork_bonds(n, layer):
            g = wl.haar_gate4(rng)
            qh.evolve_(psi, qh.gate(g, b, b + 1))
            oc.evolve(ref, oc.gate2(g, b))
            v = sv.apply2(v, g, b)
    assert psi.bond_dims() == ref.bond_dims()
    assert np.max(np.abs(_state(psi) - v)) < TOL
    for b, (lg, lr) in enumerate(zip(psi.lambdas(), ref.lam)):
        assert (lg is None) == (lr is None)
        if lg is not None:
            assert np.max(np.abs(lg - lr)) < TOL


def test_config1_parity(qh):
    arrays, schedule, tr = wl.config1()
    psi = qh.from_internal(arrays)
    psi0 = psi.copy()
    ref = oc.Chain([a.copy() for a in arrays])
    ref0 = ref.copy()
    for b, g in schedule:
        info = {}
        qh.evolve_(psi, qh.gate(g, b, b + 1), maxdim=tr["maxdim"], info=info)
        oc.evolve(ref, oc.gate2(g, b), maxdim=tr["maxdim"])
        assert info["kept"] == ref.bond_dims()[b - 1]
    assert psi.bond_dims() == ref.bond_dims()
    n = len(arrays)
    z_gpu = qh.expect_many(psi, [[qh.gate(wl.Z, i)] for i in range(1, n + 1)])
    z_ref = np.array([oc.expect(ref, [oc.gate1(wl.Z, i)]) for i in range(1, n + 1)])
    assert np.max(np.abs(z_gpu - z_ref)) < TOL
    ov_gpu = qh.overlap(psi, psi0)
    ov_ref = oc.overlap(ref, ref0)
    assert abs(ov_gpu - ov_ref) < TOL
    for i in (1, 7, 16):
        assert abs(qh.expect(psi, [qh.gate(wl.Z, i)]) - z_ref[i - 1]) < TOL


def test_vidal_mode_matches_oracle(qh):
    rng = np.random.default_rng(11)
    n = 8
    arrays = _random_chain(rng, n, 2, 8)
    psi = qh.from_internal(arrays)
    ref = oc.Chain([a.copy() for a in arrays])
    qh.canonize_(psi)
    oc.canonize(ref)
    for layer in range(4):
        for b in wl.brickwork_bonds(n, layer):
            g = wl.haar_gate4(rng)
            qh.simple_update_(psi, qh.gate(g, b, b + 1), maxdim=6)
            oc.simple_update(ref, oc.gate2(g, b), maxdim=6)
    assert psi.bond_dims() == ref.bond_dims()
    assert np.max(np.abs(_state(psi) - oc.statevector(ref))) < 1e-9
    for lg, lr in zip(psi.lambdas(), ref.lam):
        assert np.max(np.abs(lg - lr)) < 1e-9


def test_canonize_invariants(qh):
    """`Chain_test.jl:308-356`: 2n-1 tensors, value/norm invariant, Σ|Λ|² = ‖ψ‖², canonicality."""
    rng = np.random.default_rng(3)
    arrays = [rng.random((4, 4)), rng.random((4, 4, 4)), rng.random((4, 4, 4)), rng.random((4, 4, 4)), rng.random((4, 4))]
    psi = qh.MPS(arrays)
    ref = oc.MPS(arrays)
    can = qh.canonize(psi)
    assert can.ntensors() == 9
    v0 = oc.statevector(ref)
    assert np.max(np.abs(_state(can) - v0)) < 1e-10 * max(1, np.max(np.abs(v0)))
    nrm = qh.norm(can)
    assert abs(nrm - qh.norm(psi)) < 1e-10 * nrm
    for lam in can.lambdas():
        assert abs(np.sum(lam**2) - nrm**2) < 1e-10 * nrm**2
    ref_can = oc.canonize(ref.copy())
    for lg, lr in zip(can.lambdas(), ref_can.lam):
        assert np.max(np.abs(lg - lr)) < 1e-10 * nrm
    for i in range(1, 6):
        c = qh.canonize(psi)
        if i == 1:
            assert qh.isleftcanonical(c, i)
        elif i == 5:
            qh.contract_between_(c, i - 1, i, direction="right")
            arrs = c.internal_arrays()
            arrs[i - 1] = arrs[i - 1] / nrm
            assert oc.isleftcanonical(oc.Chain(arrs), i)
        else:
            qh.contract_between_(c, i - 1, i, direction="right")
            assert qh.isleftcanonical(c, i)


def test_errors(qh):
    rng = np.random.default_rng(5)
    arrays = [rng.random((2, 2)), rng.random((2, 2, 2)), rng.random((2, 2))]
    psi = qh.MPS(arrays)
    qh.canonize_site_(psi, 2, direction="right", method="svd")
    with pytest.raises(qh.MissingSchmidtCoefficientsException):
        qh.truncate_(psi, (1, 2), maxdim=1)
    t = qh.truncate(psi, (2, 3), maxdim=1)
    assert t.bond_dims()[1] == 1
    lam = psi.lambda_((2, 3))
    t = qh.truncate(psi, (2, 3), threshold=lam[1] + 0.1)
    assert t.bond_dims()[1] == 1
    with pytest.raises(qh.ArgumentError):
        qh.canonize_site_(psi, 1, direction="left")
    with pytest.raises(qh.ArgumentError):
        qh.canonize_site_(psi, 3, direction="right")
    with pytest.raises(qh.ArgumentError):
        qh.evolve_(psi, qh.gate(np.eye(4), 1, 3))
    with pytest.raises(qh.ArgumentError):
        qh.normalize_(qh.MPS(arrays), 1)


def test_batched_layer_equals_single_calls(qh):
    """A brickwork layer through qh_evolve2_batch gives the per-call result."""
    arrays, schedule, tr = wl.config1()
    a = qh.from_internal(arrays)
    b = qh.from_internal(arrays)
    n = len(arrays)
    pos = 0
    for layer in range(8):
        bonds = wl.brickwork_bonds(n, layer)
        ups = schedule[pos: pos + len(bonds)]
        pos += len(bonds)
        qh.evolve_layer_([(a, qh.gate(g, bb, bb + 1)) for bb, g in ups], maxdim=tr["maxdim"])
        for bb, g in ups:
            qh.evolve_(b, qh.gate(g, bb, bb + 1), maxdim=tr["maxdim"])
        assert a.bond_dims() == b.bond_dims()
    za = qh.expect_many(a, [[qh.gate(wl.Z, i)] for i in range(1, n + 1)])
    zb = qh.expect_many(b, [[qh.gate(wl.Z, i)] for i in range(1, n + 1)])
    assert np.max(np.abs(za - zb)) < 1e-12


def test_batched_layers_on_saturated_random_states(qh):
    """Two batched Haar brickwork layers on random MPS already at chimax = 256, as in the
    batched uneven workload: θ is 512 x 512 and full rank from the first layer, so the
    SVDs start from the least favourable case (no small spectrum to converge first).
    Kept bond dimensions and the Λ spectra of the updated bonds match the oracle."""
    rng = np.random.default_rng(21)
    grng = np.random.default_rng(22)
    n, chi = 20, 256
    arrays = [_random_chain(rng, n, 2, chi) for _ in range(2)]
    gpus = [qh.from_internal(a) for a in arrays]
    refs = [oc.Chain([x.copy() for x in a]) for a in arrays]
    for layer in range(2):
        ups = []
        for psi, ref in zip(gpus, refs):
            for b in wl.brickwork_bonds(n, layer):
                g = wl.haar_gate4(grng)
                ups.append((psi, qh.gate(g, b, b + 1)))
                oc.evolve(ref, oc.gate2(g, b), maxdim=chi)
        qh.evolve_layer_(ups, maxdim=chi)
    for psi, ref in zip(gpus, refs):
        assert psi.bond_dims() == ref.bond_dims()
        assert max(ref.bond_dims()) == chi
        checked = 0
        for lg, lr in zip(psi.lambdas(), ref.lam):
            if lr is None:
                continue
            assert lg is not None
            assert np.max(np.abs(lg - lr)) < TOL
            checked += 1
        assert checked >= n // 2


@pytest.mark.parametrize("batched", [False, True])
def test_ising_trotter_prefix_parity(qh, batched):
    """Workload 2 (n=64, cutoff 1e-12, maxdim 256) for its first 2 Trotter steps."""
    arrays, layers, tr, _ = wl.config2(steps=2)
    psi = qh.from_internal(arrays)
    ref = oc.Chain([a.copy() for a in arrays])
    for li, layer in enumerate(layers):
        if batched:
            qh.evolve_layer_([(psi, qh.gate(g, b, b + 1)) for b, g in layer], **tr)
        else:
            for b, g in layer:
                qh.evolve_(psi, qh.gate(g, b, b + 1), **tr)
        for b, g in layer:
            oc.evolve(ref, oc.gate2(g, b), **tr)
        assert psi.bond_dims() == ref.bond_dims(), li
    n = 64
    zs = [[qh.gate(wl.Z, i)] for i in range(1, n + 1)]
    z_gpu = qh.expect_many(psi, zs)
    z_ref = np.array([oc.expect(ref, [oc.gate1(wl.Z, i)]) for i in range(1, n + 1)])
    assert np.max(np.abs(z_gpu - z_ref)) < TOL


def test_config1_matches_golden_fixture(qh):
    from tests.golden import make_golden

    fx = np.load(make_golden.path("config1"), allow_pickle=False)
    arrays, schedule, tr = wl.config1()
    psi = qh.from_internal(arrays)
    psi0 = psi.copy()
    kept = []
    for b, g in schedule:
        info = {}
        qh.evolve_(psi, qh.gate(g, b, b + 1), maxdim=tr["maxdim"], info=info)
        kept.append(info["kept"])
    assert np.array_equal(np.array(kept), fx["kept"])
    z = qh.expect_many(psi, [[qh.gate(wl.Z, i)] for i in range(1, 17)])
    assert np.max(np.abs(z - fx["z"])) < TOL
    assert abs(qh.overlap(psi, psi0) - fx["overlap"]) < TOL


def test_canonize_site_reference(qh):
    """`Chain_test.jl:268-306` through the C ABI."""
    rng = np.random.default_rng(21)
    arrays = [rng.random((4, 4)), rng.random((4, 4, 4)), rng.random((4, 4))]
    qtn = qh.MPS(arrays)
    v0 = oc.statevector(oc.MPS(arrays))
    for method in ("qr", "svd"):
        for site, direction, pred in ((1, "right", qh.isleftcanonical), (2, "right", qh.isleftcanonical),
                                      (2, "left", qh.isrightcanonical), (3, "left", qh.isrightcanonical)):
            c = qh.canonize_site(qtn, site, direction=direction, method=method)
            assert pred(c, site), (method, site, direction)
            assert np.max(np.abs(_state(c) - v0)) < 1e-12 * np.max(np.abs(v0))
    assert qh.canonize_site(qtn, 2, direction="left", method="svd").ntensors() == 4
    with pytest.raises(qh.ArgumentError):
        qh.canonize_site_(qtn, 2, direction="left", method="lu")
    with pytest.raises(qh.ArgumentError):
        qh.canonize_site_(qtn, 2, direction="up")


def test_mixed_canonize_normalize_reference(qh):
    """`Chain_test.jl:358-381` through the C ABI."""
    rng = np.random.default_rng(22)
    arrays = [rng.random((4, 4))] + [rng.random((4, 4, 4)) for _ in range(3)] + [rng.random((4, 4))]
    qtn = qh.MPS(arrays)
    c = qh.mixed_canonize(qtn, 3)
    assert c.ntensors() == qtn.ntensors() + 1
    assert qh.isleftcanonical(c, 1) and qh.isleftcanonical(c, 2)
    assert all(qh.isrightcanonical(c, i) for i in (3, 4, 5))
    v0 = oc.statevector(oc.MPS(arrays))
    assert np.max(np.abs(_state(c) - v0)) < 1e-12 * np.max(np.abs(v0))
    m = qh.MPS(arrays)
    qh.normalize_(m, 3)
    assert abs(qh.norm(m) - 1.0) < 1e-12
    ref = oc.normalize(oc.MPS(arrays), 3)
    assert np.max(np.abs(m.lambda_((2, 3)) - ref.lam[1])) < 1e-12


def test_contract_between_reference(qh):
    """`Chain_test.jl:241-266` through the C ABI."""
    rng = np.random.default_rng(23)
    ref = oc.rand_chain(5, 20, rng=rng)
    qtn = qh.from_internal(ref.arrays)
    can = qh.canonize(qtn)
    with pytest.raises(qh.ArgumentError):
        qh.contract_between_(can.copy(), 1, 2, direction="dummy")
    v0 = oc.statevector(ref)
    for i in range(1, 5):
        some = qh.contract_between(can, i, i + 1)
        assert np.max(np.abs(_state(some) - v0)) < 1e-12
        assert qh.isrightcanonical(some, i)
        assert qh.isleftcanonical(qh.contract_between(can, i, i + 1, direction="right"), i + 1)
        g = can.internal_arrays()[i - 1] * can.lambda_(i)[None, None, :]
        assert np.max(np.abs(some.internal_arrays()[i - 1] - g)) < 1e-13


def test_expect_overlap_norm(qh):
    rng = np.random.default_rng(24)
    n = 7
    a_arr = _random_chain(rng, n, 2, 8, canon=False)
    b_arr = _random_chain(rng, n, 2, 4, canon=False)
    lams = [rng.random(d) + 0.1 for d in [a.shape[2] for a in a_arr[:-1]]]
    a = qh.from_internal(a_arr, lambdas=lams)
    b = qh.from_internal(b_arr)
    ra, rb = oc.Chain([x.copy() for x in a_arr], [l.copy() for l in lams]), oc.Chain([x.copy() for x in b_arr])
    va, vb = oc.statevector(ra), oc.statevector(rb)
    scale = np.max(np.abs(va)) * np.max(np.abs(vb)) * va.size
    assert abs(qh.overlap(a, b) - np.vdot(vb, va)) < 1e-12 * scale
    assert abs(qh.norm(a) - np.linalg.norm(va)) < 1e-12 * np.linalg.norm(va)
    g = wl.haar_gate4(rng)
    nv = np.vdot(va, va).real
    assert abs(qh.expect(a, [qh.gate(g, 3, 4)]) - sv.expect2(va, g, 3)) < 1e-12 * nv
    # products in list order, overlapping supports, both on one cluster and apart
    want = np.vdot(va, sv.apply1(sv.apply1(va, wl.X, 2), wl.Z, 5))
    assert abs(qh.expect(a, [qh.gate(wl.X, 2), qh.gate(wl.Z, 5)]) - want) < 1e-12 * nv
    want = np.vdot(va, sv.apply1(sv.apply2(va, g, 3), wl.Y, 4))
    assert abs(qh.expect(a, [qh.gate(g, 3, 4), qh.gate(wl.Y, 4)]) - want) < 1e-12 * nv
    want = np.vdot(va, sv.apply2(sv.apply2(va, g, 2), g, 3))
    assert abs(qh.expect(a, [qh.gate(g, 2, 3), qh.gate(g, 3, 4)]) - want) < 1e-12 * nv
    terms = [[qh.gate(wl.Z, i)] for i in range(1, n + 1)] + [[qh.gate(wl.Z, i), qh.gate(wl.Z, j)]
                                                              for i in range(1, n) for j in range(i + 1, min(n, i + 4) + 1)]
    got = qh.expect_many(a, terms)
    wants = [sv.expect1(va, wl.Z, i) for i in range(1, n + 1)]
    wants += [np.vdot(va, sv.apply1(sv.apply1(va, wl.Z, i), wl.Z, j)) for i in range(1, n)
              for j in range(i + 1, min(n, i + 4) + 1)]
    assert np.max(np.abs(got - np.array(wants))) < 1e-12 * nv


def test_evolve1_and_renormalize(qh):
    rng = np.random.default_rng(25)
    n = 6
    arrays = _random_chain(rng, n, 2, 4)
    psi = qh.from_internal(arrays)
    ref = oc.Chain([a.copy() for a in arrays])
    u = wl.haar_unitary(rng, 2)
    qh.evolve_(psi, qh.gate(u, 3))
    oc.evolve(ref, oc.gate1(u, 3))
    assert np.max(np.abs(_state(psi) - oc.statevector(ref))) < 1e-13
    # Vidal mode with renormalisation: Λ normalised on the updated bond
    qh.canonize_(psi)
    oc.canonize(ref)
    g = wl.haar_gate4(rng)
    qh.simple_update_(psi, qh.gate(g, 3, 4), maxdim=2, renormalize=True)
    oc.simple_update(ref, oc.gate2(g, 3), maxdim=2, renormalize=True)
    assert abs(np.linalg.norm(psi.lambda_(3)) - 1) < 1e-13
    assert np.max(np.abs(psi.lambda_(3) - ref.lam[2])) < 1e-12
    # mode A renormalisation re-canonises around the left site of the bond
    a = qh.from_internal(arrays)
    r = oc.Chain([x.copy() for x in arrays])
    qh.evolve_(a, qh.gate(g, 3, 4), maxdim=2, renormalize=True)
    oc.evolve(r, oc.gate2(g, 3), maxdim=2, renormalize=True)
    assert a.has_lambda() == [lam is not None for lam in r.lam]
    assert np.max(np.abs(_state(a) - oc.statevector(r))) < 1e-12
    with pytest.raises(qh.ArgumentError):
        qh.evolve_(qh.from_internal(arrays), qh.gate(g, 1, 2), maxdim=2, renormalize=True)


def test_rank_deficient_theta_without_truncation(qh):
    """θ with exactly dependent columns (middle bond 1, outer bonds 4): the
    untruncated update keeps all min(m, n) singular values like zgesdd and the
    state stays exact."""
    rng = np.random.default_rng(26)
    arrays = [rng.standard_normal((1, 2, 2)) + 0j, rng.standard_normal((2, 2, 1)) + 0j,
              rng.standard_normal((1, 2, 2)) + 0j, rng.standard_normal((2, 2, 1)) + 0j]
    psi = qh.from_internal(arrays)
    ref = oc.Chain([a.copy() for a in arrays])
    g = wl.haar_gate4(rng)
    qh.evolve_(psi, qh.gate(g, 2, 3))
    oc.evolve(ref, oc.gate2(g, 2))
    assert psi.bond_dims() == ref.bond_dims()
    assert np.max(np.abs(_state(psi) - oc.statevector(ref))) < 1e-12
    assert qh.isleftcanonical(psi, 2) and qh.isrightcanonical(psi, 3)


def _padded_product_state(rng, n, chi):
    """A product state embedded in bonds of dimension `chi`: only the (0, :, 0)
    slice of each site tensor is non-zero, so every θ has rank <= d^2 and the
    untruncated SVD keeps exactly-zero singular values."""
    dims = [1] + [chi] * (n - 1) + [1]
    arrays = []
    for i in range(n):
        a = np.zeros((dims[i], 2, dims[i + 1]), dtype=np.complex128)
        a[0, :, 0] = rng.standard_normal(2) + 1j * rng.standard_normal(2)
        arrays.append(a)
    return arrays


def _isometry_error(a, side):
    cl, d, cr = a.shape
    if side == "left":
        m = a.reshape(cl * d, cr)  # rows (l, o), columns r  (internal layout (l, o, r))
        return np.max(np.abs(m.conj().T @ m - np.eye(cr)))
    m = a.reshape(cl, d * cr)
    return np.max(np.abs(m @ m.conj().T - np.eye(cl)))


@pytest.mark.parametrize("maxdim", [None, 6])
def test_zero_singular_values_keep_isometric_factors(qh, maxdim):
    """Kept singular values at the Jacobi's deflation floor (here exact zeros of a
    padded product state) get an orthonormal completion of the singular vectors,
    as LAPACK returns: the new sites stay isometries and the state is unchanged.
    maxdim = None keeps all 8 columns of the 8 x 8 θ (6 of them zero);
    maxdim = 6 keeps the default-threshold (1e-16) survivors."""
    rng = np.random.default_rng(31)
    n = 4
    arrays = _padded_product_state(rng, n, 4)
    psi = qh.from_internal(arrays)
    ref = oc.Chain([a.copy() for a in arrays])
    g = wl.haar_gate4(rng)
    info = {}
    qh.evolve_(psi, qh.gate(g, 2, 3), maxdim=maxdim, info=info)
    oc.evolve(ref, oc.gate2(g, 2), maxdim=maxdim)
    if maxdim is None:
        assert info["kept"] == 8
        assert psi.bond_dims() == ref.bond_dims()
    assert np.max(np.abs(_state(psi) - oc.statevector(ref))) < 1e-12
    arrs = psi.internal_arrays()
    assert _isometry_error(arrs[1], "left") < 1e-12
    assert _isometry_error(arrs[2], "right") < 1e-12
    assert qh.isleftcanonical(psi, 2) and qh.isrightcanonical(psi, 3)


@pytest.mark.parametrize("direction", ["right", "left"])
def test_canonize_site_svd_on_padded_product_state(qh, direction):
    """`canonize_site!(ψ, site; method = :svd)` on a rank-1 site matrix padded to
    bond dimension 4: the factor left on the site is an isometry although three of
    its singular values are zero (`Chain_test.jl:268-306` canonicality, atol 1e-12)."""
    rng = np.random.default_rng(32)
    arrays = _padded_product_state(rng, 4, 4)
    psi = qh.from_internal(arrays)
    v0 = oc.statevector(oc.Chain([a.copy() for a in arrays]))
    site = 2 if direction == "right" else 3
    qh.canonize_site_(psi, site, direction=direction, method="svd")
    pred = qh.isleftcanonical if direction == "right" else qh.isrightcanonical
    assert pred(psi, site)
    assert _isometry_error(psi.internal_arrays()[site - 1], "left" if direction == "right" else "right") < 1e-12
    assert np.max(np.abs(_state(psi) - v0)) < 1e-12 * np.max(np.abs(v0))


def test_batched_layer_with_multichunk_descriptor_arena(qh, monkeypatch):
    """Descriptors of one batched call that overflow the arena's first chunk stay
    valid for the whole call (regression: a full arena once reallocated or wrapped
    in the middle of a call and invalidated descriptors uploaded before)."""
    monkeypatch.setenv("QH_ARENA_INITIAL_BYTES", "4096")
    ctx = qh.Context()
    rng = np.random.default_rng(17)
    states, ups, ref_states, ref_ups = [], [], [], []
    for k in range(12):
        n = 10
        arrays = _random_chain(rng, n, 2, 16)
        a = qh.from_internal(arrays, context=ctx)
        b = qh.from_internal(arrays, context=ctx)
        for bd in wl.brickwork_bonds(n, k % 2):
            g = wl.haar_gate4(rng)
            ups.append((a, qh.gate(g, bd, bd + 1)))
            ref_ups.append((b, qh.gate(g, bd, bd + 1)))
        states.append(a)
        ref_states.append(b)
    qh.evolve_layer_(ups, maxdim=16, context=ctx)   # one call, descriptors span several chunks
    for st, g in ref_ups:
        qh.evolve_(st, g, maxdim=16)
    for a, b in zip(states, ref_states):
        assert a.bond_dims() == b.bond_dims()
        for x, y in zip(a.internal_arrays(), b.internal_arrays()):
            assert np.array_equal(x, y)


def test_theta_taller_than_lds_panel(qh):
    """A 640×640 θ (χ = 320): the QR preconditioner's panels are taller than 512 rows,
    so they are factored by the 512-thread register variant (up to 1024 rows) without
    the look-ahead step."""
    from oracle import mpo as om

    rng = np.random.default_rng(23)
    n = 20
    arrays = _random_chain(rng, n, 2, 320)
    psi = qh.from_internal(arrays)
    ref = oc.Chain([a.copy() for a in arrays])
    g = wl.haar_gate4(rng)
    info = {}
    qh.evolve_(psi, qh.gate(g, 10, 11), maxdim=300, info=info)
    oc.evolve(ref, oc.gate2(g, 10), maxdim=300)
    assert psi.bond_dims() == ref.bond_dims()
    lg, lr = psi.lambda_(10), ref.lam[9]
    assert np.max(np.abs(lg - lr)) < 1e-10 * lr[0]
    z = qh.expect_many(psi, [[qh.gate(wl.Z, i)] for i in (9, 10, 11, 12)]) / qh.norm(psi) ** 2
    zr = om.local_z(ref)[[8, 9, 10, 11]]
    assert np.max(np.abs(z - zr)) < 1e-10


def test_theta_taller_than_register_panels(qh):
    """A 1100×66 θ (bond dimensions 550, 80, 33 on four sites): the first QR panels
    have more than 1024 rows and are factored in place in global memory by 512-thread
    workgroups, the later ones in registers.  The evolved state and the new Λ match
    the oracle."""
    rng = np.random.default_rng(29)
    arrays = wl.random_mps_arrays(rng, 4, 2, [550, 80, 33])
    psi = qh.from_internal(arrays)
    ref = oc.Chain([a.copy() for a in arrays])
    g = wl.haar_gate4(rng)
    qh.evolve_(psi, qh.gate(g, 2, 3))
    oc.evolve(ref, oc.gate2(g, 2))
    assert psi.bond_dims() == ref.bond_dims() == [550, 66, 33]
    lg, lr = psi.lambda_(2), ref.lam[1]
    assert np.max(np.abs(lg - lr)) < 1e-10 * lr[0]
    vg = oc.statevector(oc.Chain(psi.internal_arrays(), psi.lambdas()))
    vr = oc.statevector(ref)
    assert np.max(np.abs(vg - vr)) < 1e-10 * np.max(np.abs(vr))


def test_workload2_full_size_time_reversal(qh):
    """Workload 2 at its full size (n = 64, chi <= 256, cutoff 1e-12): 10 Strang steps
    forward, then the same layers in reverse order with the inverse gates.  The
    round trip returns to the initial product state up to the truncation: the
    infidelity is bounded by the summed Eckart-Young errors of all cuts (each cut
    moves the state by at most its discarded norm; the gates are unitary).  The
    update does not keep the chain in canonical form, so a cut can move the norm
    either way: |1 - fidelity| is what the bound holds for."""
    arrays, layers, trunc, _ = wl.config2(steps=10)
    psi = qh.from_internal(arrays)
    psi0 = psi.copy()
    disc = 0.0
    for layer in layers:
        _, d = qh.evolve_layer_([(psi, qh.gate(g, b, b + 1)) for b, g in layer], threshold=trunc["threshold"],
                                maxdim=trunc["maxdim"])
        disc += sum(d)
    assert max(psi.bond_dims()) > 64  # the state became entangled
    for layer in reversed(layers):
        inv = [(b, g.reshape(4, 4, order="F").conj().T.reshape((2, 2, 2, 2), order="F")) for b, g in layer]
        _, d = qh.evolve_layer_([(psi, qh.gate(g, b, b + 1)) for b, g in inv], threshold=trunc["threshold"],
                                maxdim=trunc["maxdim"])
        disc += sum(d)
    fid = abs(qh.overlap(psi, psi0))
    print(f"return fidelity {fid:.15f}, summed discarded norm {disc:.3e}")
    assert abs(1.0 - fid) <= disc + 1e-10
    assert abs(1.0 - fid) < 1e-5


@pytest.mark.parametrize("passes", [0, 1, 2, 3, 4])
def test_qr_preconditioning_passes(qh, passes):
    """Every number of Householder passes preconditioning the Jacobi SVD (X = Q1 R1,
    R_{j-1}^H = Q_j R_j, Jacobi on R_n^H; the odd factors end on the left singular
    vectors, the even ones on the right) evolves like the dense state vector:
    untruncated brickwork layers on 14 sites whose theta grows to 256 x 256 (the
    preconditioned path runs from 64 columns up), and the new sites stay isometries."""
    rng = np.random.default_rng(17)
    n = 14
    arrays = wl.left_canonize_normalize(wl.random_mps_arrays(rng, n, 2, wl.max_profile(n, 2, 64)))
    ctx = qh.Context()
    ctx.set_qr_passes(passes)
    psi = qh.from_internal(arrays, context=ctx)
    v = oc.statevector(oc.Chain([a.copy() for a in arrays]))
    for layer in range(3):
        for b in wl.brickwork_bonds(n, layer):
            g = wl.haar_gate4(rng)
            qh.evolve_(psi, qh.gate(g, b, b + 1))
            v = sv.apply2(v, g, b)
    assert max(psi.bond_dims()) == 128
    got = oc.statevector(oc.Chain(psi.internal_arrays(), psi.lambdas()))
    err = np.max(np.abs(got - v))
    print(f"passes {passes}: max |psi - psi_exact| = {err:.3e}")
    assert err < 1e-10
    b = n // 2
    left = psi.internal_arrays()[b - 1]
    lm = left.reshape(-1, left.shape[2], order="F")
    assert np.max(np.abs(lm.conj().T @ lm - np.eye(lm.shape[1]))) < 1e-12
    right = psi.internal_arrays()[b]
    rm = right.reshape(right.shape[0], -1, order="F")
    assert np.max(np.abs(rm @ rm.conj().T - np.eye(rm.shape[0]))) < 1e-12
    with pytest.raises(qh.ArgumentError):
        ctx.set_qr_passes(5)


@pytest.mark.parametrize("block", [8, 16, 32])
def test_jacobi_block_sizes(qh, block):
    """Every column-block width of the Jacobi tournament (8, 16 and 32 columns per
    block, `qh_ctx_set_jacobi`) evolves like the dense state vector: untruncated
    brickwork layers on 12 sites (theta up to 128 x 128: 16 blocks of 8, 8 of 16 or 4 of
    32, QR-preconditioned, and the small thetas at the chain ends with one block pair)."""
    rng = np.random.default_rng(23)
    n = 12
    arrays = wl.left_canonize_normalize(wl.random_mps_arrays(rng, n, 2, wl.max_profile(n, 2, 32)))
    ctx = qh.Context()
    ctx.set_jacobi(block=block)
    psi = qh.from_internal(arrays, context=ctx)
    v = oc.statevector(oc.Chain([a.copy() for a in arrays]))
    for layer in range(4):
        for b in wl.brickwork_bonds(n, layer):
            g = wl.haar_gate4(rng)
            qh.evolve_(psi, qh.gate(g, b, b + 1))
            v = sv.apply2(v, g, b)
    assert max(psi.bond_dims()) == 64
    err = np.max(np.abs(oc.statevector(oc.Chain(psi.internal_arrays(), psi.lambdas())) - v))
    print(f"block {block}: max |psi - psi_exact| = {err:.3e}")
    assert err < 1e-10
    for lg in psi.lambdas():
        if lg is not None:
            assert np.all(np.diff(lg) <= 0)  # singular values sorted, descending


@pytest.mark.parametrize("padded", [False, True])
def test_canonize_batched_bond_svds_match_oracle(qh, padded):
    """`canonize!` (`Chain.jl:469-497`): the Schmidt values of every bond (one batched
    SVD of the bond matrices of a QR sweep) equal the oracle's dependent SVD sweep;
    Λ_{i-1} Γ_i is a left and Γ_i Λ_i a right isometry; the state is unchanged.  With a
    first bond wider than d (padded) the reference's SVD sweep runs instead."""
    rng = np.random.default_rng(61)
    n = 10
    chis = wl.max_profile(n, 2, 24)
    if padded:
        chis[0] = 4  # site 1 is 1 x 2 x 4: wider than d = 2
    arrays = wl.random_mps_arrays(rng, n, 2, chis)
    psi = qh.from_internal(arrays)
    ref = oc.canonize(oc.Chain([a.copy() for a in arrays]))
    v0 = oc.statevector(oc.Chain([a.copy() for a in arrays]))
    qh.canonize_(psi)
    nrm = np.linalg.norm(v0)
    assert psi.bond_dims() == ref.bond_dims()
    for lg, lr in zip(psi.lambdas(), ref.lam):
        assert np.max(np.abs(lg - lr)) < 1e-10 * nrm
        assert abs(np.sum(lg ** 2) - nrm ** 2) < 1e-10 * nrm ** 2
    arrs, lams = psi.internal_arrays(), psi.lambdas()
    for i in range(1, n + 1):
        a = arrs[i - 1]
        if i <= n - 1:  # Λ_{i-1} Γ_i: orthonormal columns
            left = a * (lams[i - 2][:, None, None] if i >= 2 else 1.0)
            mm = left.reshape(-1, left.shape[2])
            assert np.max(np.abs(mm.conj().T @ mm - np.eye(mm.shape[1]))) < 1e-10, i
        if i >= 2:  # Γ_i Λ_i (Γ_n at the last site): orthonormal rows
            right = a * (lams[i - 1][None, None, :] if i <= n - 1 else 1.0)
            mm = right.reshape(right.shape[0], -1)
            assert np.max(np.abs(mm @ mm.conj().T - np.eye(mm.shape[0]))) < 1e-10, i
    got = oc.statevector(oc.Chain(psi.internal_arrays(), psi.lambdas()))
    assert np.max(np.abs(got - v0)) < 1e-10 * np.max(np.abs(v0))
