"""The device sampler `rand(Chain, Open, State; n, χ, p, eltype)` (`qh_mps_rand`,
src/Ansatz/Chain.jl:223-256) through the C ABI.

Julia's RNG stream is not reproduced, so the samples are checked by the
properties the reference's own test asserts (`test/Ansatz/Chain_test.jl:209-220`:
norm 1, no bond above χ) and by what the sampler's construction fixes: the bond
profile min(χ, p^b, p^(n-b)) -- the same as the oracle's line-by-line
restatement of the sampler (`oracle/chain.py: rand_chain`) -- sites 2..n right
isometries, site 1 an isometry divided by sqrt(p), real entries for Float64,
and the same sample for the same seed.
"""
import numpy as np
import pytest

from oracle import chain as oc

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,chi,p,eltype", [(8, 10, 2, np.float64), (8, 10, 2, np.complex128),
                                            (7, 3, 3, np.float64), (20, 256, 2, np.complex128),
                                            (12, 600, 2, np.float64)])
def test_rand_chain_properties(qh, n, chi, p, eltype):
    psi = qh.rand_chain(n, chi, p=p, eltype=eltype, seed=5)
    ref = oc.rand_chain(n, chi, p=p, dtype=eltype, rng=np.random.default_rng(5))
    assert psi.bond_dims() == ref.bond_dims()
    assert psi.bond_dims() == [min(chi, p ** b, p ** (n - b)) for b in range(1, n)]
    assert max(psi.bond_dims()) <= chi
    assert abs(qh.norm(psi) - 1.0) < 1e-12
    arrs = psi.internal_arrays()
    for i in range(2, n + 1):
        assert qh.isrightcanonical(psi, i), i
    a1 = arrs[0].reshape(p, -1)  # (o, r)
    assert np.max(np.abs(a1.conj().T @ a1 * p - np.eye(a1.shape[1]))) < 1e-12
    if eltype == np.float64:
        assert all(np.max(np.abs(a.imag)) == 0.0 for a in arrs)
    else:
        assert max(np.max(np.abs(a.imag)) for a in arrs) > 0.1


@pytest.mark.parametrize("seed", [0, 123456789])
def test_rand_chain_same_seed_same_sample(qh, seed):
    a = qh.rand_chain(9, 8, seed=seed).internal_arrays()
    b = qh.rand_chain(9, 8, seed=seed).internal_arrays()
    c = qh.rand_chain(9, 8, seed=seed + 1).internal_arrays()
    assert all(np.array_equal(x, y) for x, y in zip(a, b))
    assert not all(np.allclose(x, y) for x, y in zip(a, c))


def test_rand_chain_arguments(qh):
    with pytest.raises(qh.ArgumentError):
        qh.rand_chain(1, 4)
    with pytest.raises(qh.ArgumentError):
        qh.rand_chain(6, 0)
    with pytest.raises(qh.ArgumentError):
        qh.rand_chain(6, 4, p=5)
