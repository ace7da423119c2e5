"""`qh_overlap_noalloc` (the overlap sweep in the library's fixed per-GPU scratch,
no device allocation) against `qh_overlap` through the C ABI: two calls in a row,
the second on a smaller pair of chains, each within 1e-12 relative of the
allocating entry point; and the size limit of the scratch is an error, not a fault."""
import numpy as np
import pytest

from oracle import workloads as wl

pytestmark = pytest.mark.gpu


def _pair(qh, rng, n, chi):
    a = qh.from_internal(wl.random_mps_arrays(rng, n, 2, wl.max_profile(n, 2, chi)))
    b = qh.from_internal(wl.random_mps_arrays(rng, n, 2, wl.max_profile(n, 2, chi)))
    return a, b


def test_overlap_noalloc_matches_overlap_twice(qh):
    rng = np.random.default_rng(77)
    # first call: 24 sites at chi = 256, Λ on the bonds of one brickwork layer of the ket
    a, b = _pair(qh, rng, 24, 256)
    g = wl.haar_gate4(rng)
    for bd in wl.brickwork_bonds(24, 0):
        qh.evolve_(a, qh.gate(g, bd, bd + 1), maxdim=256)
    want = qh.overlap(a, b)
    got = qh.overlap_noalloc(a, b)
    print(f"chi 256: {got} vs {want}")
    assert abs(got - want) <= 1e-12 * abs(want)
    assert abs(qh.overlap_noalloc(a, a) - qh.overlap(a, a)) <= 1e-12 * abs(qh.overlap(a, a))
    # second call on a smaller pair: the scratch keeps the first call's contents, the
    # sweep must not read them
    c, d = _pair(qh, rng, 9, 8)
    want = qh.overlap(c, d)
    got = qh.overlap_noalloc(c, d)
    print(f"chi 8: {got} vs {want}")
    assert abs(got - want) <= 1e-12 * abs(want)


def test_overlap_noalloc_rejects_chains_beyond_its_scratch(qh):
    big = [2048, 2048, 1]
    arrays = [np.zeros((1, 2, 2048), complex), np.zeros((2048, 2, 2048), complex),
              np.zeros((2048, 2, 1), complex), np.ones((1, 2, 1), complex)]
    arrays[0][0, 0, 0] = 1.0
    a = qh.from_internal(arrays)
    assert a.bond_dims() == big
    with pytest.raises(qh.ArgumentError):
        qh.overlap_noalloc(a, a)
