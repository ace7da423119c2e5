"""The Jacobi tail passes (an item that leaves the plain sweeps near convergence: one
whole-item test, then launches of its flagged block pairs in disjoint sets).

* An item's path through the tail does not depend on its batch: a random 18-site MPS
  at chimax = 128 (θ of 256 columns, 16 blocks) gives bit-identical site tensors and Λ
  alone and batched with a 20-site chimax = 256 MPS (θ of 512 columns: the batch's
  windows are 31 rounds long, the small item's own sweeps 15), and the tail ran (its
  test kernel was timed).
* Through the tail, two truncated layers on a random 18-site chimax = 128 MPS match the oracle
  (kept dimensions, <Z_i> at 1e-10); workload 2 through saturation, with the tail on by
  default, is tests/test_gpu_saturation.py.
"""
import numpy as np
import pytest

from oracle import chain as oc
from oracle import workloads as wl

pytestmark = pytest.mark.gpu


def _layer(qh, psi, gates, n, layer):
    return [(psi, qh.gate(gates[b], b, b + 1)) for b in wl.brickwork_bonds(n, layer)]


def test_tail_passes_do_not_depend_on_the_batch(qh):
    rng = np.random.default_rng(123)
    n, nb = 18, 20  # bonds of 128 in the middle of the first chain, of 256 in the second
    a_arrays = wl.left_canonize_normalize(wl.random_mps_arrays(rng, n, 2, wl.max_profile(n, 2, 128)))
    b_arrays = wl.left_canonize_normalize(wl.random_mps_arrays(rng, nb, 2, wl.max_profile(nb, 2, 256)))
    gates = [wl.haar_gate4(rng) for _ in range(nb)]
    ctx = qh.Context()
    ctx.set_timing(True)
    alone = qh.from_internal(a_arrays, context=ctx)
    mixed = qh.from_internal(a_arrays, context=ctx)
    other = qh.from_internal(b_arrays, context=ctx)
    for layer in range(2):
        ua = _layer(qh, alone, gates, n, layer)
        qh.evolve_layer_(ua, maxdim=128, context=ctx)
        um = _layer(qh, mixed, gates, n, layer)
        uo = _layer(qh, other, gates, nb, layer)
        qh.evolve_layer_(um + uo, maxdim=[128] * len(um) + [256] * len(uo), context=ctx)
    ctx.synchronize()
    assert ctx.timing("jacobi_check")["launches"] > 0  # the tail ran
    assert alone.bond_dims() == mixed.bond_dims()
    for x, y in zip(alone.internal_arrays(), mixed.internal_arrays()):
        assert np.array_equal(x, y)
    for b in range(1, n):
        assert np.array_equal(alone.lambda_(b), mixed.lambda_(b))


def test_tail_passes_match_the_oracle(qh):
    rng = np.random.default_rng(321)
    n = 18  # bonds of 128 in the middle: θ of 256 columns
    arrays = wl.left_canonize_normalize(wl.random_mps_arrays(rng, n, 2, wl.max_profile(n, 2, 128)))
    gates = [wl.haar_gate4(rng) for _ in range(n)]
    ref = oc.Chain([a.copy() for a in arrays])
    ctx = qh.Context()
    ctx.set_timing(True)
    psi = qh.from_internal(arrays, context=ctx)
    for layer in range(2):
        qh.evolve_layer_(_layer(qh, psi, gates, n, layer), maxdim=96, context=ctx)
        for b in wl.brickwork_bonds(n, layer):
            oc.evolve(ref, oc.gate2(gates[b], b), threshold=1e-12, maxdim=96)
    ctx.synchronize()
    assert ctx.timing("jacobi_check")["launches"] > 0
    assert psi.bond_dims() == ref.bond_dims()
    z = qh.expect_many(psi, [[qh.gate(wl.Z, i)] for i in range(1, n + 1)])
    z_ref = np.array([oc.expect(ref, [oc.gate1(wl.Z, i)]) for i in range(1, n + 1)])
    assert np.max(np.abs(z - z_ref)) < 1e-10
