"""Deferred V updates of the Jacobi rounds (block width 16): a pair of a matrix with at
least 256 columns leaves V_IJ <- V_IJ Q to a V workgroup of the next launch, and a window
of rounds ends with a launch of V workgroups alone.  The products and their order are
those of the direct update, so the results are bit-identical to a context with the
deferral off (QH_JACOBI_DEFER_V=0, read when a context is created).

Workload: two truncated brickwork layers on a random 18-site MPS at chimax = 256 (θ of
512 × 512 in the bulk, about 130 block pairs per launch: the deferral is on), and the
same on a context without it; the deferred run shows the extra V-only launches.
"""
import os

import numpy as np
import pytest

from oracle import workloads as wl

pytestmark = pytest.mark.gpu


def _run(qh, ctx, arrays, gates, n):
    psi = qh.from_internal(arrays, context=ctx)
    ctx.reset_timing()
    ctx.set_timing(True)
    for layer in range(2):
        ups = [(psi, qh.gate(gates[b], b, b + 1)) for b in wl.brickwork_bonds(n, layer)]
        qh.evolve_layer_(ups, maxdim=256, context=ctx)
    ctx.synchronize()
    ctx.set_timing(False)
    return psi, ctx.timing("jacobi_round")["launches"]


def test_deferred_v_updates_are_bit_identical(qh):
    rng = np.random.default_rng(77)
    n = 18
    arrays = wl.left_canonize_normalize(wl.random_mps_arrays(rng, n, 2, wl.max_profile(n, 2, 256)))
    gates = [wl.haar_gate4(rng) for _ in range(n)]
    ctx_on = qh.Context()
    old = os.environ.get("QH_JACOBI_DEFER_V")
    os.environ["QH_JACOBI_DEFER_V"] = "0"
    try:
        ctx_off = qh.Context()
    finally:
        if old is None:
            del os.environ["QH_JACOBI_DEFER_V"]
        else:
            os.environ["QH_JACOBI_DEFER_V"] = old
    a, launches_on = _run(qh, ctx_on, arrays, gates, n)
    b, launches_off = _run(qh, ctx_off, arrays, gates, n)
    assert launches_on > launches_off  # one V-only launch per window of rounds
    assert a.bond_dims() == b.bond_dims()
    for x, y in zip(a.internal_arrays(), b.internal_arrays()):
        assert np.array_equal(x, y)
    for bond in range(1, n):
        assert np.array_equal(a.lambda_(bond), b.lambda_(bond))
