"""Reduced-precision θ contraction on MFMA (QH_PREC_FP32, QH_PREC_BF16X3)
against the FP64 oracle.  The SVD, truncation and everything after θ stay FP64.

Stated tolerances (bounds asserted here; the measured errors are printed):
  * FP32 operands and products on the matrix cores, FP32 accumulation within
    16-wide chunks of the contracted index and the chunk sums in FP64: the
    mixed-precision bound 1e-6, on the state vector of an untruncated evolution
    (relative to ‖ψ‖ = 1) and on ⟨Z_i⟩ of workload 1 (60 updates);
  * bf16x3 (hi/lo bfloat16 split, three products): 1e-4 for both -- its operands
    carry 16 significant bits, so it cannot meet 1e-6.
Kept bond dimensions are identical to the FP64 path on these full-rank cases.
"""
import numpy as np
import pytest

from oracle import chain as oc
from oracle import statevector as sv
from oracle import workloads as wl

pytestmark = pytest.mark.gpu

BOUND = {"fp32": 1e-6, "bf16x3": 1e-4}


@pytest.mark.parametrize("precision", ["fp32", "bf16x3"])
def test_reduced_precision_statevector(qh, precision):
    rng = np.random.default_rng(11)
    n = 8
    arrays = wl.left_canonize_normalize(wl.random_mps_arrays(rng, n, 2, wl.max_profile(n, 2, 16)))
    psi = qh.from_internal(arrays)
    ref = oc.Chain([a.copy() for a in arrays])
    v = oc.statevector(ref)
    for layer in range(4):
        for b in wl.brickwork_bonds(n, layer):
            g = wl.haar_gate4(rng)
            qh.evolve_(psi, qh.gate(g, b, b + 1), precision=precision)
            v = sv.apply2(v, g, b)
    got = oc.statevector(oc.Chain(psi.internal_arrays(), psi.lambdas()))
    err = np.max(np.abs(got - v))
    print(f"{precision}: max |psi - psi_fp64| = {err:.3e}")
    assert err < BOUND[precision]
    assert err > 0.0  # the reduced-precision kernel did run


@pytest.mark.parametrize("precision", ["fp32", "bf16x3"])
def test_reduced_precision_config1(qh, precision):
    arrays, schedule, tr = wl.config1()
    psi = qh.from_internal(arrays)
    ref = oc.Chain([a.copy() for a in arrays])
    for b, g in schedule:
        qh.evolve_(psi, qh.gate(g, b, b + 1), maxdim=tr["maxdim"], precision=precision)
        oc.evolve(ref, oc.gate2(g, b), maxdim=tr["maxdim"])
    assert psi.bond_dims() == ref.bond_dims()
    n = len(arrays)
    z = qh.expect_many(psi, [[qh.gate(wl.Z, i)] for i in range(1, n + 1)])
    z_ref = np.array([oc.expect(ref, [oc.gate1(wl.Z, i)]) for i in range(1, n + 1)])
    err = np.max(np.abs(z - z_ref))
    print(f"{precision}: workload 1 max |<Z_i> - <Z_i>_fp64| = {err:.3e}")
    assert err < BOUND[precision]


@pytest.mark.parametrize("precision", ["fp32", "bf16x3"])
def test_reduced_precision_vidal_and_batched(qh, precision):
    """Vidal mode (row/column Λ scales in the epilogue) through the batched entry
    point equals the per-call reduced-precision path bit for bit, and stays
    within the bound of the FP64 oracle."""
    rng = np.random.default_rng(5)
    n = 8
    arrays = wl.left_canonize_normalize(wl.random_mps_arrays(rng, n, 2, wl.max_profile(n, 2, 8)))
    a = qh.canonize(qh.from_internal(arrays))
    b = a.copy()
    ref = oc.canonize_copy(oc.Chain([x.copy() for x in arrays]))
    for layer in range(2):
        gates = [(bd, wl.haar_gate4(rng)) for bd in wl.brickwork_bonds(n, layer)]
        qh.evolve_layer_([(a, qh.gate(g, bd, bd + 1)) for bd, g in gates], iscanonical=True, precision=precision)
        for bd, g in gates:
            qh.evolve_(b, qh.gate(g, bd, bd + 1), iscanonical=True, precision=precision)
            oc.evolve(ref, oc.gate2(g, bd), iscanonical=True)
    for x, y in zip(a.internal_arrays(), b.internal_arrays()):
        assert np.array_equal(x, y)
    va = oc.statevector(oc.Chain(a.internal_arrays(), a.lambdas()))
    vr = oc.statevector(ref)
    err = np.max(np.abs(va - vr))
    print(f"{precision}: Vidal mode, batched, max |psi - psi_fp64| = {err:.3e}")
    assert err < BOUND[precision]
