"""Parity of the HIP path where the workloads' bond dimensions saturate.

* Workload 2 (n = 64, Ising Trotter, cut-off 1e-12, maxdim 256) through 6 Strang
  steps: χ reaches 256 at layer 14, so the cut-off decides the kept dimensions
  while the state grows and the cap decides them afterwards.  Per call and
  batched, against the oracle's fixture (`tests/golden/config2_saturation.npz`).
  Over the first 3 Strang steps: identical kept dimension for every update and
  every unnormalised ⟨Z_i⟩ within 1e-10.  Beyond that the cut-off at 1e-12 (about
  1e-13 of σ₁) decides kept dimensions at the rounding level, and two correct
  FP64 SVDs part: the fixture holds the same run with LAPACK's zgesvd in place of
  zgesdd (14 of 570 kept dimensions differ, ⟨Z_i⟩ by 1e-9, Λ by 1e-2).  The GPU
  must stay within that spread (stated factors below).
* A reduced batched uneven workload (16 MPS, n ~ U{8..20}, χmax in {16..256},
  4 Haar layers, every layer of every MPS in ONE batched call) against the
  oracle's fixture (`tests/golden/config3_reduced.npz`): FP64 within 1e-10,
  the FP32-MFMA θ form within the bound stated below; kept dimensions identical.
* The mode-A Λ rule (`src/Ansatz/Chain.jl:616`: θ contracts the bond's own Λ,
  not its neighbours') pinned by a mathematical identity: in a state whose
  environment is isometric on both sides of the bond -- with non-trivial Λ's on
  the neighbouring bonds -- the discarded norm of a truncated update equals
  ‖Gψ − ψ_trunc‖ (Eckart–Young on the Schmidt matrix).  A neighbouring Λ
  entering θ breaks the identity (checked on the oracle in test_oracle.py).
"""
import numpy as np
import pytest

from oracle import chain as oc
from oracle import statevector as sv
from oracle import workloads as wl
from tests.golden import make_golden

pytestmark = pytest.mark.gpu

TOL = 1e-10
# FP32 θ (v_mfma_f32_16x16x4_f32, chunk sums in FP64) against the FP64 oracle on the
# reduced batched workload, on unnormalised <Z_i> of normalised states and on the
# overlap with the initial state; the SVD and truncation stay FP64: the
# mixed-precision bound.
FP32_BOUND = 1e-6


def _z_terms(qh, n):
    return [[qh.gate(wl.Z, i)] for i in range(1, n + 1)]


def _contested_decades(window, ref, other, diff, thr):
    """Largest |log10(σ/thr)| over the singular values one run keeps and the other
    drops, read from the fixture's window of zgesdd's values around each cut
    (window[u, t] = σ_{ref[u] - W + t + 1}, 1-based); inf if a step leaves the window."""
    w = window.shape[1] // 2
    worst = 0.0
    for u in diff:
        lo, hi = sorted((int(ref[u]), int(other[u])))
        for k in range(lo, hi):  # 0-based index k = σ_{k+1}
            t = k - (int(ref[u]) - w)
            if not 0 <= t < window.shape[1] or not np.isfinite(window[u, t]):
                return np.inf
            worst = max(worst, abs(np.log10(window[u, t] / thr)))
    return worst


@pytest.mark.parametrize("batched", [False, True])
def test_config2_through_saturation_matches_golden(qh, batched):
    fx = np.load(make_golden.path("config2_saturation"), allow_pickle=False)
    arrays, layers, tr, _ = wl.config2(steps=int(fx["steps"]))
    psi = qh.from_internal(arrays)
    psi0 = psi.copy()
    kept = []
    z_exact = None
    for li, layer in enumerate(layers):
        if batched:
            k, _ = qh.evolve_layer_([(psi, qh.gate(g, b, b + 1)) for b, g in layer], **tr)
            kept.extend(k)
        else:
            for b, g in layer:
                info = {}
                qh.evolve_(psi, qh.gate(g, b, b + 1), info=info, **tr)
                kept.append(info["kept"])
        if li + 1 == 3 * int(fx["exact_steps"]):
            n_exact = len(kept)
            z_exact = qh.expect_many(psi, _z_terms(qh, psi.n))
    kept = np.array(kept)
    ref, alt = fx["kept"], fx["kept_gesvd"]
    thr = tr["threshold"]
    margin = fx["threshold_margin"]
    bad = np.nonzero(kept != ref)[0]
    bad_alt = np.nonzero(alt != ref)[0]
    z = qh.expect_many(psi, _z_terms(qh, psi.n))
    dz, dz_alt = np.max(np.abs(z - fx["z"])), np.max(np.abs(fx["z_gesvd"] - fx["z"]))
    dov = abs(qh.overlap(psi, psi0) - fx["overlap"])
    dov_alt = abs(fx["overlap_gesvd"] - fx["overlap"])
    dlam = max(np.max(np.abs(psi.lambda_(int(b)) - fx[f"lambda{b}"])) for b in fx["lambda_bonds"])
    dlam_alt = max(np.max(np.abs(fx[f"lambda{b}_gesvd"] - fx[f"lambda{b}"])) for b in fx["lambda_bonds"])
    print(f"kept-dimension differences from zgesdd: GPU {bad.size} (first at update {bad[0] if bad.size else '-'}), "
          f"zgesvd {bad_alt.size} (first at {bad_alt[0] if bad_alt.size else '-'}) of {ref.size}")
    print(f"max |d<Z_i>|: GPU {dz:.2e}, zgesvd {dz_alt:.2e}; |d overlap|: GPU {dov:.2e}, zgesvd {dov_alt:.2e}; "
          f"max |d Lambda|: GPU {dlam:.2e}, zgesvd {dlam_alt:.2e}")
    # (1) before the cut-off decides near the rounding level (the first exact_steps Strang
    #     steps, chi up to 85): identical kept dimensions and <Z_i> within 1e-10
    assert np.array_equal(kept[:n_exact], ref[:n_exact])
    assert np.max(np.abs(z_exact - fx["z_exact"])) < TOL
    # (2) through saturation: within the spread of two LAPACK FP64 SVD drivers on the same
    #     inputs -- differing kept dimensions only at cut-off decisions, at most 3x as many
    #     as zgesvd's, and only across singular values within a decade of the cut-off
    #     (the values one side keeps and the other drops: zgesdd's σ_{min+1} .. σ_{max} of
    #     the two cuts); <Z_i>, overlap and Λ no further from zgesdd than 10x zgesvd's
    #     distance (and <Z_i>, overlap within 1e-8 in any case)
    assert np.all(np.isfinite(margin[bad]))
    assert bad.size <= 3 * max(bad_alt.size, 1)
    for name, other, diff in (("GPU", kept, bad), ("zgesvd", alt, bad_alt)):
        dec = _contested_decades(fx["sv_window"], ref, other, diff, thr)
        print(f"{name}: kept-dimension steps {list(other[diff] - ref[diff])}, "
              f"contested singular values at most {dec:.2f} decades from the cut-off")
        assert dec <= 1.0
    assert dz < max(10 * dz_alt, TOL) and dz < 1e-8
    assert dov < max(10 * dov_alt, TOL) and dov < 1e-8
    assert dlam < max(10 * dlam_alt, TOL)
    assert max(psi.bond_dims()) == 256
    assert abs(qh.overlap(psi, psi) - fx["norm2"]) < 1e-8


@pytest.mark.parametrize("precision,bound", [("fp64", TOL), ("fp32", FP32_BOUND)])
def test_config3_reduced_batched_matches_golden(qh, precision, bound):
    fx = np.load(make_golden.path("config3_reduced"), allow_pickle=False)
    items = wl.config3(**make_golden.CONFIG3_REDUCED)
    assert int(fx["count"]) == len(items)
    states = [qh.from_internal(it["arrays"]) for it in items]
    starts = [s.copy() for s in states]
    kept = [[] for _ in items]
    depth = len(items[0]["layers"])
    for layer in range(depth):
        ups, mds, owner = [], [], []
        for m, (it, psi) in enumerate(zip(items, states)):
            for b, g in it["layers"][layer]:
                ups.append((psi, qh.gate(g, b, b + 1)))
                mds.append(it["chi"])
                owner.append(m)
        k, _ = qh.evolve_layer_(ups, maxdim=mds, precision=precision)  # every MPS in one call
        for m, kk in zip(owner, k):
            kept[m].append(kk)
    worst_z = worst_ov = 0.0
    for m, (it, psi) in enumerate(zip(items, states)):
        assert np.array_equal(np.array(kept[m]), fx[f"kept{m}"]), m
        z = qh.expect_many(psi, _z_terms(qh, psi.n))
        worst_z = max(worst_z, float(np.max(np.abs(z - fx[f"z{m}"]))))
        worst_ov = max(worst_ov, abs(qh.overlap(psi, starts[m]) - fx[f"overlap{m}"]))
    print(f"{precision}: max |<Z_i> - oracle| = {worst_z:.2e}, max |overlap - oracle| = {worst_ov:.2e}")
    assert worst_z < bound
    assert worst_ov < bound
    if precision != "fp64":
        assert worst_z > 0.0  # the reduced-precision θ did run


def test_mode_a_discarded_norm_is_state_error(qh):
    rng = np.random.default_rng(41)
    bond, keep = 4, 3
    arrays, lams = wl.isometric_environment_state(rng, bond=bond)
    g = wl.haar_gate4(rng)
    v_exact = sv.apply2(oc.statevector(oc.Chain([a.copy() for a in arrays], [l for l in lams])), g, bond)
    psi = qh.from_internal(arrays, lambdas=lams)
    info = {}
    qh.evolve_(psi, qh.gate(g, bond, bond + 1), maxdim=keep, info=info)
    assert info["kept"] == keep
    v_trunc = oc.statevector(oc.Chain(psi.internal_arrays(), psi.lambdas()))
    err = np.linalg.norm(v_exact - v_trunc)
    print(f"discarded {info['discarded']:.6e}, |G psi - psi_trunc| {err:.6e}")
    assert info["discarded"] > 1e-3
    assert abs(err - info["discarded"]) < 1e-10 * err
