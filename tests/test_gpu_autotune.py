"""The SVD schedule tuner (`qh_ctx_autotune`): it picks, per size class, a Jacobi
block width and a number of QR preconditioning passes by measurement, and the
schedule it leaves changes only the rounding of the results -- the reduced batched
uneven workload through a tuned context still matches the oracle's fixture at
1e-10 with identical kept dimensions (its matrices fall in several size classes, so
one batch mixes pass counts)."""
import numpy as np
import pytest

from oracle import workloads as wl
from tests.golden import make_golden

pytestmark = pytest.mark.gpu


def test_autotune_report_and_tuned_batch_parity(qh):
    ctx = qh.Context()
    rep = ctx.autotune(max_q=512, batch=8)
    print(rep)
    assert [r["q"] for r in rep] == [64, 128, 256, 512]
    for r in rep:
        assert r["block"] in (8, 16) and 1 <= r["qr_passes"] <= 4
        assert 0.0 < r["best_ms"] <= r["previous_ms"]
    fx = np.load(make_golden.path("config3_reduced"), allow_pickle=False)
    items = wl.config3(**make_golden.CONFIG3_REDUCED)
    states = [qh.from_internal(it["arrays"], context=ctx) for it in items]
    kept = [[] for _ in items]
    for layer in range(len(items[0]["layers"])):
        ups, mds, owner = [], [], []
        for m, (it, psi) in enumerate(zip(items, states)):
            for b, g in it["layers"][layer]:
                ups.append((psi, qh.gate(g, b, b + 1)))
                mds.append(it["chi"])
                owner.append(m)
        k, _ = qh.evolve_layer_(ups, maxdim=mds, context=ctx)
        for m, kk in zip(owner, k):
            kept[m].append(kk)
    for m, psi in enumerate(states):
        assert np.array_equal(np.array(kept[m]), fx[f"kept{m}"]), m
        z = qh.expect_many(psi, [[qh.gate(wl.Z, i)] for i in range(1, psi.n + 1)])
        assert np.max(np.abs(z - fx[f"z{m}"])) < 1e-10


def test_explicit_settings_replace_the_tuned_schedule(qh):
    """After qh_ctx_set_jacobi / qh_ctx_set_qr_passes a tuned context computes bit for
    bit what an untuned context with the same settings computes."""
    rng = np.random.default_rng(71)
    n = 12
    arrays = wl.left_canonize_normalize(wl.random_mps_arrays(rng, n, 2, wl.max_profile(n, 2, 64)))
    gates = [wl.haar_gate4(rng) for _ in range(n)]
    tuned, plain = qh.Context(), qh.Context()
    tuned.autotune(max_q=128, batch=4)
    for ctx in (tuned, plain):
        ctx.set_jacobi(block=16)
        ctx.set_qr_passes(3)
    out = []
    for ctx in (tuned, plain):
        psi = qh.from_internal(arrays, context=ctx)
        for layer in range(2):
            for b in wl.brickwork_bonds(n, layer):
                qh.evolve_(psi, qh.gate(gates[b], b, b + 1), maxdim=48)
        out.append(psi.internal_arrays())
    for a, b in zip(*out):
        assert np.array_equal(a, b)
