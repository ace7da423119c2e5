"""Request loop on the device: each request's result is the same alone and in the
mix (bit-identical, the stated tolerance being zero), matches the oracle within
1e-10 with identical kept bond dimensions, and a failing request (non-finite
gate) is reported with its status while the others are unaffected."""
import numpy as np
import pytest

from test_serving import OracleEngine

pytestmark = pytest.mark.gpu


def _requests(sv):
    reqs = sv.make_requests(21, 6, qps=100.0, n_range=(8, 14), chi_choices=(8, 16, 32), depth_range=(3, 8))
    # one request deep enough to reach χ = 64 (θ 128×128: QR-preconditioned Jacobi with several block pairs)
    reqs.append(sv.Request(rid=6, n=12, chimax=64, depth=12, seed=99, arrival=reqs[-1].arrival, zz=[(3, 7), (5, 6)]))
    return reqs


def test_serving_alone_equals_mix(qh):
    from qrochet_hip import serving as sv

    reqs = _requests(sv)
    mix = sv.Server().run(reqs, realtime=False)
    assert [r.status for r in mix] == ["ok"] * len(reqs), [r.message for r in mix]
    assert max(max(r.bond_dims) for r in mix) == 64
    for r, m in zip(reqs, mix):
        print("request", r.rid, "n", r.n, "chimax", r.chimax, "depth", r.depth, flush=True)
        alone = sv.Server().run([r], realtime=False)[0]
        assert alone.bond_dims == m.bond_dims
        assert np.array_equal(alone.values, m.values)
        assert m.updates == r.updates()


def test_serving_matches_oracle(qh):
    from qrochet_hip import serving as sv

    reqs = _requests(sv)
    gpu = sv.Server().run(reqs, realtime=False)
    cpu = sv.Server(OracleEngine).run(reqs, realtime=False)
    for g, c in zip(gpu, cpu):
        assert g.status == c.status == "ok"
        assert g.bond_dims == c.bond_dims
        np.testing.assert_allclose(g.values, c.values, rtol=0, atol=1e-10)


def test_serving_isolates_nonfinite_request(qh):
    from qrochet_hip import serving as sv

    reqs = _requests(sv)[:4]
    for r in reqs:
        r.arrival = 0.0
    bad = sv.Request(rid=40, n=10, chimax=16, depth=5, seed=5, arrival=0.0, poison_layer=2)
    srv = sv.Server()
    res = {r.rid: r for r in srv.run(reqs + [bad], realtime=False)}
    assert res[40].status == "numeric", res[40].message
    assert srv.stats["isolations"] == 1
    ref = {r.rid: r for r in sv.Server().run(reqs, realtime=False)}
    for rid, r in ref.items():
        assert res[rid].status == "ok"
        assert res[rid].bond_dims == r.bond_dims
        assert np.array_equal(res[rid].values, r.values)
    # the library is usable after the failure and no device memory leaked into later calls
    again = sv.Server().run(reqs[:1], realtime=False)[0]
    assert np.array_equal(again.values, ref[reqs[0].rid].values)
