"""Request loop (qrochet_hip.serving) on the CPU: admission, failure isolation,
batch invariance of the scheduling, and the request generator.

The loop runs here over an engine built from the oracle (no GPU): the engine
interface is the loop's only dependency, so its scheduling and error paths are
the ones the device runs.
"""
import numpy as np
import pytest

from oracle import chain as oc
from qrochet_hip import _lib as L
from qrochet_hip import serving as sv
from qrochet_hip.errors import QHError

Z = np.diag([1.0, -1.0]).astype(np.complex128)


class OracleEngine:
    """CPU stand-in for HipEngine (same contract: a failed layer changes no state)."""

    def __init__(self, total=1 << 30):
        self.total = total
        self.layers = 0

    def total_memory(self):
        return self.total

    def create(self, req):
        return sv_product(req.n, req.d)

    def layer(self, work):
        self.layers += 1
        staged = []
        for psi, gates, thr, md in work:
            phi = psi.copy()
            for b, g in gates:
                if not np.all(np.isfinite(g)):
                    raise QHError(L.QH_ERR_NUMERIC, "non-finite entries in the matrix to decompose")
                oc.evolve(phi, oc.gate2(g, b), threshold=thr, maxdim=md)
            staged.append((psi, phi))
        for psi, phi in staged:  # commit only after every update succeeded
            psi.arrays, psi.lam = phi.arrays, phi.lam

    def measure(self, psi, req):
        vals = [oc.expect(psi, [oc.gate1(Z, i)]) for i in range(1, req.n + 1)]
        vals += [oc.expect(psi, [oc.gate1(Z, i), oc.gate1(Z, j)]) for i, j in req.zz]
        return np.array(vals), psi.bond_dims()


def sv_product(n, d):
    e0 = np.zeros(d, dtype=np.complex128)
    e0[0] = 1.0
    return oc.MPS([e0.reshape(d, 1)] + [e0.reshape(d, 1, 1)] * (n - 2) + [e0.reshape(d, 1)])


def small_requests(seed=7, count=5):
    return sv.make_requests(seed, count, qps=50.0, n_range=(6, 9), chi_choices=(4, 8), depth_range=(2, 5))


def test_footprint_closed_form():
    # n=4, d=2, χmax=2: χ = (1,2,2,2,1); sites 2·(2+4+4+2)·16 B, Λ (2+2+2)·8 B;
    # workspace: bond 2 alone (θ 4×4) on one parity, bonds 1 and 3 (θ 2×4, 4×2) on the other
    assert sv.bond_profile(4, 2, 2) == [1, 2, 2, 2, 1]
    assert sv.footprint(4, 2, 2) == 384 + 48 + sv.WORKSPACE_FACTOR * 16 * 16
    # the largest request of the serving distribution fits one MI355X with room to spare
    big = sv.footprint(128, 2, 1024)
    assert 4 << 30 < big < 64 << 30


def test_request_generator():
    reqs = sv.make_requests(3, 400, qps=2.0)
    t = np.array([r.arrival for r in reqs])
    assert np.all(np.diff(t) > 0)
    assert abs(np.mean(np.diff(t)) - 0.5) < 0.1  # Poisson inter-arrivals, mean 1/qps
    assert {r.chimax for r in reqs} <= {32, 64, 128, 256, 512, 1024}
    assert len({r.chimax for r in reqs}) == 6
    assert all(8 <= r.n <= 128 and 1 <= r.depth <= 40 for r in reqs)
    assert all(1 <= j - i <= 4 and j <= r.n for r in reqs for i, j in r.zz)
    # gates are a pure function of (seed, layer) and unitary
    g1 = reqs[0].gates(3)
    g2 = reqs[0].gates(3)
    assert all(b1 == b2 and np.array_equal(a, c) for (b1, a), (b2, c) in zip(g1, g2))
    assert all(np.allclose(g.conj().T @ g, np.eye(4), atol=1e-12) for _, g in g1)
    assert [b for b, _ in g1] == sv.layer_bonds(reqs[0].n, 3)


def test_alone_equals_mix():
    reqs = small_requests()
    mix = sv.Server(OracleEngine).run(reqs, realtime=False)
    assert all(r.status == "ok" for r in mix)
    for r, m in zip(reqs, mix):
        alone = sv.Server(OracleEngine).run([r], realtime=False)[0]
        assert alone.bond_dims == m.bond_dims
        assert np.array_equal(alone.values, m.values)
        assert m.updates == r.updates()


def test_admission_refuses_oversized_and_serves_the_rest():
    reqs = small_requests()
    budget = max(sv.footprint(r.n, r.d, r.chimax) for r in reqs) * 2
    big = sv.Request(rid=99, n=64, chimax=256, depth=2, seed=1, arrival=0.0)
    assert sv.footprint(big.n, big.d, big.chimax) > budget
    srv = sv.Server(OracleEngine, hbm_budget=budget)
    res = srv.run(reqs + [big], realtime=False)
    by = {r.rid: r for r in res}
    assert by[99].status == "admission" and by[99].values is None
    ref = sv.Server(OracleEngine).run(reqs, realtime=False)
    for r in ref:
        assert by[r.rid].status == "ok"
        assert np.array_equal(by[r.rid].values, r.values)
    assert srv.stats["refused"] == 1


def test_budget_bounds_concurrency():
    reqs = small_requests(count=4)
    for r in reqs:
        r.arrival = 0.0
    fp = {r.rid: sv.footprint(r.n, r.d, r.chimax) for r in reqs}
    budget = max(fp.values())
    res = sv.Server(OracleEngine, hbm_budget=budget).run(reqs, realtime=False)
    # at every admission the running requests' projected footprints fit the budget
    for r in res:
        live = [q for q in res if q.admitted <= r.admitted < q.finished]
        assert sum(fp[q.rid] for q in live) <= budget
    # one slot: strictly one request after another, same values
    res2 = sv.Server(OracleEngine, max_active=1).run(reqs, realtime=False)
    spans = sorted((r.admitted, r.finished) for r in res2)
    assert all(spans[k][1] <= spans[k + 1][0] for k in range(len(spans) - 1))
    for a, b in zip(res, res2):
        assert np.array_equal(a.values, b.values)


def test_failed_request_is_isolated():
    reqs = small_requests()
    for r in reqs:
        r.arrival = 0.0
    bad = sv.Request(rid=50, n=7, chimax=8, depth=4, seed=11, arrival=0.0, poison_layer=1)
    srv = sv.Server(OracleEngine)
    res = {r.rid: r for r in srv.run(reqs + [bad], realtime=False)}
    assert res[50].status == "numeric" and "non-finite" in res[50].message
    assert res[50].updates == len(sv.layer_bonds(7, 0))  # layer 0 done, layer 1 failed
    assert srv.stats["isolations"] == 1
    ref = {r.rid: r for r in sv.Server(OracleEngine).run(reqs, realtime=False)}
    for rid, r in ref.items():
        assert res[rid].status == "ok"
        assert np.array_equal(res[rid].values, r.values)


def test_latency_summary():
    rs = [sv.Result(i, "ok", arrival=0.0, finished=float(i + 1)) for i in range(100)]
    rs.append(sv.Result(200, "numeric", arrival=0.0, finished=1e9))
    s = sv.latency_summary(rs)
    assert s["served"] == 100
    assert s["p50_s"] == pytest.approx(50.5)
    assert s["p99_s"] == pytest.approx(99.01)
    assert s["max_s"] == 100.0


def test_lanes_split_by_chimax_and_share_nothing():
    srv = sv.Server(OracleEngine, lanes=(64, 256, 1024))
    mk = lambda chi: sv.Request(rid=0, n=8, chimax=chi, depth=1, seed=0)  # noqa: E731
    assert [srv.lane_of(mk(c)) for c in (4, 64, 65, 256, 512, 1024, 4096)] == [0, 0, 1, 1, 2, 2, 2]
    # the same requests through one lane or three give the same values
    reqs = sv.make_requests(5, 6, qps=50.0, n_range=(6, 8), chi_choices=(2, 4, 8), depth_range=(2, 4))
    one = sv.Server(OracleEngine, lanes=None).run(reqs, realtime=False)
    three = sv.Server(OracleEngine, lanes=(2, 4, 8)).run(reqs, realtime=False)
    for a, b in zip(one, three):
        assert a.status == b.status == "ok"
        assert np.array_equal(a.values, b.values)


class _BrokenEngine(OracleEngine):
    def layer(self, work):
        raise RuntimeError("engine failure outside the status-code contract")


def test_lane_crash_propagates():
    with pytest.raises(RuntimeError, match="outside the status-code contract"):
        sv.Server(_BrokenEngine).run(small_requests(count=2), realtime=False)
