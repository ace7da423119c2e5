"""Continuous-batching request loop over HBM-resident MPS states.

The reference has no serving layer: a Qrochet.jl user calls ``evolve!`` gate by
gate on one ``Chain`` (src/Ansatz/Chain.jl:543-661) and ``expect`` at the end
(src/Ansatz/Chain.jl:724-735).  Here many such jobs share the GPU:

* a request is (n sites, d = 2, χmax, depth, seed) plus an observable list: a
  brickwork circuit of Haar-random two-site gates applied to |0…0⟩, truncated
  at χmax (and singular values below ``threshold``), then ⟨Z_i⟩ and ⟨Z_i Z_j⟩;
* requests are admitted when they have arrived and their projected device
  footprint (state at the χmax profile plus one layer of SVD workspace) fits the
  memory budget next to the requests already running; one whose footprint alone
  exceeds the budget is refused with status ``"admission"``;
* admitted requests join a lane by χmax; every iteration of a lane advances each
  of its running requests by one brickwork layer, all of them in ONE batched
  update pipeline (`qh_evolve2_batch`), so the work of requests that arrived at
  different times overlaps on the device; lanes run concurrently on their own
  streams; a request that reaches its depth is measured (`qh_expect_many`) and
  released at once, and its memory admits the next one;
* a failure inside a batch (non-finite tensor, non-convergence) is isolated by
  re-running that iteration request by request: the batched call writes no state
  before every decomposition of the batch has succeeded, so the failed request
  gets its status code and the others continue unaffected.

Results do not depend on the batch a request shares: every kernel of the
pipeline treats each update on its own (per-matrix Jacobi tolerance, sorting
window and deflation floor, fixed-order reductions), so a request's values are
bit-identical alone and in the mix.
"""
from __future__ import annotations

import collections
import dataclasses
import threading
import time

import numpy as np

from . import _lib as L
from .chain import MPS, Context, default_context, evolve_layer_, expect_many, gate
from .errors import QHError

PAULI_Z = np.diag([1.0, -1.0]).astype(np.complex128)

# Bytes of SVD workspace per complex element of a two-site matrix: θ, its QR
# factors and rotations, and the new site tensors (about ten matrices of θ's size).
WORKSPACE_FACTOR = 10


def haar_unitary(dim: int, rng: np.random.Generator) -> np.ndarray:
    """Haar-random U(dim): QR of a complex Ginibre matrix with R's diagonal phases
    moved into Q (Mezzadri, Notices AMS 54 (2007) 592)."""
    z = (rng.standard_normal((dim, dim)) + 1j * rng.standard_normal((dim, dim))) / np.sqrt(2.0)
    q, r = np.linalg.qr(z)
    ph = np.diagonal(r) / np.abs(np.diagonal(r))
    return q * ph[None, :]


def layer_bonds(n: int, layer: int) -> list[int]:
    """Bonds (1-based, bond b joins sites b and b+1) of brickwork layer ``layer``
    (0-based): odd bonds on even layers, even bonds on odd layers."""
    return list(range(1 + (layer % 2), n, 2))


@dataclasses.dataclass
class Request:
    rid: int
    n: int
    chimax: int
    depth: int
    seed: int
    arrival: float = 0.0  # seconds after the start of the run
    zz: list = dataclasses.field(default_factory=list)  # (i, j) pairs, 1-based, 1 <= j - i <= 4
    d: int = 2
    threshold: float = 1e-12
    poison_layer: int = -1  # layer whose first gate is replaced by NaN (failure-path tests)

    def gates(self, layer: int) -> list[tuple[int, np.ndarray]]:
        """(bond, 4×4 gate) pairs of one layer, from a stream seeded by (seed, layer)."""
        rng = np.random.default_rng([self.seed, layer])
        out = [(b, haar_unitary(self.d * self.d, rng)) for b in layer_bonds(self.n, layer)]
        if layer == self.poison_layer and out:
            out[0] = (out[0][0], np.full((4, 4), np.nan, dtype=np.complex128))
        return out

    def observables(self) -> list:
        """⟨Z_i⟩ for every site, then ⟨Z_i Z_j⟩ for every listed pair."""
        terms = [[gate(PAULI_Z, i)] for i in range(1, self.n + 1)]
        terms += [[gate(PAULI_Z, i), gate(PAULI_Z, j)] for i, j in self.zz]
        return terms

    def updates(self) -> int:
        return sum(len(layer_bonds(self.n, l)) for l in range(self.depth))


@dataclasses.dataclass
class Result:
    rid: int
    status: str  # "ok" | "admission" | "numeric" | "not_converged" | "error"
    message: str = ""
    values: np.ndarray | None = None  # ⟨Z_1⟩ … ⟨Z_n⟩, then the ⟨Z_i Z_j⟩ in request order
    bond_dims: list | None = None
    updates: int = 0
    arrival: float = 0.0
    admitted: float = float("nan")
    finished: float = float("nan")

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


def bond_profile(n: int, d: int, chimax: int) -> list[int]:
    """Largest reachable bond dimensions χ_0 … χ_n (χ_0 = χ_n = 1) under the cap χmax."""
    chi = [1] * (n + 1)
    for b in range(1, n):
        left = d ** min(b, 64)
        right = d ** min(n - b, 64)
        chi[b] = min(left, right, chimax)
    return chi


def footprint(n: int, d: int, chimax: int) -> int:
    """Projected device bytes of a request: site tensors and Λ vectors at the χmax
    profile, plus one layer of SVD workspace (the larger of the two parities)."""
    chi = bond_profile(n, d, chimax)
    state = sum(d * chi[i - 1] * chi[i] * 16 for i in range(1, n + 1)) + sum(chi[b] * 8 for b in range(1, n))
    ws = [0, 0]
    for b in range(1, n):
        ws[(b - 1) % 2] += WORKSPACE_FACTOR * (d * chi[b - 1]) * (d * chi[b + 1]) * 16
    return state + max(ws)


def make_requests(seed: int, count: int, qps: float, n_range=(8, 128), chi_choices=(32, 64, 128, 256, 512, 1024),
                  depth_range=(1, 40), zz_per_site: float = 1.0) -> list[Request]:
    """Poisson arrivals at ``qps``; n ~ U{n_range}, χmax log-uniform over ``chi_choices``
    (powers of two), depth ~ U{depth_range}; ⟨Z_iZ_j⟩ for a random selection of
    about ``zz_per_site``·n pairs with 1 <= j - i <= 4."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for rid in range(count):
        t += rng.exponential(1.0 / qps)
        n = int(rng.integers(n_range[0], n_range[1] + 1))
        chimax = int(chi_choices[rng.integers(len(chi_choices))])
        depth = int(rng.integers(depth_range[0], depth_range[1] + 1))
        cands = [(i, j) for i in range(1, n + 1) for j in range(i + 1, min(i + 4, n) + 1)]
        k = min(len(cands), int(round(zz_per_site * n)))
        pick = sorted(rng.choice(len(cands), size=k, replace=False).tolist()) if k else []
        out.append(Request(rid=rid, n=n, chimax=chimax, depth=depth, seed=int(rng.integers(2**31)), arrival=t,
                           zz=[cands[p] for p in pick]))
    return out


def product_state(n: int, d: int = 2, context: Context | None = None):
    """|0…0⟩ as an MPS with bond dimension 1 (reference-shaped arrays (o, r), (o, l, r), (o, l))."""
    e0 = np.zeros(d, dtype=np.complex128)
    e0[0] = 1.0
    arrays = [e0.reshape(d, 1)] + [e0.reshape(d, 1, 1) for _ in range(n - 2)] + [e0.reshape(d, 1)]
    return MPS(arrays, context=context)


_STATUS = {L.QH_ERR_NUMERIC: "numeric", L.QH_ERR_NOT_CONVERGED: "not_converged", L.QH_ERR_OOM: "oom"}


class HipEngine:
    """What a lane runs on: its own context (one HIP stream), states in device
    memory, one `qh_evolve2_batch` per layer of the lane's batch, `qh_expect_many`
    at completion."""

    def __init__(self, context: Context | None = None):
        self.ctx = context or Context()

    def total_memory(self) -> int:
        return self.ctx.mem_info()[1]

    def create(self, req: Request):
        return product_state(req.n, req.d, self.ctx)

    def layer(self, work):
        """``work``: list of (state, [(bond, gate)], threshold, maxdim).  Raises
        QHError without having changed any state if one update fails."""
        ups, ths, mds = [], [], []
        for psi, gates, thr, md in work:
            ups += [(psi, gate(g, b, b + 1)) for b, g in gates]
            ths += [thr] * len(gates)
            mds += [md] * len(gates)
        evolve_layer_(ups, threshold=ths, maxdim=mds, context=self.ctx)

    def measure(self, psi, req: Request):
        return expect_many(psi, req.observables()), psi.bond_dims()


class _Running:
    __slots__ = ("req", "psi", "layer", "fp", "res")

    def __init__(self, req, psi, fp, res):
        self.req, self.psi, self.fp, self.res = req, psi, fp, res
        self.layer = 0


class _Lane(threading.Thread):
    """One batching lane: a host thread driving its own engine (own HIP stream),
    advancing its running requests one brickwork layer per batched call."""

    def __init__(self, server: "Server", index: int, engine):
        super().__init__(name=f"qh-lane-{index}", daemon=True)
        self.srv, self.index, self.engine = server, index, engine
        self.inbox: collections.deque = collections.deque()
        self.active: list[_Running] = []
        self.error: BaseException | None = None

    def run(self):
        try:
            self._loop()
        except BaseException as e:  # reported by Server.run
            self.error = e
            with self.srv.cv:
                self.srv.cv.notify_all()

    def _loop(self):
        srv = self.srv
        while True:
            with srv.cv:
                while not self.inbox and not self.active and not srv.closing:
                    srv.cv.wait()
                if srv.closing and not self.inbox and not self.active:
                    return
                new = list(self.inbox)
                self.inbox.clear()
            for r, fp, res in new:
                try:
                    psi = self.engine.create(r)
                except QHError as e:
                    res.status, res.message = _STATUS.get(e.code, "error"), str(e)
                    srv._finish(r, fp, res)
                    continue
                self.active.append(_Running(r, psi, fp, res))
            if not self.active:
                continue
            self._step()
            srv._count("iterations")
            if srv._log_iter:
                srv._log_iter(self.index, len(self.active), sum(a.res.updates for a in self.active))
            done = [a for a in self.active if a.layer >= a.req.depth or a.res.status != "ok"]
            for a in done:
                if a.res.status == "ok":
                    try:
                        a.res.values, a.res.bond_dims = self.engine.measure(a.psi, a.req)
                    except QHError as e:
                        a.res.status, a.res.message = _STATUS.get(e.code, "error"), str(e)
                a.psi = None  # releases the state
                self.active.remove(a)
                srv._finish(a.req, a.fp, a.res)

    @staticmethod
    def _work(a: _Running):
        return (a.psi, a.req.gates(a.layer), a.req.threshold, a.req.chimax)

    def _step(self):
        work = [self._work(a) for a in self.active]
        try:
            self.engine.layer(work)
        except QHError:
            # isolate the failure: no state was written, so rerun request by request
            self.srv._count("isolations")
            for a, w in zip(self.active, work):
                try:
                    self.engine.layer([w])
                except QHError as e:
                    a.res.status, a.res.message = _STATUS.get(e.code, "error"), str(e)
                    continue
                a.layer += 1
                a.res.updates += len(w[1])
            return
        for a, w in zip(self.active, work):
            a.layer += 1
            a.res.updates += len(w[1])


class Server:
    """The request loop.

    ``lanes``: χmax bounds of the batching lanes; a request joins the first lane
    whose bound is >= its χmax (the last lane takes the rest).  Each lane batches
    only requests of similar size, because a batched call lasts as long as its
    largest decomposition, and runs on its own stream from its own host thread, so
    lanes share the GPU's compute units concurrently: a χ=32 request is not held
    behind the layer of a χ=1024 one.  ``engine_factory()`` makes one engine per
    lane (default: `HipEngine`, a fresh context each).

    ``hbm_budget`` (bytes) bounds the summed projected footprint of admitted
    requests (default: 85 % of the device memory); ``max_active`` optionally
    bounds their number.  Admission is first come, first served."""

    def __init__(self, engine_factory=None, hbm_budget: int | None = None, max_active: int | None = None,
                 lanes=(64, 256, 1024)):
        self.engine_factory = engine_factory or HipEngine
        self.lane_bounds = tuple(lanes) if lanes else (1 << 30,)
        self._hbm_budget = hbm_budget
        self.budget = None
        self.max_active = max_active
        self.stats = collections.Counter()
        self.cv = threading.Condition()
        self.closing = False

    def lane_of(self, req: Request) -> int:
        for i, b in enumerate(self.lane_bounds):
            if req.chimax <= b:
                return i
        return len(self.lane_bounds) - 1

    def _count(self, key, k=1):
        with self.cv:
            self.stats[key] += k

    def _finish(self, req, fp, res):
        with self.cv:
            res.finished = self._now()
            self._results[req.rid] = res
            self._used -= fp
            self._inflight -= 1
            self.cv.notify_all()
        if self._log:
            self._log(res)

    def run(self, requests, realtime: bool = True, log=None, log_iter=None) -> list[Result]:
        """Serve ``requests`` (arrival times in seconds after the call).  With
        ``realtime`` False every request is taken as arrived at once, in arrival
        order (values are the same; latencies are then not meaningful).  ``log(result)``
        is called per finished request, ``log_iter(lane, running, updates)`` after
        every batched layer of a lane."""
        engines = [self.engine_factory() for _ in self.lane_bounds]
        if self.budget is None:
            self.budget = int(self._hbm_budget if self._hbm_budget is not None
                              else 0.85 * engines[0].total_memory())
        pending = collections.deque(sorted(requests, key=lambda r: (r.arrival, r.rid)))
        waiting: collections.deque = collections.deque()
        self._results: dict[int, Result] = {}
        self._used = 0
        self._inflight = 0
        self._log = log
        self._log_iter = log_iter
        self.closing = False
        t0 = time.perf_counter()
        self._now = lambda: time.perf_counter() - t0
        lanes = [_Lane(self, i, e) for i, e in enumerate(engines)]
        for ln in lanes:
            ln.start()
        try:
            with self.cv:
                while pending or waiting or self._inflight:
                    for ln in lanes:
                        if ln.error is not None:
                            raise ln.error
                    t = self._now()
                    while pending and (not realtime or pending[0].arrival <= t):
                        waiting.append(pending.popleft())
                    while waiting:
                        r = waiting[0]
                        fp = footprint(r.n, r.d, r.chimax)
                        if fp > self.budget:
                            waiting.popleft()
                            self._results[r.rid] = Result(
                                r.rid, "admission", f"projected footprint {fp} B exceeds the budget {self.budget} B",
                                arrival=r.arrival, finished=self._now())
                            self.stats["refused"] += 1
                            continue
                        if self._used + fp > self.budget or (
                                self.max_active is not None and self._inflight >= self.max_active):
                            break
                        waiting.popleft()
                        self._used += fp
                        self._inflight += 1
                        self.stats["admitted"] += 1
                        res = Result(r.rid, "ok", arrival=r.arrival if realtime else 0.0, admitted=self._now())
                        lanes[self.lane_of(r)].inbox.append((r, fp, res))
                        self.cv.notify_all()
                    if not (pending or waiting or self._inflight):
                        break
                    timeout = None
                    if pending and realtime:
                        timeout = max(0.0, pending[0].arrival - self._now())
                    self.cv.wait(timeout)
        finally:
            with self.cv:
                self.closing = True
                self.cv.notify_all()
            for ln in lanes:
                ln.join()
        for ln in lanes:
            if ln.error is not None:
                raise ln.error
        return [self._results[r.rid] for r in sorted(requests, key=lambda r: r.rid)]


def latency_summary(results: list[Result]) -> dict:
    """p50 / p99 latency (seconds) over the requests served successfully."""
    lat = np.array([r.latency for r in results if r.status == "ok"])
    if lat.size == 0:
        return {"served": 0}
    return {
        "served": int(lat.size),
        "p50_s": float(np.percentile(lat, 50)),
        "p99_s": float(np.percentile(lat, 99)),
        "mean_s": float(lat.mean()),
        "max_s": float(lat.max()),
    }
