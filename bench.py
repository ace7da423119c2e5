"""Benchmark of the MPS two-site update path on one MI355X.

Workload (one bench step): open MPS, n=64, d=2, χmax=256, complex FP64 state,
from a random product state; 20 second-order (Strang) Trotter steps of
h_b = J_b Z⊗Z + (h_b/2)(X⊗I + I⊗X), J_b, h_b ~ U[0.5, 1.5], dt = 0.05, each
step = odd(dt/2), even(dt), odd(dt/2) brickwork layers -> 1900 two-site
gate+SVD updates, truncated with singular-value cutoff 1e-12 and maxdim 256
(`evolve!` default mode).  One layer is one batched pipeline (all bonds of the
layer in one launch sequence).

Printed JSON: updates/s over the timed steps, plus
  * roofline of the dominant kernel (the Jacobi round) from HIP events on the
    library's stream over the timed region,
  * the plain FP64 per-call path (one evolve! per bond) on a bounded sample of
    saturated (χ=256) layers, for the batched/plain ratio,
  * the CPU oracle (numpy/LAPACK zgesdd restatement) on a bounded sample of the
    same saturated layers (`cpu_baseline`),
  * the request loop (`serving`): Poisson arrivals at a fixed QPS of brickwork
    Haar-circuit requests, p50/p99 latency and updates/s.  The default run draws
    12 requests at 0.1 QPS from the full serving distribution (n ~ U{8..128},
    χmax log-uniform in {32..1024}, depth ~ U{1..40}); `--serving sample` uses a
    lighter one (n ~ U{8..64}, χmax up to 256) and `--serving-only` prints the
    leg as its own line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (os.path.join(ROOT, "qrochet.jl_amd"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402

FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def update_flops(d, chil, chim, chir):
    """Algorithmic work of one two-site update (θ build + gate, thin SVD) in real
    flops: F_gemm = 8 d² χl χm χr + 8 d⁴ χl χr (complex MAC = 8 flops) and the
    Golub–Van Loan R-SVD count ×4 for complex, F_svd = 4 (14 m k² + 8 k³) with
    m = max(d χl, d χr), k = min(d χl, d χr)."""
    f_gemm = 8.0 * d * d * chil * chim * chir + 8.0 * d ** 4 * chil * chir
    m, k = max(d * chil, d * chir), min(d * chil, d * chir)
    return f_gemm, 4.0 * (14.0 * m * k * k + 8.0 * k ** 3)


def update_bytes(d, chil, chim, chir, kept):
    """Algorithmic HBM bytes of one FP64 update: read both site tensors and the
    gate, write the two new site tensors and the new Λ."""
    return 16.0 * (d * chil * chim + d * chim * chir + d ** 4) + 16.0 * (d * chil * kept + kept * d * chir) + 8.0 * kept


def run_workload(qh, arrays, layers, trunc, precision="fp64", stop_layer=None, account=None):
    """Runs the layers on a fresh upload of `arrays`.  `account`, if a dict, receives
    the algorithmic flops and bytes of every update from the bond dimensions
    before and after each layer (host metadata only: no device synchronisation)."""
    psi = qh.from_internal(arrays)
    nup = 0
    for li, layer in enumerate(layers):
        if stop_layer is not None and li >= stop_layer:
            break
        ups = [(psi, qh.gate(g, b, b + 1)) for b, g in layer]
        before = psi.bond_dims() if account is not None else None
        kept, _ = qh.evolve_layer_(ups, threshold=trunc["threshold"], maxdim=trunc["maxdim"], precision=precision)
        if account is not None:
            dims = [1] + before + [1]
            for (b, _), k in zip(layer, kept):
                fg, fs = update_flops(2, dims[b - 1], dims[b], dims[b + 1])
                account["f_gemm"] = account.get("f_gemm", 0.0) + fg
                account["f_svd"] = account.get("f_svd", 0.0) + fs
                account["bytes"] = account.get("bytes", 0.0) + update_bytes(2, dims[b - 1], dims[b], dims[b + 1], k)
        nup += len(ups)
    psi.ctx.synchronize()
    return psi, nup


SERVING_SAMPLE = dict(n_range=(8, 64), chi_choices=(32, 64, 128, 256), depth_range=(1, 40))
SERVING_FULL = dict(n_range=(8, 128), chi_choices=(32, 64, 128, 256, 512, 1024), depth_range=(1, 40))


def serving_leg(kind, requests, qps, seed=2025):
    """Run the request loop on `requests` Poisson arrivals at `qps` (real time)."""
    from qrochet_hip import serving as sv

    dist = SERVING_FULL if kind == "full" else SERVING_SAMPLE
    reqs = sv.make_requests(seed, requests, qps, **dist)
    srv = sv.Server()
    t0 = time.perf_counter()
    last = [0.0]

    def on_layer(lane, running, updates):  # at most one line per 20 s
        now = time.perf_counter() - t0
        if now - last[0] > 20.0:
            last[0] = now
            log(f"t={now:.0f} s: lane {lane} finished a layer ({running} requests running)")

    res = srv.run(reqs, realtime=True,
                  log=lambda r: log(f"request {r.rid} {r.status} latency {r.latency:.2f} s ({r.updates} updates)"),
                  log_iter=on_layer)
    wall = time.perf_counter() - t0
    ok = [r for r in res if r.status == "ok"]
    updates = sum(r.updates for r in ok)
    out = {
        "workload": (f"{'full' if kind == 'full' else 'sample of the'} serving distribution: Poisson arrivals, "
                     f"n~U{{{dist['n_range'][0]}..{dist['n_range'][1]}}}, d=2, chimax log-uniform in "
                     f"{{{','.join(map(str, dist['chi_choices']))}}}, depth~U{{{dist['depth_range'][0]}.."
                     f"{dist['depth_range'][1]}}}, Haar brickwork from |0..0>, cutoff 1e-12, observables Z_i and "
                     f"~n random Z_iZ_j (|i-j|<=4), seed {seed}"),
        "qps": qps,
        "requests": requests,
        "lanes_chimax": list(srv.lane_bounds),
    }
    out.update({k: (round(v, 4) if isinstance(v, float) else v) for k, v in sv.latency_summary(res).items()})
    out.update({
        "refused": srv.stats["refused"],
        "failed": sum(r.status not in ("ok", "admission") for r in res),
        "updates": updates,
        "wall_s": round(wall, 3),
        "updates_per_s": round(updates / wall, 2),
        "iterations": srv.stats["iterations"],
    })
    return out


def batched_leg(qh, wl, precision, count, depth):
    """Batched uneven workload: `count` independent MPS (n ~ U{8..64}, χmax uniform
    over {16..256}, random left-canonized initial states at the χmax profile),
    Haar brickwork of `depth` layers; each layer of ALL MPS is one batched call."""
    items = wl.config3(count=count, depth=depth)
    log(f"batched workload: {count} MPS generated")
    ctx = qh.default_context()
    # the SVD schedule (Jacobi block width, QR passes per size class) by the tuner
    t_tune = time.perf_counter()
    tuned = ctx.autotune(max_q=512, batch=32)
    t_tune = time.perf_counter() - t_tune
    log(f"autotune: {tuned} ({t_tune:.1f} s)")
    states = [qh.from_internal(it["arrays"]) for it in items]
    ctx.synchronize()
    n_up = sum(len(l) for it in items for l in it["layers"])
    ctx.reset_timing()
    ctx.set_timing(True)
    t0 = time.perf_counter()
    for layer in range(depth):
        ups, mds = [], []
        for st, it in zip(states, items):
            for b, g in it["layers"][layer]:
                ups.append((st, qh.gate(g, b, b + 1)))
                mds.append(it["chi"])
        qh.evolve_layer_(ups, maxdim=mds, precision=precision)
        ctx.synchronize()
        log(f"layer {layer}: {len(ups)} updates, {time.perf_counter() - t0:.2f} s")
    elapsed = time.perf_counter() - t0
    ctx.set_timing(False)
    classes = ["jacobi_round", "theta", "qr", "qr_apply", "select", "gather"]
    return {
        "metric": "2-site gate+SVD updates/s, batched uneven workload (one launch pipeline per layer of all MPS)",
        "value": round(n_up / elapsed, 2), "unit": "updates/s", "n_gpus": 1, "higher_is_better": True,
        "vs_baseline": None, "dtype": "c128", "data": "synthetic",
        "config": {"workload": f"{count} open MPS, n~U{{8..64}}, d=2, chimax uniform in {{16,32,64,128,256}}, "
                               f"random left-canonized states at the chimax profile, Haar brickwork depth {depth}, "
                               f"maxdim=chimax, threshold default", "precision": precision,
                   "updates": n_up, "wall_s": round(elapsed, 3)},
        "kernel_ms": {k: round(ctx.timing(k)["ms"], 2) for k in classes},
        "autotune": {"classes": tuned, "seconds": round(t_tune, 2)},
    }


def _mpo_run(qh, qm, wl, n, chi, D, sweeps):
    arrays, marr = wl.config4(n=n, chi=chi, D=D)
    ctx = qh.default_context()
    psi = qh.from_internal(arrays)
    h = qm.MPO(marr)
    z_terms = [[qh.gate(wl.Z, i)] for i in range(1, n + 1)]
    res = {}
    for prec in ("fp64", "fp32", "bf16x3"):
        qm.compress(h, psi, sweeps=1, precision=prec)  # warm-up
        ctx.synchronize()
        t0 = time.perf_counter()
        phi, ov = qm.compress(h, psi, sweeps=sweeps, precision=prec)
        ctx.synchronize()
        dt = time.perf_counter() - t0
        nrm2 = qh.norm(phi) ** 2
        z = np.real(qh.expect_many(phi, z_terms)) / nrm2
        res[prec] = {"seconds": dt, "overlap": ov, "z": z}
        log(f"mpo fit n={n} chi={chi} {prec}: {dt:.2f} s, overlap {ov.real:.12e}")
    return arrays, marr, res


def _mpo_errors(res, ref_ov, ref_z, sweeps):
    return {p: {"s_per_sweep": round(r["seconds"] / sweeps, 4),
                "overlap_rel_err": float(abs(r["overlap"] - ref_ov) / abs(ref_ov)),
                "max_abs_dZ": float(np.max(np.abs(r["z"] - ref_z)))} for p, r in res.items()}


def mpo_leg(qh, wl, n, chi, D, sweeps, cpu, cpu_chi):
    """MPO application + variational compression back to χ (workload 4).  At the full
    size each precision is compared with the FP64 GPU path; on a bounded sample
    (χ = cpu_chi, same n and D) all three are compared with the FP64 CPU oracle,
    which also gives the CPU baseline (the oracle at χ = 512 takes minutes per sweep)."""
    from qrochet_hip import mpo as qm

    _, _, res = _mpo_run(qh, qm, wl, n, chi, D, sweeps)
    ref = res["fp64"]
    out = {
        "metric": "MPO application + variational compression: seconds per sweep and error per precision",
        "value": round(ref["seconds"] / sweeps, 4), "unit": "s/sweep (FP64)", "n_gpus": 1,
        "higher_is_better": False, "vs_baseline": None, "dtype": "c128", "data": "synthetic",
        "config": {"workload": f"random Hermitian nearest-neighbour MPO (D={D}, complex Gaussian / sqrt(D d), "
                               f"Hermitian d x d blocks) on a random normalised MPS n={n}, d=2, chi={chi}; one-site "
                               f"variational fit back to chi={chi} from phi0 = psi, {sweeps} sweeps",
                   "n": n, "chi": chi, "D": D, "sweeps": sweeps},
        "precision_vs_fp64_gpu": _mpo_errors(res, ref["overlap"], ref["z"], sweeps),
        "overlap_fp64": [ref["overlap"].real, ref["overlap"].imag],
    }
    if cpu:
        from oracle import chain as oc
        from oracle import mpo as om

        arrays, marr, res_s = _mpo_run(qh, qm, wl, n, cpu_chi, D, sweeps)
        t0 = time.perf_counter()
        ws = om.from_arrays(marr)
        psi_ref = oc.Chain([a.copy() for a in arrays])
        phi_ref = psi_ref.copy()
        for s in range(sweeps):  # one sweep per call, continuing from the previous φ
            phi_ref, ovs = om.fit(ws, psi_ref, phi_ref, sweeps=1)
            log(f"cpu oracle sweep {s} (chi={cpu_chi}): {time.perf_counter() - t0:.1f} s")
        ct = time.perf_counter() - t0
        out["precision_vs_cpu_oracle"] = {"chi": cpu_chi, **_mpo_errors(res_s, ovs[-1], np.real(om.local_z(phi_ref)),
                                                                          sweeps)}
        out["cpu_baseline"] = {"value": round(ct / sweeps, 3), "unit": "s/sweep", "kind": "port",
                               "cores": os.cpu_count(),
                               "sample": f"the same fit at chi={cpu_chi} in numpy (einsum/zgemm), {sweeps} sweeps"}
    return out


def contractions_leg(qh, wl, psi, reps=3):
    """`expect` / `overlap` / `norm` / `canonize!` on a saturated workload state
    (n = 64, chi = 256): all <Z_i> plus every <Z_i Z_j> with |i - j| <= 4 from one
    `qh_expect_many`, the overlap with a copy, the norm, and `canonize!` of a copy
    (the Vidal form: a right-to-left QR sweep and a left-to-right SVD sweep).
    Median wall time over `reps` calls (each ends with its result on the host)."""
    ctx = psi.ctx
    n = psi.n
    terms = [[qh.gate(wl.Z, i)] for i in range(1, n + 1)]
    terms += [[qh.gate(wl.Z, i), qh.gate(wl.Z, j)] for i in range(1, n) for j in range(i + 1, min(n, i + 4) + 1)]
    other = psi.copy()

    def timed(fn):
        ts = []
        for _ in range(reps):
            ctx.synchronize()
            t0 = time.perf_counter()
            fn()
            ctx.synchronize()
            ts.append(time.perf_counter() - t0)
        return round(float(np.median(ts)) * 1e3, 3)

    out = {"state": f"workload state after 10 Trotter steps (n={n}, max bond {max(psi.bond_dims())})",
           "expect_many_terms": len(terms),
           "expect_many_ms": timed(lambda: qh.expect_many(psi, terms)),
           "overlap_ms": timed(lambda: qh.overlap(psi, other)),
           "norm_ms": timed(lambda: qh.norm(psi))}
    copies = [psi.copy() for _ in range(reps)]
    it = iter(copies)
    out["canonize_ms"] = timed(lambda: qh.canonize_(next(it)))
    return out


def cpu_baseline_leg(sat_arrays, sat_lams, layer, trunc, n_updates, reps):
    """The CPU oracle (numpy/scipy LAPACK zgesdd restatement of the update) on the
    first `n_updates` updates of one saturated layer, repeated `reps` times from the
    same state with every host thread and then with one thread; medians."""
    from oracle import chain as oc

    try:
        from threadpoolctl import threadpool_info, threadpool_limits
    except ImportError:  # pragma: no cover - threadpoolctl ships with scikit-learn
        threadpool_info = threadpool_limits = None
    ups = layer[:n_updates]

    def one_rep():
        ref = oc.Chain([a.copy() for a in sat_arrays], [None if l is None else l.copy() for l in sat_lams])
        c0 = time.perf_counter()
        for b, g in ups:
            oc.evolve(ref, oc.gate2(g, b), threshold=trunc["threshold"], maxdim=trunc["maxdim"])
        return len(ups) / (time.perf_counter() - c0)

    threads = None
    if threadpool_info is not None:
        threads = max((i.get("num_threads") or 0) for i in threadpool_info()) or None
    one_rep()  # warm-up (LAPACK workspace queries, page faults)
    rates = [one_rep() for _ in range(reps)]
    rates1 = []
    if threadpool_limits is not None:
        with threadpool_limits(limits=1):
            one_rep()
            rates1 = [one_rep() for _ in range(reps)]
    return {"value": round(float(np.median(rates)), 3), "unit": "updates/s", "cores": threads or os.cpu_count(),
            "kind": "port",
            "value_1_thread": round(float(np.median(rates1)), 3) if rates1 else None,
            "reps": reps, "spread": [round(min(rates), 3), round(max(rates), 3)],
            "spread_1_thread": [round(min(rates1), 3), round(max(rates1), 3)] if rates1 else None,
            "sample": f"the first {len(ups)} updates of a saturated (chi=256, theta 512x512) layer of the same "
                      f"workload from the same state, numpy/scipy-openblas zgesdd restatement of the update, "
                      f"median of {reps} repetitions after one warm-up, all host threads and 1 thread"}


def sample_legs(qh, wl, ctx, arrays, layers, trunc, args):
    """The legs beside the timed steps on a saturated state of the same workload: the
    batched FP64 sample, the FP32 / bf16x3 θ forms, the contractions, the per-call FP64
    form and the CPU baseline.  Returns (extra fields, cpu_baseline)."""
    # saturated sample: layers of Trotter steps 11.. (χ = 256 in the bulk)
    sat_start = max(0, min(30, len(layers) - args.sample_layers))
    sample = layers[sat_start: sat_start + args.sample_layers]
    sample_updates = sum(len(l) for l in sample)
    state_sat, _ = run_workload(qh, arrays, layers, trunc, stop_layer=sat_start)
    sat_arrays = state_sat.internal_arrays()
    sat_lams = state_sat.lambdas()

    log("saturated state ready")
    extra = {}
    # batched on the sample
    psi_b = state_sat.copy()
    ctx.synchronize()
    tb0 = time.perf_counter()
    for layer in sample:
        qh.evolve_layer_([(psi_b, qh.gate(g, b, b + 1)) for b, g in layer], threshold=trunc["threshold"],
                         maxdim=trunc["maxdim"])
    ctx.synchronize()
    tb = time.perf_counter() - tb0
    extra["saturated_batched_updates_per_s"] = round(sample_updates / tb, 2)
    z_terms = [[qh.gate(wl.Z, i)] for i in range(1, psi_b.n + 1)]
    z64 = qh.expect_many(psi_b, z_terms)
    mixed = {}
    for prec in ("fp32", "bf16x3"):
        psi_m = state_sat.copy()
        ctx.synchronize()
        tm0 = time.perf_counter()
        for layer in sample:
            qh.evolve_layer_([(psi_m, qh.gate(g, b, b + 1)) for b, g in layer], threshold=trunc["threshold"],
                             maxdim=trunc["maxdim"], precision=prec)
        ctx.synchronize()
        tm = time.perf_counter() - tm0
        zm = qh.expect_many(psi_m, z_terms)
        mixed[prec] = {"updates_per_s": round(sample_updates / tm, 2),
                       "max_abs_dZ_vs_fp64": float(np.max(np.abs(zm - z64))),
                       "sum_bond_dims": int(sum(psi_m.bond_dims())),
                       "sum_bond_dims_fp64": int(sum(psi_b.bond_dims()))}
    extra["mixed_precision_sample"] = mixed
    extra["contractions"] = contractions_leg(qh, wl, psi_b)
    log("contraction sample done")
    if not args.no_plain:
        psi_p = state_sat.copy()
        ctx.synchronize()
        tp0 = time.perf_counter()
        for layer in sample:
            for b, g in layer:
                qh.evolve_(psi_p, qh.gate(g, b, b + 1), threshold=trunc["threshold"], maxdim=trunc["maxdim"])
        ctx.synchronize()
        tp = time.perf_counter() - tp0
        extra["saturated_plain_fp64_updates_per_s"] = round(sample_updates / tp, 2)
        extra["batched_over_plain"] = round(tp / tb, 3)
        # the mixed-precision target: batched FP32-contraction updates against the plain
        # (per-call, unbatched) FP64 form, at <= 1e-6 error in <Z_i>; and, separately,
        # against the batched FP64 form (the SVD is FP64 in both, so this ratio is ~1)
        m32 = mixed["fp32"]
        extra["mixed_precision_target"] = {
            "mixed_over_plain_fp64": round(m32["updates_per_s"] * tp / sample_updates, 3),
            "mixed_over_batched_fp64": round(m32["updates_per_s"] * tb / sample_updates, 3),
            "max_abs_dZ_vs_fp64": m32["max_abs_dZ_vs_fp64"], "bound": 1e-6, "speedup_goal": 5.0,
            "met": bool(m32["updates_per_s"] * tp / sample_updates >= 5.0 and m32["max_abs_dZ_vs_fp64"] <= 1e-6),
            "note": "speed from batching; the SVD stays FP64 (FP32 Jacobi sweeps cannot resolve the "
                    "spectrum below ~1e-7 of sigma_1, where half the kept values of a saturated theta lie)"}
        extra["saturated_sample"] = f"{args.sample_layers} layers ({sample_updates} updates) at chi<=256 after 10 Trotter steps"

    log("plain sample done")
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline_leg(sat_arrays, sat_lams, sample[0], trunc, args.cpu_updates, args.cpu_reps)

    return extra, cpu


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--trotter", type=int, default=20, help="Trotter steps per bench step")
    ap.add_argument("--sample-layers", type=int, default=3, help="saturated layers for plain/CPU samples")
    ap.add_argument("--cpu-updates", type=int, default=8, help="updates per CPU-baseline repetition")
    ap.add_argument("--cpu-reps", type=int, default=5, help="CPU-baseline repetitions (median reported)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-samples", action="store_true",
                    help="only the timed steps (no saturated sample, contraction, plain or CPU legs): "
                         "a profile of the process then covers exactly the timed launches")
    ap.add_argument("--serving", choices=("sample", "full", "off"), default="full",
                    help="request-loop leg: the full serving distribution (default: 12 requests at 0.1 QPS), "
                         "a lighter sample (n <= 64, chimax <= 256), or none")
    ap.add_argument("--serving-only", action="store_true", help="run only the request loop, print its own line")
    ap.add_argument("--serving-requests", type=int, default=None)
    ap.add_argument("--qps", type=float, default=None)
    ap.add_argument("--workload", choices=("tebd", "batched", "mpo", "contractions"), default="tebd",
                    help="tebd: the n=64, chi=256 Trotter evolution (default); batched: 512 uneven MPS; "
                         "mpo: MPO application + variational compression precision study; contractions: "
                         "expect_many / overlap / norm / canonize! on the chi=256 workload state")
    ap.add_argument("--mpo-n", type=int, default=32)
    ap.add_argument("--mpo-chi", type=int, default=512)
    ap.add_argument("--mpo-D", type=int, default=8)
    ap.add_argument("--mpo-sweeps", type=int, default=2)
    ap.add_argument("--mpo-cpu-chi", type=int, default=256, help="chi of the sample checked against the CPU oracle")
    ap.add_argument("--precision", choices=("fp64", "fp32", "bf16x3"), default=None,
                    help="contraction precision of the batched workload (default fp32)")
    ap.add_argument("--batch-count", type=int, default=512)
    ap.add_argument("--batch-depth", type=int, default=10)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first run a plain 1 GiB device copy (known bytes) for HBM counter calibration")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); multi-GPU sharding is not built.",
              file=sys.stderr)
        sys.exit(2)

    import qrochet_hip as qh

    if args.calibrate_copy:
        import torch

        src = torch.empty(1 << 27, dtype=torch.float64, device="cuda").fill_(1.0)
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.mul(src, 1.0, out=dst)  # an elementwise kernel: 1 GiB read, 1 GiB written
        torch.cuda.synchronize()
        log("calibration copy: 3 launches, each 1 GiB read + 1 GiB written (torch elementwise mul by 1)")
        del src, dst

    if args.workload == "mpo":
        from qrochet_hip import workloads as wl

        print(json.dumps(mpo_leg(qh, wl, args.mpo_n, args.mpo_chi, args.mpo_D, args.mpo_sweeps, not args.no_cpu,
                                 args.mpo_cpu_chi)))
        return

    if args.workload == "contractions":
        from qrochet_hip import workloads as wl

        arrays, layers, trunc, _ = wl.config2(steps=args.trotter)
        psi, _ = run_workload(qh, arrays, layers, trunc, stop_layer=min(30, len(layers)))
        out = contractions_leg(qh, wl, psi)
        print(json.dumps({"metric": "expect_many (all Z_i and Z_iZ_j, |i-j|<=4) / overlap / norm / canonize! "
                                    "wall time on the workload-2 state after 10 Trotter steps",
                          "value": out["expect_many_ms"], "unit": "ms", "n_gpus": 1, "higher_is_better": False,
                          "vs_baseline": None, "dtype": "c128", "data": "synthetic",
                          "config": {"workload": out["state"]}, "contractions": out}))
        return

    if args.workload == "batched":
        from qrochet_hip import workloads as wl

        print(json.dumps(batched_leg(qh, wl, args.precision or "fp32", args.batch_count, args.batch_depth)))
        return

    if args.serving != "off":
        full = args.serving == "full"
        n_req = args.serving_requests or (12 if full else 48)
        qps = args.qps or (0.1 if full else 4.0)
    if args.serving_only:
        if args.serving == "off":
            sys.exit("--serving-only needs --serving sample|full")
        sv_out = serving_leg(args.serving, n_req, qps)
        print(json.dumps({
            "metric": "p99 latency of brickwork-circuit requests at fixed QPS (request loop)",
            "value": sv_out.get("p99_s"), "unit": "s", "n_gpus": 1, "higher_is_better": False,
            "vs_baseline": None, "dtype": "c128", "data": "synthetic",
            "config": {"workload": sv_out["workload"], "qps": qps, "requests": n_req},
            "serving": sv_out,
        }))
        return

    from qrochet_hip import workloads as wl

    arrays, layers, trunc, _ = wl.config2(steps=args.trotter)
    n_updates = sum(len(l) for l in layers)
    ctx = qh.default_context()

    # algorithmic flops / bytes of one step (every step does the same updates: the
    # inputs and the arithmetic are identical, so are the kept dimensions)
    account = {}
    for w in range(args.warmup):
        tw = time.perf_counter()
        psi, _ = run_workload(qh, arrays, layers, trunc, account=account if w == 0 else None)
        log(f"warmup {w}: {time.perf_counter() - tw:.2f} s, max bond {max(psi.bond_dims())}")
    ctx.synchronize()
    ctx.reset_timing()
    ctx.set_timing(True)
    ctx.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        psi, nup = run_workload(qh, arrays, layers, trunc, account=account if not account else None)
        log(f"timed step {k} done at {time.perf_counter() - t0:.2f} s")
    ctx.synchronize()
    t1 = time.perf_counter()
    ctx.set_timing(False)
    classes = ["jacobi_round", "jacobi_check", "theta", "qr", "qr_apply", "select", "gather", "gemm", "jacobi_sweep"]
    timers = {k: ctx.timing(k) for k in classes}
    elapsed = t1 - t0
    value = n_updates * args.steps / elapsed
    final_dims = psi.bond_dims()

    # Roofline of the dominant kernel, the Jacobi round.  `achieved` follows the
    # algorithmic count of a two-site update (F_gemm + F_svd at each update's own
    # (χl, χm, χr), summed over the timed steps) divided by the summed duration of
    # the Jacobi launches (HIP events on the library's stream), against the FP64
    # peak: the fraction of the FP64 roof the update path would reach if all of
    # its algorithmic work ran in the time of this one kernel.
    jr = timers["jacobi_round"]
    jr_s = jr["ms"] * 1e-3
    f_alg = (account["f_gemm"] + account["f_svd"]) * args.steps
    b_alg = account["bytes"] * args.steps
    launches = max(jr["launches"], 1)
    achieved_tf = f_alg / jr_s / 1e12 if jr_s > 0 else 0.0
    roof = {"bound": "mfma", "achieved": round(achieved_tf, 3), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": round(achieved_tf / FP64_PEAK_TFLOPS, 4)}
    # HBM bytes per launch measured by rocprofv3 FETCH_SIZE / WRITE_SIZE (copy-calibrated,
    # and counting L2 misses served by the Infinity Cache as well as HBM reads)
    roof["traffic"] = None
    tf = os.path.join(ROOT, "profiles", "k_jacobi_round_traffic.json")
    if os.path.exists(tf):
        with open(tf) as fh:
            t = json.load(fh)
        roof["traffic"] = t["traffic_bytes_per_launch"]
        roof["traffic_source"] = f"profiles/k_jacobi_round_traffic.json ({t['workload']})"
    roof["kernel"] = "k_jacobi_round<32, false> (block one-sided Jacobi round, FP64)"
    roof["avg_launch_ms"] = round(jr["ms"] / launches, 5)
    roof["launches"] = jr["launches"]
    roof["algorithmic_flops_per_launch"] = f_alg / launches
    roof["algorithmic_bytes_per_launch"] = b_alg / launches
    # what the kernel itself streams and computes (counted on the device: Gram reads of
    # every checked block pair, X/V reads and writes of every rotated one)
    roof["kernel_streamed_bytes_per_launch"] = jr["bytes"] / launches
    roof["kernel_streamed_over_algorithmic_bytes"] = round(jr["bytes"] / b_alg, 1) if b_alg > 0 else None
    roof["kernel_executed_flops_per_launch"] = jr["flops"] / launches
    roof["kernel_executed_tflops"] = round(jr["flops"] / jr_s / 1e12, 3) if jr_s > 0 else None
    # end to end: roofline time of all updates of the timed steps over the wall time
    t_star = account["f_gemm"] * args.steps / (FP64_PEAK_TFLOPS * 1e12) + \
        account["f_svd"] * args.steps / (FP64_PEAK_TFLOPS * 1e12)
    end_to_end = {"t_star_s": round(t_star, 4), "t_measured_s": round(elapsed, 4),
                  "frac": round(t_star / elapsed, 4),
                  "algorithmic_gflop_per_step": round((account["f_gemm"] + account["f_svd"]) / 1e9, 1)}

    extra, cpu = {}, None
    if not args.no_samples:
        extra, cpu = sample_legs(qh, wl, ctx, arrays, layers, trunc, args)

    out = {
        "metric": "2-site gate+SVD updates/s at chi=256 (MPS TEBD, Ising Trotter workload)",
        "value": round(value, 3),
        "unit": "updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128",
        "data": "synthetic",
        "config": {
            "workload": f"open MPS n=64 d=2 chimax=256, Ising Trotter {args.trotter} Strang steps dt=0.05, "
                        f"{n_updates} two-site updates per step, cutoff 1e-12, one batched launch pipeline per layer",
            "n": 64, "d": 2, "chimax": 256, "updates_per_step": n_updates, "precision": "fp64",
            "final_max_bond": max(final_dims),
        },
        "roofline": roof,
        "end_to_end_roofline": end_to_end,
        "cpu_baseline": cpu,
        "kernel_ms": {k: round(v["ms"], 3) for k, v in timers.items()},
        "jacobi_sweeps": timers["jacobi_sweep"]["launches"],
    }
    out.update(extra)
    if args.serving != "off":
        log(f"serving leg: {n_req} requests at {qps} QPS")
        out["serving"] = serving_leg(args.serving, n_req, qps)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
