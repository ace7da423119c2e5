"""Generates the golden fixtures of tests/golden/*.npz from the oracle.

The reference ships no golden vectors (its tests are property tests), so these
fixtures record the oracle's outputs on seeded inputs *after* the oracle has
been pinned by tests/test_oracle.py (statevector brute force, closed forms and
the reference's own invariants).  They freeze those outputs so that later
changes to the oracle, the workload generators or numpy's streams are caught,
and they give the GPU tests a second, fixed comparison point.

Run:  python -m tests.golden.make_golden
"""
from __future__ import annotations

import os
import sys

import numpy as np
import scipy.linalg as sla

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from oracle import chain as oc  # noqa: E402
from oracle import statevector as sv  # noqa: E402
from oracle import workloads as wl  # noqa: E402


def path(name):
    return os.path.join(HERE, f"{name}.npz")


def config1_outputs():
    """Workload 1: n=16, χ=32, 8 Haar brickwork layers, maxdim=32 (60 updates)."""
    arrays, schedule, tr = wl.config1()
    checksum = float(sum(np.sum(np.abs(a) ** 2 * (1 + i)) for i, a in enumerate(arrays)))
    psi = oc.Chain([a.copy() for a in arrays])
    psi0 = psi.copy()
    kept = []
    for b, g in schedule:
        oc.evolve(psi, oc.gate2(g, b), maxdim=tr["maxdim"])
        kept.append(psi.bond_dims()[b - 1])
    n = psi.n
    z = np.array([oc.expect(psi, [oc.gate1(wl.Z, i)]) for i in range(1, n + 1)])
    return dict(
        input_checksum=checksum,
        kept=np.array(kept, dtype=np.int64),
        z=z,
        overlap=np.complex128(oc.overlap(psi, psi0)),
        final_dims=np.array(psi.bond_dims(), dtype=np.int64),
    )


class _SpectrumRecorder:
    """Records the full spectrum of every SVD the oracle's `evolve` takes (to report
    how close each threshold decision is to the cut-off)."""

    def __enter__(self):
        self.last = None
        self._orig = oc._svd

        def rec(m):
            u, s, vh = self._orig(m)
            self.last = s
            return u, s, vh

        oc._svd = rec
        return self

    def __exit__(self, *exc):
        oc._svd = self._orig


CONFIG2_STEPS = 6
CONFIG2_EXACT_STEPS = 3
CONFIG2_LAMBDA_BONDS = (16, 32, 48)
CONFIG2_WINDOW = 8  # singular values kept on either side of each update's cut


def _gesvd(m):
    return sla.svd(m, full_matrices=False, lapack_driver="gesvd")


def _config2_run(steps, svd=None):
    """Workload 2 through `steps` Strang steps with the oracle, optionally with another
    LAPACK SVD driver; kept dimension and cut-off distance per update, ⟨Z_i⟩ after
    CONFIG2_EXACT_STEPS and after `steps` steps, Λ of bulk bonds, overlaps."""
    arrays, layers, tr, _ = wl.config2(steps=steps)
    psi = oc.Chain([a.copy() for a in arrays])
    psi0 = psi.copy()
    kept, margin, window = [], [], []
    z_exact = None
    orig = oc._svd
    if svd is not None:
        oc._svd = svd
    try:
        with _SpectrumRecorder() as rec:
            for li, layer in enumerate(layers):
                for b, g in layer:
                    oc.evolve(psi, oc.gate2(g, b), **tr)
                    k = psi.bond_dims()[b - 1]
                    sv_ = rec.last
                    kept.append(k)
                    w = np.full(2 * CONFIG2_WINDOW, np.nan)
                    for t in range(2 * CONFIG2_WINDOW):
                        if 0 <= k - CONFIG2_WINDOW + t < len(sv_):
                            w[t] = sv_[k - CONFIG2_WINDOW + t]
                    window.append(w)
                    if k < min(len(sv_), tr["maxdim"]):
                        margin.append(min(abs(np.log(sv_[k - 1] / tr["threshold"])),
                                          abs(np.log(sv_[k] / tr["threshold"]))))
                    else:
                        margin.append(np.inf)
                if li + 1 == 3 * CONFIG2_EXACT_STEPS:
                    z_exact = np.array([oc.expect(psi, [oc.gate1(wl.Z, i)]) for i in range(1, psi.n + 1)])
    finally:
        oc._svd = orig
    n = psi.n
    out = dict(kept=np.array(kept, dtype=np.int64), threshold_margin=np.array(margin), z_exact=z_exact,
               sv_window=np.array(window),
               z=np.array([oc.expect(psi, [oc.gate1(wl.Z, i)]) for i in range(1, n + 1)]),
               norm2=np.complex128(oc.overlap(psi, psi)), overlap=np.complex128(oc.overlap(psi, psi0)),
               final_dims=np.array(psi.bond_dims(), dtype=np.int64))
    for b in CONFIG2_LAMBDA_BONDS:
        out[f"lambda{b}"] = psi.lam[b - 1]
    return out


def config2_saturation_outputs(steps=CONFIG2_STEPS):
    """Workload 2 (n=64, Ising Trotter, cutoff 1e-12, maxdim 256) through `steps`
    Strang steps: the bond dimension reaches 256 at layer 14 (of 3·steps), so the
    cut-off decides the kept dimensions while the state grows and the cap decides
    them once it has saturated.  Records, per update in schedule order, the kept
    dimension and -- where the cut-off (not the cap) decided it -- the distance of
    the decisive singular values from the cut-off, |log(σ/1e-12)| of the smallest
    kept and the largest dropped one; the singular values σ_{k-7} .. σ_{k+8} around
    each cut k (NaN beyond the spectrum); all unnormalised ⟨Z_i⟩ (`expect`) after
    CONFIG2_EXACT_STEPS steps and at the end, Λ of three bulk bonds, ⟨ψ|ψ⟩ and the
    overlap with the initial product state.

    The same run with LAPACK's zgesvd in place of zgesdd (both complex FP64,
    `*_gesvd` entries) measures how far two correct FP64 SVDs carry the state
    apart: once the cut-off at 1e-12 (about 1e-13 of σ₁) decides kept dimensions,
    the decisions depend on rounding, and in the default (not canonical) gauge the
    differences grow over the layers."""
    main = _config2_run(steps)
    other = _config2_run(steps, _gesvd)
    out = dict(steps=np.int64(steps), exact_steps=np.int64(CONFIG2_EXACT_STEPS),
               lambda_bonds=np.array(CONFIG2_LAMBDA_BONDS, dtype=np.int64), **main)
    for k in ("kept", "z", "overlap", "norm2", "sv_window") + tuple(f"lambda{b}" for b in CONFIG2_LAMBDA_BONDS):
        out[f"{k}_gesvd"] = other[k]
    out["z_exact_gesvd"] = other["z_exact"]
    return out


CONFIG3_REDUCED = dict(seed=33, count=16, depth=4, nmin=8, nmax=20)


def config3_reduced_outputs():
    """A reduced batched uneven workload (the shape of workload 3 at a size the
    oracle finishes in seconds): 16 MPS, n ~ U{8..20}, chimax drawn from
    {16, 32, 64, 128, 256}, left-canonized random states at the full bond profile,
    4 Haar brickwork layers, maxdim = chimax (threshold default).  Records every
    update's kept dimension (MPS by MPS, layer by layer), each MPS's unnormalised
    ⟨Z_i⟩ and its overlap with its initial state."""
    items = wl.config3(**CONFIG3_REDUCED)
    out = dict(count=np.int64(len(items)))
    for m, it in enumerate(items):
        psi = oc.Chain([a.copy() for a in it["arrays"]])
        psi0 = psi.copy()
        kept = []
        for layer in it["layers"]:
            for b, g in layer:
                oc.evolve(psi, oc.gate2(g, b), maxdim=it["chi"])
                kept.append(psi.bond_dims()[b - 1])
        out[f"kept{m}"] = np.array(kept, dtype=np.int64)
        out[f"z{m}"] = np.array([oc.expect(psi, [oc.gate1(wl.Z, i)]) for i in range(1, psi.n + 1)])
        out[f"overlap{m}"] = np.complex128(oc.overlap(psi, psi0))
    return out


def small_exact():
    rng = np.random.default_rng(99)
    n = 6
    arrays = wl.random_mps_arrays(rng, n, 2, wl.max_profile(n, 2, 4))
    psi = oc.Chain([a.copy() for a in arrays])
    v = oc.statevector(psi)
    gates, bonds = [], []
    for layer in range(4):
        for b in wl.brickwork_bonds(n, layer):
            g = wl.haar_gate4(rng)
            gates.append(g)
            bonds.append(b)
            v = sv.apply2(v, g, b)
    out = dict(n=n, ngates=len(gates), bonds=np.array(bonds), final_state=v)
    for i, a in enumerate(arrays):
        out[f"site{i}"] = a
    for k, g in enumerate(gates):
        out[f"gate{k}"] = g
    return out


def main():
    np.savez(path("config1"), **config1_outputs())
    np.savez(path("small_exact"), **small_exact())
    np.savez(path("config2_saturation"), **config2_saturation_outputs())
    np.savez(path("config3_reduced"), **config3_reduced_outputs())
    print("wrote", *(path(k) for k in ("config1", "small_exact", "config2_saturation", "config3_reduced")))


if __name__ == "__main__":
    main()
