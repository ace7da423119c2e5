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
    print("wrote", path("config1"), path("small_exact"))


if __name__ == "__main__":
    main()
