"""Brute-force statevector reference for small chains (n <= ~16).

TEST INFRASTRUCTURE ONLY (see `oracle/chain.py`).  Used to pin the MPS
oracle: with no truncation an MPS update must reproduce the exact state.

State layout: numpy array with one axis per site (axis k = site k+1).
Gate layout: (out_l, out_r, in_l, in_r), the order `Dense(Operator(), G;
sites=[i, i+1, i', (i+1)'])` gives (`src/Ansatz/Dense.jl:21-34`).
"""
from __future__ import annotations

import numpy as np


def product_state(vectors):
    v = np.asarray(vectors[0], dtype=np.complex128)
    for w in vectors[1:]:
        v = np.multiply.outer(v, np.asarray(w, dtype=np.complex128))
    return v


def apply1(psi, g, site):
    """ψ'[..., a, ...] = Σ_b G[a, b] ψ[..., b, ...] on 1-based ``site``."""
    ax = site - 1
    out = np.tensordot(g, psi, axes=(1, ax))
    return np.moveaxis(out, 0, ax)


def apply2(psi, g4, site):
    """ψ'[.., a, b, ..] = Σ G[a, b, c, e] ψ[.., c, e, ..] on sites (site, site+1)."""
    ax = site - 1
    out = np.tensordot(g4, psi, axes=([2, 3], [ax, ax + 1]))
    return np.moveaxis(out, [0, 1], [ax, ax + 1])


def expect1(psi, op, site):
    """⟨ψ|O_site|ψ⟩ (unnormalised)."""
    return complex(np.vdot(psi, apply1(psi, op, site)))


def expect2(psi, op4, site):
    return complex(np.vdot(psi, apply2(psi, op4, site)))


def schmidt_values(psi, cut):
    """Singular values across the cut after ``cut`` sites."""
    d = psi.shape
    m = psi.reshape(int(np.prod(d[:cut])), -1)
    return np.linalg.svd(m, compute_uv=False)
