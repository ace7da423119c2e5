"""CPU restatement of the MPO × MPS path (exact application and one-site
variational fitting) in complex FP64.

TEST INFRASTRUCTURE: the checker for qh_mpo_apply / qh_mpo_fit; nothing in the
library imports it.  The reference has no MPO application or compression routine
(its MPO is `Chain(Operator(), Open(), arrays)`, src/Ansatz/Chain.jl:20,133-172,
index order (o, i, l, r), :34; `merge`, src/Quantum.jl:330-348, would join the
networks), so this is the build's own algorithm, pinned by closed forms in
tests/test_oracle_mpo.py: dense H·ψ for small chains, exactness of the fit when
φ's bonds can hold H|ψ⟩, the identity MPO, and the variational monotonicity.

Internal layouts: MPS site A[l, o, r] (as oracle.chain), MPO site W[l, o, i, r].
"""
from __future__ import annotations

import numpy as np

from . import chain as oc

ORDER = ("o", "i", "l", "r")
# pairwise contraction order without numpy's default cap on intermediate sizes (the
# cap makes einsum fall back to the naive 8-index loop for these networks)
_OPT = ("greedy", 1e13)


def from_arrays(arrays, order=ORDER):
    """Reference-shaped MPO arrays -> internal W[l, o, i, r] (boundary bonds 1)."""
    n = len(arrays)
    out = []
    for k, a in enumerate(arrays):
        a = np.asarray(a, dtype=np.complex128)
        names = list(order)
        if k == 0:
            names = [x for x in names if x != "l"]
        if k == n - 1:
            names = [x for x in names if x != "r"]
        assert a.ndim == len(names)
        if "l" not in names:
            a = a[..., None]
            names.append("l")
        if "r" not in names:
            a = a[..., None]
            names.append("r")
        out.append(np.ascontiguousarray(np.transpose(a, [names.index(x) for x in ("l", "o", "i", "r")])))
    return out


def to_arrays(ws, order=ORDER):
    """Internal W[l, o, i, r] -> reference-shaped arrays in ``order``."""
    n = len(ws)
    out = []
    for k, w in enumerate(ws):
        names = ["l", "o", "i", "r"]
        a = w
        if k == 0:
            a = a[0]
            names = ["o", "i", "r"]
        if k == n - 1:
            a = a[..., 0]
            names = [x for x in names if x != "r"]
        want = [x for x in order if x in names]
        out.append(np.ascontiguousarray(np.transpose(a, [names.index(x) for x in want])))
    return out


def random_hermitian_mpo(rng, n, d, D):
    """Reference-shaped arrays (o, i, l, r) of a random Hermitian nearest-neighbour
    MPO: complex Gaussian entries × 1/sqrt(D·d), each d×d block W[w, :, :, v]
    replaced by its Hermitian part, which makes H Hermitian (operators on
    different sites commute)."""
    dims = [1] + [D] * (n - 1) + [1]
    ws = []
    for k in range(n):
        shp = (dims[k], d, d, dims[k + 1])
        w = (rng.standard_normal(shp) + 1j * rng.standard_normal(shp)) / np.sqrt(2.0 * D * d)
        w = 0.5 * (w + np.conj(np.transpose(w, (0, 2, 1, 3))))
        ws.append(w)
    return to_arrays(ws)


def dense(ws):
    """H as a tensor with axes (o_1 … o_n, i_1 … i_n), site k on axes k-1 and n+k-1.
    A periodic MPO (first left bond > 1, `pMPO`) is traced over its wrap bond."""
    n = len(ws)
    t = ws[0]  # (w, o, i, r)
    for w in ws[1:]:
        t = np.tensordot(t, w, axes=(t.ndim - 1, 0))  # (w, …, o, i, r')
    t = np.trace(t, axis1=0, axis2=t.ndim - 1)  # (o1, i1, o2, i2, …, on, in)
    return np.transpose(t, list(range(0, 2 * n, 2)) + list(range(1, 2 * n, 2)))


def from_arrays_periodic(arrays, order=ORDER):
    """`pMPO` arrays (all 4-D) -> internal W[l, o, i, r]."""
    return [np.ascontiguousarray(np.transpose(np.asarray(a, dtype=np.complex128),
                                              [list(order).index(x) for x in ("l", "o", "i", "r")])) for a in arrays]


def apply_dense(ws, v):
    """H·v for a statevector tensor v (axis k = site k+1)."""
    n = len(ws)
    h = dense(ws)
    return np.tensordot(h, v, axes=(list(range(n, 2 * n)), list(range(n))))


def apply_exact(ws, psi: oc.Chain) -> oc.Chain:
    """H|ψ⟩ as an MPS with bond (b, w) -> b + χ_b·w; ψ's Λ vectors absorbed."""
    xs = oc._absorbed(psi)
    out = []
    for w, a in zip(ws, xs):
        t = np.einsum("woiv,bic->wbovc", w, a)
        D, B, d, V, Cc = t.shape
        out.append(t.reshape(D * B, d, V * Cc))
    return oc.Chain(out, periodic=psi.periodic)


def _qr_cols(m):
    q, _ = np.linalg.qr(m)
    return q


def fit(ws, psi: oc.Chain, phi0: oc.Chain, sweeps=2):
    """One-site variational fitting of H|ψ⟩ with φ's bond dimensions; returns
    (φ, overlaps) where overlaps[s] = ⟨φ|Hψ⟩ after sweep s (centre on site 1).
    Same sweep schedule as qh_mpo_fit: right-orthonormalise φ, then per sweep
    left→right (sites 1..n-1: local tensor, left-orthonormalise) and right→left
    (sites n..2: local tensor, right-orthonormalise), then site 1 = local tensor."""
    n = len(ws)
    A = oc._absorbed(psi)
    B = [a.copy() for a in oc._absorbed(phi0)]
    d = B[0].shape[1]
    for i in range(n - 1, 0, -1):
        cl, _, cr = B[i].shape
        q = _qr_cols(B[i].reshape(cl, d * cr).conj().T)  # (o r) × l
        r = B[i].reshape(cl, d * cr) @ q
        B[i] = q.conj().T.reshape(cl, d, cr)
        B[i - 1] = np.einsum("lox,xy->loy", B[i - 1], r)
    one = np.ones((1, 1, 1), dtype=np.complex128)
    L = [None] * (n + 1)
    R = [None] * (n + 1)
    L[0] = one
    R[n] = one

    def right_env(i):  # R[i-1] from R[i], site i (1-based)
        return np.einsum("bic,woiv,aox,cxv->baw", A[i - 1], ws[i - 1], B[i - 1].conj(), R[i], optimize=_OPT)

    def left_env(i):  # L[i] from L[i-1], site i
        return np.einsum("baw,bic,woiv,aox->cxv", L[i - 1], A[i - 1], ws[i - 1], B[i - 1].conj(), optimize=_OPT)

    def local(i):
        return np.einsum("baw,bic,woiv,cxv->aox", L[i - 1], A[i - 1], ws[i - 1], R[i], optimize=_OPT)

    for i in range(n, 1, -1):
        R[i - 1] = right_env(i)
    overlaps = []
    for _ in range(sweeps):
        for i in range(1, n):
            m = local(i)
            a, _, x = m.shape
            B[i - 1] = _qr_cols(m.reshape(a * d, x)).reshape(a, d, x)
            L[i] = left_env(i)
        for i in range(n, 1, -1):
            m = local(i)
            a, _, x = m.shape
            q = _qr_cols(m.reshape(a, d * x).conj().T)
            B[i - 1] = q.conj().T.reshape(a, d, x)
            R[i - 1] = right_env(i)
        m = local(1)
        B[0] = m
        overlaps.append(complex(np.vdot(m, m)))
    return oc.Chain(B), overlaps


def local_z(psi: oc.Chain, op=None):
    """⟨ψ|O_k|ψ⟩ / ⟨ψ|ψ⟩ for a one-site operator on every site (default Z), from one
    pair of environment sweeps (the same numbers as `expect` per site,
    src/Ansatz/Chain.jl:724-735, normalised)."""
    op = np.diag([1.0, -1.0]).astype(np.complex128) if op is None else op
    xs = oc._absorbed(psi)
    n = len(xs)
    left = [np.ones((1, 1), dtype=np.complex128)]
    for a in xs:
        left.append(np.einsum("ab,aoc,bod->cd", left[-1], a.conj(), a, optimize=_OPT))
    right = [None] * (n + 1)
    right[n] = np.ones((1, 1), dtype=np.complex128)
    for k in range(n - 1, -1, -1):
        a = xs[k]
        right[k] = np.einsum("cd,aoc,bod->ab", right[k + 1], a.conj(), a, optimize=_OPT)
    nrm = left[n][0, 0]
    return np.array([np.einsum("ab,aoc,op,bpd,cd->", left[k], xs[k].conj(), op, xs[k], right[k + 1], optimize=_OPT)
                     for k in range(n)]) / nrm
