"""Open MPO on the device and MPO × MPS (exact and variationally compressed).

`MPO(arrays)` mirrors `MPO(arrays)` = `Chain(Operator(), Open(), arrays)`
(src/Ansatz/Chain.jl:20,133-172): default index order (o, i, l, r)
(Chain.jl:34), boundary arrays 3-D ((o, i, r) first, (o, i, l) last).  The
reference has no routine that applies an MPO to an MPS; `apply` and `fit_` are
this library's (C-ABI `qh_mpo_apply`, `qh_mpo_fit`).
"""
from __future__ import annotations

import ctypes as C
import weakref

import numpy as np

from . import _lib as L
from .chain import _PREC, Chain, Context, default_context
from .errors import ArgumentError, DimensionMismatch

_ORDER = ("o", "i", "l", "r")


def _oilr(a, k, n, order):
    """Array of site k (0-based) in ``order`` -> (o, i, l, r) with size-1 boundary bonds."""
    a = np.asarray(a, dtype=np.complex128)
    names = list(order)
    if k == 0:
        names = [x for x in names if x != "l"]
    if k == n - 1:
        names = [x for x in names if x != "r"]
    if a.ndim != len(names):
        want = 3 if k in (0, n - 1) else 4
        raise ArgumentError(f"site {k + 1}: array must have {want} dimensions")
    if "l" not in names:
        a = a[..., None]
        names.append("l")
    if "r" not in names:
        a = a[..., None]
        names.append("r")
    return np.transpose(a, [names.index(x) for x in _ORDER])


class MPOChain:
    """An open MPO resident in device memory (`qh_mpo_create`)."""

    def __init__(self, handle, ctx: Context, n: int, d: int, bond_dims):
        self.handle = handle
        self.ctx = ctx
        self.n = n
        self.d = d
        self._bonds = list(bond_dims)
        self._fin = weakref.finalize(self, L.load().qh_mpo_free, handle)

    def bond_dims(self):
        return list(self._bonds)

    def __repr__(self):
        return f"MPO (inputs={self.n}, outputs={self.n})"


def MPO(arrays, order=_ORDER, context: Context | None = None) -> MPOChain:
    ctx = context or default_context()
    n = len(arrays)
    if n < 2:
        raise ArgumentError("an MPO needs at least 2 sites")
    if sorted(order) != sorted(_ORDER):
        raise ArgumentError("order must be a permutation of (o, i, l, r)")
    full = [_oilr(a, k, n, order) for k, a in enumerate(arrays)]
    d = full[0].shape[0]
    for k, f in enumerate(full):
        if f.shape[0] != d or f.shape[1] != d:
            raise DimensionMismatch("all sites must share the physical dimension")
        if k < n - 1 and f.shape[3] != full[k + 1].shape[2]:
            raise DimensionMismatch(f"bond {k + 1}: {f.shape[3]} != {full[k + 1].shape[2]}")
    bufs = [np.asfortranarray(f).ravel(order="F").view(np.float64) for f in full]
    dims = [f.shape[3] for f in full[:-1]]
    cdims = (C.c_int * (n - 1))(*dims)
    ptrs = (L._dp * n)(*[b.ctypes.data_as(L._dp) for b in bufs])
    h = C.c_void_p()
    ctx.check(L.load().qh_mpo_create(ctx.handle, n, d, cdims, ptrs, C.byref(h)))
    return MPOChain(h, ctx, n, d, dims)


def pMPO(arrays, order=_ORDER, context: Context | None = None) -> MPOChain:
    """`pMPO(arrays)` = `Chain(Operator(), Periodic(), arrays)` (src/Ansatz/Chain.jl:21,102-131):
    every array 4-D; site n's right bond is site 1's left bond."""
    ctx = context or default_context()
    n = len(arrays)
    if n < 2:
        raise ArgumentError("an MPO needs at least 2 sites")
    if sorted(order) != sorted(_ORDER):
        raise ArgumentError("order must be a permutation of (o, i, l, r)")
    if not all(np.ndim(a) == 4 for a in arrays):
        raise AssertionError("All arrays must have 4 dimensions")
    full = [np.transpose(np.asarray(a, dtype=np.complex128), [list(order).index(x) for x in _ORDER]) for a in arrays]
    d = full[0].shape[0]
    for k, f in enumerate(full):
        if f.shape[0] != d or f.shape[1] != d:
            raise DimensionMismatch("all sites must share the physical dimension")
        if f.shape[3] != full[(k + 1) % n].shape[2]:
            raise DimensionMismatch(f"bond {k + 1}: {f.shape[3]} != {full[(k + 1) % n].shape[2]}")
    bufs = [np.asfortranarray(f).ravel(order="F").view(np.float64) for f in full]
    dims = [f.shape[3] for f in full]
    cdims = (C.c_int * n)(*dims)
    ptrs = (L._dp * n)(*[b.ctypes.data_as(L._dp) for b in bufs])
    h = C.c_void_p()
    ctx.check(L.load().qh_pmpo_create(ctx.handle, n, d, cdims, ptrs, C.byref(h)))
    return MPOChain(h, ctx, n, d, dims[:-1])


def apply(mpo: MPOChain, psi: Chain) -> Chain:
    """H|ψ⟩ exactly, as a new MPS with bond dimensions χ_b·D_b."""
    out = C.c_void_p()
    psi.ctx.check(L.load().qh_mpo_apply(psi.ctx.handle, mpo.handle, psi.handle, C.byref(out)))
    return Chain(out, psi.ctx)


def fit_(mpo: MPOChain, psi: Chain, phi: Chain, sweeps: int = 2, precision: str = "fp64") -> complex:
    """Variational compression of H|ψ⟩ into ``phi`` (its bond dimensions are kept);
    returns ⟨φ|Hψ⟩.  ``precision`` applies to the environment contractions."""
    ov = (C.c_double * 2)()
    psi.ctx.check(L.load().qh_mpo_fit(psi.ctx.handle, mpo.handle, psi.handle, phi.handle, sweeps, _PREC[precision], ov))
    return complex(ov[0], ov[1])


def compress(mpo: MPOChain, psi: Chain, sweeps: int = 2, precision: str = "fp64"):
    """H|ψ⟩ compressed back to ψ's bond dimensions: (φ, ⟨φ|Hψ⟩), initial guess ψ."""
    phi = psi.copy()
    ov = fit_(mpo, psi, phi, sweeps=sweeps, precision=precision)
    return phi, ov
