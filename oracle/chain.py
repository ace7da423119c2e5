"""CPU oracle: a complex-FP64 restatement of Qrochet's MPS (`Chain`) update path.

TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library imports this module;
only `tests/`, `__graft_entry__.smoke()` and the `cpu_baseline` leg of
`bench.py` use it, and only as the checker / the timed CPU restatement.

The reference is Julia (Qrochet.jl v0.1.2 on Tenet 0.6 / LinearAlgebra).  Its
arithmetic lives in Tenet and LAPACK, which are not available here, so this
module restates the semantics of `src/Ansatz/Chain.jl` with numpy/scipy
(LAPACK zgesdd for SVD -- Julia's `svd` default -- and Householder QR).

Representation (open boundary only, like the reference's algorithms):
  * ``arrays[i]``: complex128 array of shape ``(chi_l, d, chi_r)`` for site
    ``i+1``; the first site has ``chi_l == 1`` and the last ``chi_r == 1``
    (the reference's 2-D boundary tensors, `Chain.jl:64-100`).
  * ``lam[b]``: ``None`` or a real vector on bond ``b`` (between sites b+1 and
    b+2).  It is the "singular values on a hyperindex" tensor that `svd!`
    leaves behind (`Chain.jl:338`); the state is
    ``A_1 diag(lam_1) A_2 diag(lam_2) ... A_n`` with absent lam = identity.

Public functions take 1-based site ids like the reference.

Parity status of each function (how it is pinned) is listed in
`tests/test_oracle.py`; functions whose behaviour depends on Tenet internals
that cannot be checked here are marked "parity unpinned" below.
"""
from __future__ import annotations

import copy as _copy
from dataclasses import dataclass, field

import numpy as np
import scipy.linalg as sla


class ArgumentError(ValueError):
    """Julia `ArgumentError`."""


class MissingSchmidtCoefficientsException(RuntimeError):
    """`src/Ansatz.jl:91-99`."""

    def __init__(self, bond):
        super().__init__(f"Can't access the spectrum on bond {tuple(bond)}")
        self.bond = tuple(bond)


# ---------------------------------------------------------------------------
# sites and gates
# ---------------------------------------------------------------------------
@dataclass(frozen=True)
class Site:
    """`src/Quantum.jl:10-27` (1-D sites only)."""

    id: int
    dual: bool = False

    def adjoint(self) -> "Site":
        return Site(self.id, not self.dual)


@dataclass
class Dense:
    """Operator container, `src/Ansatz/Dense.jl:21-34`.

    ``array`` axis k is attached to ``sites[k]``; outputs are non-dual sites,
    inputs are dual sites (`ext/QrochetQuacExt.jl:11`: ``[outs..., ins...]``).
    ``socket`` is "operator" for gates; "state" exists only to exercise the
    validation in `evolve!` (`Chain.jl:552-554`).
    """

    array: np.ndarray
    sites: list
    socket: str = "operator"

    @property
    def outputs(self):
        return [s for s in self.sites if not s.dual]

    @property
    def inputs(self):
        return [s for s in self.sites if s.dual]

    def canonical(self):
        """Gate array permuted to (out_1..out_k, in_1..in_k), lanes ascending."""
        lanes = sorted({s.id for s in self.sites})
        order = [self.sites.index(Site(i, False)) for i in lanes] + [
            self.sites.index(Site(i, True)) for i in lanes
        ]
        return np.transpose(np.asarray(self.array, dtype=np.complex128), order), lanes


def gate2(matrix_or_array, i: int) -> Dense:
    """2-site gate on sites (i, i+1) from a (d,d,d,d) array (out_l,out_r,in_l,in_r)
    or from a (d^2,d^2) matrix reshaped Julia-style (site i = fastest index)."""
    g = np.asarray(matrix_or_array, dtype=np.complex128)
    if g.ndim == 2:
        d = int(round(np.sqrt(g.shape[0])))
        g = g.reshape((d, d, d, d), order="F")
    return Dense(g, [Site(i), Site(i + 1), Site(i, True), Site(i + 1, True)])


def gate1(matrix, i: int) -> Dense:
    return Dense(np.asarray(matrix, dtype=np.complex128), [Site(i), Site(i, True)])


# ---------------------------------------------------------------------------
# the chain
# ---------------------------------------------------------------------------
@dataclass
class Chain:
    arrays: list
    lam: list = field(default_factory=list)
    # Periodic boundary (`pMPS`, `Chain.jl:19,36-62`): site 1's left bond is site n's
    # right bond (dimension arrays[0].shape[0]); Λ vectors live on bonds 1..n-1 only
    periodic: bool = False

    def __post_init__(self):
        if not self.lam:
            self.lam = [None] * (len(self.arrays) - 1)
        assert len(self.lam) == len(self.arrays) - 1

    @property
    def n(self) -> int:
        return len(self.arrays)

    @property
    def d(self) -> int:
        return self.arrays[0].shape[1]

    def bond_dims(self):
        return [a.shape[2] for a in self.arrays[:-1]]

    def copy(self) -> "Chain":
        return Chain([a.copy() for a in self.arrays], [None if l is None else l.copy() for l in self.lam],
                     self.periodic)

    def ntensors(self) -> int:
        """Number of tensors in the reference TN (sites + Λ vectors)."""
        return self.n + sum(l is not None for l in self.lam)


def MPS(arrays, order=("o", "l", "r")) -> Chain:
    """`MPS(arrays)` / `Chain(::State, ::Open, arrays; order)` (`Chain.jl:18,64-100`).

    ``arrays[0]`` and ``arrays[-1]`` are 2-D, bulk arrays 3-D, with axes named by
    ``order`` (a permutation of o, l, r; boundary tensors drop the missing one).
    """
    n = len(arrays)
    assert np.ndim(arrays[0]) == 2, "First array must have 2 dimensions"
    assert all(np.ndim(a) == 3 for a in arrays[1:-1]), "All arrays must have 3 dimensions"
    assert np.ndim(arrays[-1]) == 2, "Last array must have 2 dimensions"
    if sorted(order) != ["l", "o", "r"]:
        raise ArgumentError("order must be a permutation of (o, l, r)")
    out = []
    for i, a in enumerate(arrays):
        a = np.asarray(a, dtype=np.complex128)
        if i == 0:
            o = [x for x in order if x != "l"]
        elif i == n - 1:
            o = [x for x in order if x != "r"]
        else:
            o = list(order)
        full = a
        names = list(o)
        if "l" not in names:
            full = full[..., None]
            names.append("l")
        if "r" not in names:
            full = full[..., None]
            names.append("r")
        full = np.transpose(full, [names.index("l"), names.index("o"), names.index("r")])
        out.append(np.ascontiguousarray(full))
    for i in range(n - 1):
        assert out[i].shape[2] == out[i + 1].shape[0], "bond dimension mismatch"
    return Chain(out)


def pMPS(arrays, order=("o", "l", "r")) -> Chain:
    """`pMPS(arrays)` / `Chain(::State, ::Periodic, arrays; order)` (`Chain.jl:19,36-62`):
    every array 3-D; site n's right bond is site 1's left bond."""
    assert all(np.ndim(a) == 3 for a in arrays), "All arrays must have 3 dimensions"
    if sorted(order) != ["l", "o", "r"]:
        raise ArgumentError("order must be a permutation of (o, l, r)")
    out = [np.ascontiguousarray(np.transpose(np.asarray(a, dtype=np.complex128),
                                             [list(order).index(x) for x in ("l", "o", "r")])) for a in arrays]
    for i in range(len(out)):
        assert out[i].shape[2] == out[(i + 1) % len(out)].shape[0], "bond dimension mismatch"
    return Chain(out, periodic=True)


def to_arrays(psi: Chain, order=("o", "l", "r")):
    """Inverse of `MPS`: reference-shaped arrays (boundary tensors 2-D)."""
    res = []
    for i, a in enumerate(psi.arrays):
        names = ["l", "o", "r"]
        o = list(order)
        if i == 0:
            o = [x for x in o if x != "l"]
        if i == psi.n - 1:
            o = [x for x in o if x != "r"]
        t = a
        if i == 0:
            t = t[0]
            names = ["o", "r"]
        if i == psi.n - 1:
            t = t[..., 0]
            names = names[:-1]
        res.append(np.transpose(t, [names.index(x) for x in o]))
    return res


def _open_only(psi: Chain):
    """canonize_site!/canonize!/mixed_canonize! dispatch on `::Open` only
    (`Chain.jl:336,339,461,469,500,509`)."""
    if psi.periodic:
        raise ArgumentError("defined for open chains only")


def _check_site(psi: Chain, site: int):
    if not (1 <= site <= psi.n):
        raise AssertionError(f"Site {site} not found")


# ---------------------------------------------------------------------------
# Λ absorption   (Chain.jl:309-333)
# ---------------------------------------------------------------------------
def contract_between(psi: Chain, site1: int, site2: int, direction: str = "left", delete_lambda: bool = True) -> Chain:
    """`contract!(tn, :between, site1, site2; direction, delete_Λ)` (`Chain.jl:309-333`).

    direction=:left multiplies Λ into the tensor at ``site1`` (its right
    index), :right into the tensor at ``site2`` (its left index).
    """
    _check_site(psi, site1)
    _check_site(psi, site2)
    if abs(site1 - site2) != 1:
        return psi  # no shared index -> `tensors(; between)` is `nothing`
    b = min(site1, site2) - 1
    lam = psi.lam[b]
    if lam is None:
        return psi
    if direction == "right":
        psi.arrays[site2 - 1] = psi.arrays[site2 - 1] * lam[:, None, None] if site2 > site1 else psi.arrays[site2 - 1] * lam[None, None, :]
    elif direction == "left":
        psi.arrays[site1 - 1] = psi.arrays[site1 - 1] * lam[None, None, :] if site2 > site1 else psi.arrays[site1 - 1] * lam[:, None, None]
    else:
        raise ArgumentError(f"Unknown direction=:{direction}")
    if delete_lambda:
        psi.lam[b] = None
    return psi


# ---------------------------------------------------------------------------
# factorisations
# ---------------------------------------------------------------------------
def _svd(m):
    """Thin SVD, LAPACK zgesdd (Julia `svd` default: DivideAndConquer)."""
    return sla.svd(m, full_matrices=False, lapack_driver="gesdd", check_finite=False)


def _qr(m):
    """Thin Householder QR (`LinearAlgebra.qr` -> zgeqrt); k = min(rows, cols)."""
    q, r = np.linalg.qr(m, mode="reduced")
    return q, r


def _pinv_diag(lam, atol):
    """`diag(pinv(Diagonal(λ); atol))`: entries with |λ| > atol inverted, else 0."""
    out = np.zeros_like(lam, dtype=np.float64)
    mask = np.abs(lam) > atol
    with np.errstate(divide="ignore"):
        inv = 1.0 / lam[mask]
    inv[~np.isfinite(inv)] = 0.0
    out[mask] = inv
    return out


def canonize_site(psi: Chain, site: int, direction: str, method: str = "qr") -> Chain:
    """`canonize_site!(::Open, tn, site; direction, method)` (`Chain.jl:339-376`).

    QR (or SVD) of the site tensor with the bond on the ``direction`` side as
    the column index; the R factor (or S·Vᴴ, leaving S as a Λ vector on that
    bond) is contracted -- together with any Λ already on that bond -- into
    the neighbouring tensor.
    """
    _open_only(psi)
    _check_site(psi, site)
    i = site - 1
    if direction == "left":
        if site == 1:
            raise ArgumentError("Cannot right-canonize left-most tensor")
    elif direction == "right":
        if site == psi.n:
            raise ArgumentError("Cannot left-canonize right-most tensor")
    else:
        raise ArgumentError(f"Unknown direction=:{direction}")
    if method not in ("svd", "qr"):
        raise ArgumentError(f"Unknown factorization method=:{method}")

    a = psi.arrays[i]
    cl, d, cr = a.shape
    if direction == "right":
        m = a.reshape(cl * d, cr)
        b = i  # bond to the right
        old = psi.lam[b]
        if method == "qr":
            q, r = _qr(m)
            k = q.shape[1]
            psi.arrays[i] = q.reshape(cl, d, k)
            if old is not None:
                r = r * old[None, :]
            nxt = psi.arrays[i + 1]
            psi.arrays[i + 1] = np.tensordot(r, nxt, axes=(1, 0))
            psi.lam[b] = None
        else:
            u, s, vh = _svd(m)
            k = u.shape[1]
            psi.arrays[i] = u.reshape(cl, d, k)
            if old is not None:
                vh = vh * old[None, :]
            psi.arrays[i + 1] = np.tensordot(vh, psi.arrays[i + 1], axes=(1, 0))
            psi.lam[b] = s.astype(np.float64)
    else:
        # rows (o, r), cols l   (left_inds = [rightindex, phys], right_inds = [leftindex])
        m = a.transpose(1, 2, 0).reshape(d * cr, cl)
        b = i - 1  # bond to the left
        old = psi.lam[b]
        if method == "qr":
            q, r = _qr(m)
            k = q.shape[1]
            psi.arrays[i] = q.reshape(d, cr, k).transpose(2, 0, 1).copy()
            # A_{i-1}[l', o', l] Λ[l] R[k, l]
            w = r if old is None else r * old[None, :]
            psi.arrays[i - 1] = np.tensordot(psi.arrays[i - 1], w, axes=(2, 1))
            psi.lam[b] = None
        else:
            u, s, vh = _svd(m)
            k = u.shape[1]
            psi.arrays[i] = u.reshape(d, cr, k).transpose(2, 0, 1).copy()
            w = vh if old is None else vh * old[None, :]
            psi.arrays[i - 1] = np.tensordot(psi.arrays[i - 1], w, axes=(2, 1))
            psi.lam[b] = s.astype(np.float64)
    return psi


# ---------------------------------------------------------------------------
# truncation (Chain.jl:390-422)
# ---------------------------------------------------------------------------
def _bond_index(psi: Chain, bond) -> int:
    s1, s2 = (int(bond[0]), int(bond[1]))
    if not (1 <= s1 < psi.n and s2 == s1 + 1):
        raise ArgumentError(f"Invalid bond {tuple(bond)}")
    return s1 - 1


def truncate(psi: Chain, bond, threshold=None, maxdim=None) -> Chain:
    """`truncate!(qtn, bond; threshold, maxdim)` (`Chain.jl:390-422`).

    Keeps indices ``1:min(χ, maxdim)`` (all when maxdim is nothing), then
    filters ``|s_i| > threshold`` (absolute; default 1e-16).
    """
    b = _bond_index(psi, bond)
    lam = psi.lam[b]
    if lam is None:
        raise MissingSchmidtCoefficientsException(bond)
    size = lam.shape[0]
    extent = np.arange(min(size, maxdim) if maxdim is not None else size)
    thr = 1e-16 if threshold is None else threshold
    keep = extent[np.abs(lam[extent]) > thr]
    psi.arrays[b] = psi.arrays[b][:, :, keep]
    psi.arrays[b + 1] = psi.arrays[b + 1][keep, :, :]
    psi.lam[b] = lam[keep]
    return psi


def truncated_weight(lam, threshold=None, maxdim=None):
    """sqrt of the sum of squared dropped values (Eckart-Young error of the cut)."""
    size = lam.shape[0]
    extent = np.arange(min(size, maxdim) if maxdim is not None else size)
    thr = 1e-16 if threshold is None else threshold
    keep = np.zeros(size, bool)
    keep[extent[np.abs(lam[extent]) > thr]] = True
    return float(np.sqrt(np.sum(np.abs(lam[~keep]) ** 2)))


# ---------------------------------------------------------------------------
# canonical-form predicates (Chain.jl:424-458)
# ---------------------------------------------------------------------------
def isleftcanonical(psi: Chain, site: int, atol: float = 1e-12) -> bool:
    a = psi.arrays[site - 1]
    cl, d, cr = a.shape
    m = a.reshape(cl * d, cr)
    c = m.conj().T @ m
    return bool(np.linalg.norm(c - np.eye(cr)) <= atol)


def isrightcanonical(psi: Chain, site: int, atol: float = 1e-12) -> bool:
    a = psi.arrays[site - 1]
    cl, d, cr = a.shape
    m = a.reshape(cl, d * cr)
    c = m @ m.conj().T
    return bool(np.linalg.norm(c - np.eye(cl)) <= atol)


# ---------------------------------------------------------------------------
# canonical forms (Chain.jl:460-536)
# ---------------------------------------------------------------------------
def canonize(psi: Chain) -> Chain:
    """`canonize!(::Open, tn)` (`Chain.jl:469-497`): Vidal form Γ₁Λ₁Γ₂…Γₙ."""
    _open_only(psi)
    n = psi.n
    for i in range(n, 1, -1):
        canonize_site(psi, i, "left", "qr")
    lams = []
    for i in range(1, n):
        canonize_site(psi, i, "right", "svd")
        lam = psi.lam[i - 1]
        psi.lam[i - 1] = None
        psi.arrays[i] = psi.arrays[i] * lam[:, None, None]
        lams.append(lam)
    for i in range(2, n + 1):
        lam = lams[i - 2]
        psi.arrays[i - 1] = psi.arrays[i - 1] * _pinv_diag(lam, 1e-64)[:, None, None]
        psi.lam[i - 2] = lam
    return psi


def mixed_canonize(psi: Chain, center: int) -> Chain:
    """`mixed_canonize!(::Open, tn, center)` (`Chain.jl:509-524`)."""
    _open_only(psi)
    for i in range(1, center):
        canonize_site(psi, i, "right", "qr")
    for i in range(psi.n, center, -1):
        canonize_site(psi, i, "left", "qr")
    canonize_site(psi, center, "left", "svd")
    return psi


def normalize(psi: Chain, root: int) -> Chain:
    """`normalize!(tn, root)` (`Chain.jl:532-536`): Λ on (root-1, root) ← Λ/‖Λ‖₂."""
    mixed_canonize(psi, root)
    b = root - 2
    lam = psi.lam[b]
    psi.lam[b] = lam / np.linalg.norm(lam)
    return psi


# ---------------------------------------------------------------------------
# gate application (Chain.jl:543-722)
# ---------------------------------------------------------------------------
def _validate_gate(psi: Chain, gate: Dense):
    """Checks of `evolve!` (`Chain.jl:551-581`)."""
    if gate.socket != "operator":
        raise ArgumentError(f"Gate must be an operator, but got {gate.socket}")
    ins = {s.id for s in gate.inputs}
    outs = {s.id for s in gate.outputs}
    if ins != outs:
        raise ArgumentError("Gate inputs and outputs must be the same")
    if not ins <= set(range(1, psi.n + 1)):
        raise ArgumentError("Gate inputs must be a subset of the TN sites")
    lanes = sorted(ins)
    if len(lanes) == 2 and lanes[1] != lanes[0] + 1:
        raise ArgumentError("Gate lanes must be contiguous")
    if len(lanes) > 2 or len(lanes) == 0:
        raise ArgumentError(f"Invalid number of lanes {len(lanes)}, maximum is 2")
    return lanes


def evolve(psi: Chain, gate: Dense, threshold=None, maxdim=None, iscanonical=False, renormalize=False) -> Chain:
    """`evolve!(qtn, gate; threshold, maxdim, iscanonical, renormalize)` (`Chain.jl:543-584`)."""
    lanes = _validate_gate(psi, gate)
    g, _ = gate.canonical()
    if len(lanes) == 1:
        _evolve_1site(psi, lanes[0], g)
    else:
        _evolve_2site(psi, lanes[0], g, threshold, maxdim, iscanonical, renormalize)
    return psi


def simple_update(psi: Chain, gate: Dense, threshold=None, maxdim=None, renormalize=False) -> Chain:
    """Vidal-form TEBD step = `evolve!(…; iscanonical=true)` (`Chain.jl:606-722`)."""
    return evolve(psi, gate, threshold=threshold, maxdim=maxdim, iscanonical=True, renormalize=renormalize)


def _evolve_1site(psi: Chain, site: int, g):
    """`evolve_1site!` (`Chain.jl:586-603`): A'[l,o,r] = Σ_o' G[o,o'] A[l,o',r]."""
    a = psi.arrays[site - 1]
    psi.arrays[site - 1] = np.einsum("ab,lbr->lar", g, a)


def theta(psi: Chain, site: int, g=None, iscanonical=False):
    """Two-site wave function on bond (site, site+1), optionally gated.

    Mode A (`Chain.jl:616`): contracts every tensor on the bond index, i.e.
    A_l, the bond's own Λ (if any) and A_r -- neighbouring Λ's are excluded.
    Mode B (`contract_2sitewf!`, `Chain.jl:669-685`): Λ_{i-1} Γ_i Λ_i Γ_{i+1} Λ_{i+1}.
    Returns θ with axes (l, o_l, o_r, r).
    """
    i = site - 1
    al = psi.arrays[i]
    ar = psi.arrays[i + 1]
    lam = psi.lam[i]
    if lam is not None:
        al = al * lam[None, None, :]
    if iscanonical:
        if i >= 1 and psi.lam[i - 1] is not None:
            al = al * psi.lam[i - 1][:, None, None]
        if i + 1 <= psi.n - 2 and psi.lam[i + 1] is not None:
            ar = ar * psi.lam[i + 1][None, None, :]
    th = np.tensordot(al, ar, axes=(2, 0))  # (l, o1, o2, r)
    if g is not None:
        # θ'[l,a,b,r] = Σ G[a,b,c,e] θ[l,c,e,r]  (gate inputs contract the state's sites)
        th = np.einsum("abce,lcer->labr", g, th)
    return th


def _evolve_2site(psi: Chain, site: int, g, threshold, maxdim, iscanonical, renormalize):
    """`evolve_2site!` (`Chain.jl:606-661`) incl. `unpack_2sitewf!` (`:693-722`)."""
    i = site - 1
    th = theta(psi, site, g, iscanonical)
    cl, d, _, cr = th.shape
    u, s, vh = _svd(th.reshape(cl * d, d * cr))
    k = s.shape[0]
    if iscanonical:
        if i >= 1 and psi.lam[i - 1] is not None:
            u = u * np.repeat(_pinv_diag(psi.lam[i - 1], 1e-32), d)[:, None]
        if i + 1 <= psi.n - 2 and psi.lam[i + 1] is not None:
            vh = vh * np.tile(_pinv_diag(psi.lam[i + 1], 1e-32), d)[None, :]
    psi.arrays[i] = u.reshape(cl, d, k)
    psi.arrays[i + 1] = vh.reshape(k, d, cr)
    psi.lam[i] = s.astype(np.float64)
    if threshold is not None or maxdim is not None:
        truncate(psi, (site, site + 1), threshold=threshold, maxdim=maxdim)
        if renormalize and iscanonical:
            psi.lam[i] = psi.lam[i] / np.linalg.norm(psi.lam[i])
        elif renormalize:
            normalize(psi, site)


# ---------------------------------------------------------------------------
# contractions (Chain.jl:724-752, Ansatz.jl:101-109)
# ---------------------------------------------------------------------------
def _absorbed(psi: Chain):
    """Site tensors with each bond's Λ multiplied into its left site."""
    out = []
    for i, a in enumerate(psi.arrays):
        if i < psi.n - 1 and psi.lam[i] is not None:
            a = a * psi.lam[i][None, None, :]
        out.append(a)
    return out


def overlap(a: Chain, b: Chain) -> complex:
    """`overlap(a, b)` (`Chain.jl:737-748`): Σ a·conj(b) = ⟨b|a⟩ (full contraction)."""
    if a.n != b.n:
        raise AssertionError("Ansatzes must have the same sites")
    xa, xb = _absorbed(a), _absorbed(b)
    if a.periodic or b.periodic:
        # carry the two wrap indices (bra p, ket q) through the chain, then trace
        wb, wa = xb[0].shape[0], xa[0].shape[0]
        e = np.einsum("pl,qm->pqlm", np.eye(wb), np.eye(wa)).astype(np.complex128)
        for ta, tb in zip(xa, xb):
            t = np.tensordot(e, ta, axes=(3, 0))  # (p, q, l_b, o, r_a)
            e = np.einsum("lor,pqlos->pqrs", tb.conj(), t)
        return complex(np.einsum("pqpq->", e))
    e = np.ones((1, 1), dtype=np.complex128)  # e[l_b, l_a]
    for ta, tb in zip(xa, xb):
        t = np.tensordot(e, ta, axes=(1, 0))  # (l_b, o, r_a)
        e = np.tensordot(tb.conj(), t, axes=([0, 1], [0, 1]))  # (r_b, r_a)
    return complex(e[0, 0])


def norm(psi: Chain) -> float:
    """`norm(ψ)` (`Ansatz.jl:101-109`): |sqrt(⟨ψ|ψ⟩)|."""
    return float(abs(np.sqrt(overlap(psi, psi))))


def expect(psi: Chain, observables) -> complex:
    """`expect(ψ, observables)` (`Chain.jl:724-735`): apply every observable to a
    copy with `evolve!` (defaults: no truncation, mode A) and contract with ψ'.
    Unnormalised ⟨ψ|O_k…O_1|ψ⟩."""
    phi = psi.copy()
    for o in observables:
        evolve(phi, o)
    return overlap(phi, psi)


def statevector(psi: Chain) -> np.ndarray:
    """Full contraction incl. Λ's (`transform(tn, HyperFlatten())` + `contract`).
    Axis k is site k+1.  Periodic chains trace over the wrap bond."""
    xs = _absorbed(psi)
    if psi.periodic:
        v = xs[0]  # (w, o, r)
        for t in xs[1:]:
            v = np.tensordot(v, t, axes=(v.ndim - 1, 0))
        return np.trace(v, axis1=0, axis2=v.ndim - 1)
    v = xs[0][0]  # (o, r)
    for t in xs[1:]:
        v = np.tensordot(v, t, axes=(v.ndim - 1, 0))
    return v[..., 0]


# ---------------------------------------------------------------------------
# sampler (Chain.jl:223-256) -- structure only; Julia's RNG stream is not
# reproducible here (parity unpinned for values, pinned for properties)
# ---------------------------------------------------------------------------
def _gramschmidt_rows(a):
    """Orthonormalise the rows of ``a`` (the role of `Muscle.gramschmidt!`)."""
    q, _ = np.linalg.qr(a.conj().T, mode="reduced")
    return q.conj().T


def rand_chain(n: int, chi: int, p: int = 2, dtype=np.float64, rng=None) -> Chain:
    """`rand(Chain, Open, State; n, χ, p, eltype)` (`Chain.jl:223-256`)."""
    rng = np.random.default_rng() if rng is None else rng
    arrays = []
    for i in range(1, n + 1):
        after_mid = i > n // 2
        ii = (n + 1 - abs(2 * i - n - 1)) // 2
        cl = min(chi, p ** (ii - 1))
        cr = min(chi, p**ii)
        if n % 2 == 1 and i == n // 2 + 1:
            cl, cr = cl, cl
        elif after_mid:
            cl, cr = cr, cl
        if i == 1:
            cl, cr = cr, 1
        if np.issubdtype(dtype, np.complexfloating):
            raw = rng.random((cl, cr * p)) + 1j * rng.random((cl, cr * p))
        else:
            raw = rng.random((cl, cr * p))
        a = _gramschmidt_rows(raw).reshape(cl, cr, p, order="F")
        arrays.append(np.transpose(a, (2, 0, 1)))  # (p, χl, χr)
    arrays[0] = arrays[0].reshape(p, p, order="F")
    arrays[-1] = arrays[-1].reshape(p, p, order="F")
    arrays[0] = arrays[0] / np.sqrt(p)
    return MPS(arrays)


# ---------------------------------------------------------------------------
# non-mutating variants (Chain.jl:299,335,378,460,499)
# ---------------------------------------------------------------------------
def canonize_copy(psi):
    return canonize(psi.copy())


def truncate_copy(psi, bond, **kw):
    return truncate(psi.copy(), bond, **kw)


def canonize_site_copy(psi, site, **kw):
    return canonize_site(psi.copy(), site, **kw)


def mixed_canonize_copy(psi, center):
    return mixed_canonize(psi.copy(), center)


def contract_between_copy(psi, s1, s2, **kw):
    return contract_between(psi.copy(), s1, s2, **kw)


deepcopy = _copy.deepcopy
