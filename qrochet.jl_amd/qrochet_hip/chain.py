"""Host-side mirror of Qrochet's `Chain` API over the libqrochet_hip.so C ABI.

Method names, argument meaning and error behaviour follow
`src/Ansatz/Chain.jl` / `src/Ansatz.jl` / `src/Ansatz/Dense.jl` of Qrochet.jl
v0.1.2 (cited per function).  Julia's mutating `f!` is spelled `f_` here;
non-mutating variants copy first, like the reference (`Chain.jl:299,335,378,
460,499`).  Every numerical operation runs on the GPU through the C ABI; the
MPS tensors stay resident in device memory between calls.
"""
from __future__ import annotations

import ctypes as C
import math
import weakref
from dataclasses import dataclass

import numpy as np

from . import _lib as L
from .errors import (
    AdmissionError,
    ArgumentError,
    DimensionMismatch,
    MissingSchmidtCoefficientsException,
    QHError,
)


# ---------------------------------------------------------------------------
# sites, sockets, operators  (src/Quantum.jl:10-43, src/Ansatz/Dense.jl)
# ---------------------------------------------------------------------------
@dataclass(frozen=True)
class Site:
    id: int
    dual: bool = False

    def adjoint(self) -> "Site":
        return Site(self.id, not self.dual)

    def __repr__(self):
        return f"{self.id}{chr(39) if self.dual else ''}"


def site(s: str) -> Site:
    """`site"…"` string macro (src/Quantum.jl:29-43), 1-D sites only."""
    dual = s.endswith("'")
    return Site(int(s.rstrip("'")), dual)


class Socket:
    pass


class State(Socket):
    def __repr__(self):
        return "State()"


class Operator(Socket):
    def __repr__(self):
        return "Operator()"


class Dense:
    """`Dense(::Operator, array; sites)` (src/Ansatz/Dense.jl:21-34).

    Axis k of ``array`` is attached to ``sites[k]``; non-dual sites are outputs,
    dual sites inputs.
    """

    def __init__(self, socket, array, sites):
        array = np.asarray(array)
        assert array.ndim > 0
        assert all(s > 1 for s in array.shape)
        sites = [s if isinstance(s, Site) else Site(int(s)) for s in sites]
        assert len(sites) == array.ndim
        self.socket = socket
        self.array = array
        self.sites = sites

    @property
    def outputs(self):
        return [s for s in self.sites if not s.dual]

    @property
    def inputs(self):
        return [s for s in self.sites if s.dual]

    def lanes(self):
        return sorted({s.id for s in self.sites})

    def canonical(self) -> np.ndarray:
        """Column-major flat buffer (out_1..out_k, in_1..in_k), lanes ascending."""
        lanes = self.lanes()
        m = len(lanes)
        g = np.asarray(self.array, dtype=np.complex128)
        if len(self.sites) == 2 * m and all(s.id == lanes[k % m] and s.dual == (k >= m)
                                            for k, s in enumerate(self.sites)):
            flat = np.asfortranarray(g).ravel(order="F")  # already in canonical order
        else:
            try:
                order = ([self.sites.index(Site(i, False)) for i in lanes] +
                         [self.sites.index(Site(i, True)) for i in lanes])
            except ValueError:
                raise ArgumentError("Gate inputs and outputs must be the same") from None
            flat = np.asfortranarray(np.transpose(g, order)).ravel(order="F")
        return flat.copy() if np.may_share_memory(flat, self.array) else flat

    def copy(self):
        return Dense(self.socket, self.array.copy(), list(self.sites))


_SITES: dict = {}


def _site_of(i, dual):
    """Site(i, dual), shared (sites are immutable)."""
    key = (int(i), bool(dual))
    s = _SITES.get(key)
    if s is None:
        s = _SITES[key] = Site(*key)
    return s


def gate(array_or_matrix, *lanes) -> Dense:
    """Operator on ``lanes``: a (d,…,d) array in (outs…, ins…) order, or a
    (d^k × d^k) matrix whose row/column index has the first lane fastest
    (Julia column-major `reshape`)."""
    a = np.asarray(array_or_matrix, dtype=np.complex128)
    k = len(lanes)
    if a.ndim == 2 and k == 2:
        d = int(round(math.sqrt(a.shape[0])))
        a = a.reshape((d, d, d, d), order="F")
    sites = [_site_of(i, False) for i in lanes] + [_site_of(i, True) for i in lanes]
    return Dense(Operator(), a, sites)


# ---------------------------------------------------------------------------
# context
# ---------------------------------------------------------------------------
def _raise(rc, ctx_handle, bond=None):
    lib = L.load()
    msg = lib.qh_last_error(ctx_handle)
    msg = msg.decode() if msg else ""
    if rc == L.QH_ERR_ARGUMENT:
        raise ArgumentError(msg)
    if rc == L.QH_ERR_MISSING_SCHMIDT:
        raise MissingSchmidtCoefficientsException(bond, msg)
    if rc == L.QH_ERR_DIMENSION:
        raise DimensionMismatch(msg)
    if rc == L.QH_ERR_ASSERTION:
        raise AssertionError(msg)
    if rc == L.QH_ERR_ADMISSION:
        raise AdmissionError(rc, msg)
    raise QHError(rc, msg)


class Context:
    """One HIP stream on one device (`qh_ctx_create`)."""

    def __init__(self, device: int = 0):
        lib = L.load()
        h = C.c_void_p()
        rc = lib.qh_ctx_create(device, C.byref(h))
        if rc != 0:
            _raise(rc, None)
        self.handle = h
        self.device = device
        self._fin = weakref.finalize(self, lib.qh_ctx_destroy, h)

    def check(self, rc, bond=None):
        if rc != 0:
            _raise(rc, self.handle, bond)

    def synchronize(self):
        self.check(L.load().qh_ctx_synchronize(self.handle))

    def set_timing(self, on: bool):
        self.check(L.load().qh_ctx_set_timing(self.handle, 1 if on else 0))

    def reset_timing(self):
        self.check(L.load().qh_ctx_reset_timing(self.handle))

    def timing(self, cls: str):
        ms = C.c_double()
        n = C.c_int64()
        fl = C.c_double()
        by = C.c_double()
        self.check(
            L.load().qh_ctx_get_timing(self.handle, cls.encode(), C.byref(ms), C.byref(n), C.byref(fl), C.byref(by))
        )
        return dict(ms=ms.value, launches=n.value, flops=fl.value, bytes=by.value)

    def set_jacobi(self, block=0, max_sweeps=0):
        self.check(L.load().qh_ctx_set_jacobi(self.handle, block, max_sweeps))

    def set_qr_passes(self, passes: int):
        """Householder QR passes preconditioning the Jacobi SVD (0..4, default 3)."""
        self.check(L.load().qh_ctx_set_qr_passes(self.handle, int(passes)))

    def autotune(self, max_q=512, batch=32):
        """Tune the batched SVD's schedule (Jacobi block width, QR passes) per size class
        up to ``max_q`` columns (`qh_ctx_autotune`); returns one dict per class."""
        rep = (C.c_double * 40)()
        n = C.c_int()
        self.check(L.load().qh_ctx_autotune(self.handle, int(max_q), int(batch), rep, C.byref(n)))
        return [dict(q=int(rep[5 * k]), block=int(rep[5 * k + 1]), qr_passes=int(rep[5 * k + 2]),
                     best_ms=rep[5 * k + 3], previous_ms=rep[5 * k + 4]) for k in range(n.value)]

    def set_graphs(self, enable: bool):
        """Replay each Jacobi sweep's rounds as a captured HIP graph (default) or launch
        them one by one (`qh_ctx_set_graphs`)."""
        self.check(L.load().qh_ctx_set_graphs(self.handle, 1 if enable else 0))

    def set_pivoting(self, enable: bool):
        """First QR pass of the SVD's preconditioner on the columns ordered by decreasing
        norm (default) or in their own order (`qh_ctx_set_pivoting`)."""
        self.check(L.load().qh_ctx_set_pivoting(self.handle, 1 if enable else 0))

    def autotune_reset(self):
        self.check(L.load().qh_ctx_autotune_reset(self.handle))

    def mem_info(self):
        """(free, total) bytes of the device."""
        f = C.c_size_t()
        t = C.c_size_t()
        self.check(L.load().qh_ctx_mem_info(self.handle, C.byref(f), C.byref(t)))
        return f.value, t.value


_default = {}


def default_context(device: int = 0) -> Context:
    if device not in _default:
        _default[device] = Context(device)
    return _default[device]


# ---------------------------------------------------------------------------
# Chain
# ---------------------------------------------------------------------------
def _sid(s) -> int:
    return s.id if isinstance(s, Site) else int(s)


def _olr(a, i, n, order):
    """Reference-shaped site array -> (o, l, r) Fortran-ordered complex128."""
    a = np.asarray(a, dtype=np.complex128)
    names = list(order)
    if i == 0:
        names = [x for x in names if x != "l"]
    if i == n - 1:
        names = [x for x in names if x != "r"]
    if a.ndim != len(names):
        raise AssertionError(
            "First array must have 2 dimensions" if i == 0 else
            "Last array must have 2 dimensions" if i == n - 1 else "All arrays must have 3 dimensions"
        )
    full = a
    if "l" not in names:
        full = full[..., None]
        names = names + ["l"]
    if "r" not in names:
        full = full[..., None]
        names = names + ["r"]
    full = np.transpose(full, [names.index("o"), names.index("l"), names.index("r")])
    return np.asfortranarray(full)


class Chain:
    """`Chain` with `State()` socket and `Open()` boundary (src/Ansatz/Chain.jl:6-100),
    resident in device memory.  Build with `MPS(arrays)`."""

    def __init__(self, handle, ctx: Context):
        self.handle = handle
        self.ctx = ctx
        lib = L.load()
        n = C.c_int()
        d = C.c_int()
        lib.qh_mps_nsites(handle, C.byref(n), C.byref(d))
        self.n = n.value
        self.d = d.value
        per = C.c_int()
        wrap = C.c_int()
        lib.qh_mps_is_periodic(handle, C.byref(per), C.byref(wrap))
        self.periodic = bool(per.value)
        self._wrap = wrap.value
        self._fin = weakref.finalize(self, lib.qh_mps_free, handle)

    # --- queries -----------------------------------------------------------
    def nsites(self):
        return self.n

    nlanes = nsites

    def sites(self):
        return [Site(i) for i in range(1, self.n + 1)]

    def bond_dims(self):
        dims = (C.c_int * (self.n - 1))()
        has = (C.c_int * (self.n - 1))()
        L.load().qh_mps_bond_dims(self.handle, dims, has)
        return list(dims)

    def has_lambda(self):
        dims = (C.c_int * (self.n - 1))()
        has = (C.c_int * (self.n - 1))()
        L.load().qh_mps_bond_dims(self.handle, dims, has)
        return [bool(h) for h in has]

    def ntensors(self):
        """Tensors of the reference TN: sites plus Λ vectors."""
        return self.n + sum(self.has_lambda())

    def footprint(self):
        b = C.c_size_t()
        L.load().qh_mps_footprint(self.handle, C.byref(b))
        return b.value

    def wrap_dim(self):
        """Dimension of the bond between site n and site 1 (1 for an open chain)."""
        return self._wrap

    def _site_shape(self, i):
        full = [self._wrap] + self.bond_dims() + [self._wrap]
        return self.d, full[i - 1], full[i]

    def site_array(self, i, order=("o", "l", "r")):
        """Download the tensor at site ``i`` in the reference's shape/order."""
        i = _sid(i)
        _, cl, cr = self._site_shape(i)
        buf = np.empty(self.d * cl * cr, dtype=np.complex128)
        self.ctx.check(L.load().qh_mps_get_site(self.handle, i, buf.ctypes.data_as(C.POINTER(C.c_double))))
        t = buf.reshape((self.d, cl, cr), order="F")
        names = ["o", "l", "r"]
        if i == 1 and not self.periodic:
            t = t[:, 0, :]
            names = ["o", "r"]
        if i == self.n and not self.periodic:
            t = t[..., 0]
            names = names[:-1]
        o = [x for x in order if x in names]
        return np.transpose(t, [names.index(x) for x in o])

    def arrays(self, order=("o", "l", "r")):
        return [self.site_array(i, order) for i in range(1, self.n + 1)]

    def internal_arrays(self):
        """Site tensors as (χl, d, χr) arrays (boundary bonds of size 1)."""
        out = []
        for i in range(1, self.n + 1):
            a = self.site_array(i)
            if i == 1 and not self.periodic:
                a = a[:, None, :]
            if i == self.n and not self.periodic:
                a = a[..., None]
            out.append(np.transpose(a, (1, 0, 2)))
        return out

    def lambda_(self, bond):
        """Λ vector on ``bond`` = (i, i+1) or i (`tensors(ψ; between)`, src/Ansatz.jl:78-89)."""
        b = _sid(bond[0]) if isinstance(bond, (tuple, list)) else _sid(bond)
        chi = self.bond_dims()[b - 1] if 1 <= b < self.n else 0
        buf = np.empty(max(chi, 1), dtype=np.float64)
        self.ctx.check(
            L.load().qh_mps_get_lambda(self.handle, b, buf.ctypes.data_as(C.POINTER(C.c_double))), bond=(b, b + 1)
        )
        return buf[:chi]

    def lambdas(self):
        return [self.lambda_(b) if h else None for b, h in zip(range(1, self.n), self.has_lambda())]

    def copy(self) -> "Chain":
        """`copy`/`deepcopy` (src/Ansatz/Chain.jl:11)."""
        h = C.c_void_p()
        self.ctx.check(L.load().qh_mps_clone(self.ctx.handle, self.handle, C.byref(h)))
        return Chain(h, self.ctx)

    __copy__ = copy

    def __deepcopy__(self, memo):
        return self.copy()

    def __repr__(self):
        return f"MPS (inputs=0, outputs={self.n})"


def MPS(arrays, order=("o", "l", "r"), lambdas=None, context: Context | None = None) -> Chain:
    """`MPS(arrays)` = `Chain(State(), Open(), arrays; order)` (src/Ansatz/Chain.jl:18,64-100),
    uploaded to the device.  ``lambdas`` optionally gives Λ vectors per bond (None = absent)."""
    ctx = context or default_context()
    n = len(arrays)
    if n < 2:
        raise ArgumentError("a Chain needs at least 2 sites")
    if sorted(order) != ["l", "o", "r"]:
        raise ArgumentError("order must be a permutation of (o, l, r)")
    full = [_olr(a, i, n, order) for i, a in enumerate(arrays)]
    d = full[0].shape[0]
    for i in range(n - 1):
        if full[i].shape[2] != full[i + 1].shape[1]:
            raise DimensionMismatch(f"bond {i + 1}: {full[i].shape[2]} != {full[i + 1].shape[1]}")
        if full[i].shape[0] != d:
            raise DimensionMismatch("all sites must share the physical dimension")
    dims = (C.c_int * (n - 1))(*[full[i].shape[2] for i in range(n - 1)])
    ptrs = (L._dp * n)(*[f.ctypes.data_as(L._dp) for f in full])
    lam_keep = []
    lam_ptrs = None
    if lambdas is not None:
        lam_ptrs = (L._dp * (n - 1))()
        for b, lam in enumerate(lambdas):
            if lam is None:
                lam_ptrs[b] = None
            else:
                v = np.ascontiguousarray(lam, dtype=np.float64)
                if v.shape != (full[b].shape[2],):
                    raise DimensionMismatch("Λ length must match the bond dimension")
                lam_keep.append(v)
                lam_ptrs[b] = v.ctypes.data_as(L._dp)
    h = C.c_void_p()
    ctx.check(L.load().qh_mps_create(ctx.handle, n, d, dims, ptrs, lam_ptrs, C.byref(h)))
    return Chain(h, ctx)


def rand_chain(n: int, chi: int, p: int = 2, eltype=np.float64, seed: int = 0,
               context: Context | None = None) -> Chain:
    """`rand(Chain, Open, State; n, χ, p, eltype)` (src/Ansatz/Chain.jl:223-256), sampled
    on the device: bond b of dimension min(χ, p^b, p^(n-b)), right-canonical from site
    2, norm 1 (for χ >= p).  ``seed`` fixes the sample (Julia's RNG stream is not
    reproduced)."""
    ctx = context or default_context()
    cplx = np.issubdtype(np.dtype(eltype), np.complexfloating)
    h = C.c_void_p()
    ctx.check(L.load().qh_mps_rand(ctx.handle, int(n), int(chi), int(p), 1 if cplx else 0,
                                   C.c_ulonglong(int(seed) & 0xFFFFFFFFFFFFFFFF), C.byref(h)))
    return Chain(h, ctx)


def pMPS(arrays, order=("o", "l", "r"), context: Context | None = None) -> Chain:
    """`pMPS(arrays)` = `Chain(State(), Periodic(), arrays; order)` (src/Ansatz/Chain.jl:19,36-62):
    every array 3-D; site n's right bond is site 1's left bond."""
    ctx = context or default_context()
    n = len(arrays)
    if n < 2:
        raise ArgumentError("a Chain needs at least 2 sites")
    if sorted(order) != ["l", "o", "r"]:
        raise ArgumentError("order must be a permutation of (o, l, r)")
    if not all(np.ndim(a) == 3 for a in arrays):
        raise AssertionError("All arrays must have 3 dimensions")
    full = [np.asfortranarray(np.transpose(np.asarray(a, dtype=np.complex128),
                                           [list(order).index(x) for x in ("o", "l", "r")])) for a in arrays]
    d = full[0].shape[0]
    for i in range(n):
        if full[i].shape[0] != d:
            raise DimensionMismatch("all sites must share the physical dimension")
        if full[i].shape[2] != full[(i + 1) % n].shape[1]:
            raise DimensionMismatch(f"bond {i + 1}: {full[i].shape[2]} != {full[(i + 1) % n].shape[1]}")
    dims = (C.c_int * n)(*[full[i].shape[2] for i in range(n)])
    ptrs = (L._dp * n)(*[f.ctypes.data_as(L._dp) for f in full])
    h = C.c_void_p()
    ctx.check(L.load().qh_pmps_create(ctx.handle, n, d, dims, ptrs, C.byref(h)))
    return Chain(h, ctx)


def from_internal(arrays, lambdas=None, context=None) -> Chain:
    """Build from (χl, d, χr) arrays (boundary bonds of size 1)."""
    ref = []
    n = len(arrays)
    for i, a in enumerate(arrays):
        t = np.transpose(np.asarray(a), (1, 0, 2))  # (o, l, r)
        if i == 0:
            t = t[:, 0, :]
        if i == n - 1:
            t = t[..., 0]
        ref.append(t)
    return MPS(ref, lambdas=lambdas, context=context)


# ---------------------------------------------------------------------------
# algorithms  (src/Ansatz/Chain.jl:299-752)
# ---------------------------------------------------------------------------
_PREC = {"fp64": L.QH_PREC_FP64, "fp32": L.QH_PREC_FP32, "bf16x3": L.QH_PREC_BF16X3}
_DIR = {"left": L.QH_LEFT, "right": L.QH_RIGHT}
_METHOD = {"qr": L.QH_METHOD_QR, "svd": L.QH_METHOD_SVD}


def _lanes_checked(psi: Chain, gate: Dense):
    """Gate validation of `evolve!` (src/Ansatz/Chain.jl:551-581)."""
    if not isinstance(gate.socket, Operator):
        raise ArgumentError(f"Gate must be an operator, but got {gate.socket}")
    ins = {s.id for s in gate.inputs}
    outs = {s.id for s in gate.outputs}
    if ins != outs:
        raise ArgumentError(f"Gate inputs ({sorted(ins)}) and outputs ({sorted(outs)}) must be the same")
    if not all(1 <= i <= psi.n for i in ins):
        raise ArgumentError(f"Gate inputs ({sorted(ins)}) must be a subset of the TN sites")
    lanes = sorted(ins)
    if len(lanes) == 2:
        if lanes[1] != lanes[0] + 1:
            raise ArgumentError("Gate lanes must be contiguous")
    elif len(lanes) != 1:
        raise ArgumentError(f"Invalid number of lanes {len(lanes)}, maximum is 2")
    return lanes


def _trunc_args(threshold, maxdim):
    thr = float("nan") if threshold is None else float(threshold)
    md = -1 if maxdim is None else int(maxdim)
    if maxdim is not None and md < 1:
        raise ArgumentError("maxdim must be positive")
    return thr, md


def evolve_(psi: Chain, gate: Dense, threshold=None, maxdim=None, iscanonical=False, renormalize=False,
            precision="fp64", info=None) -> Chain:
    """`evolve!(qtn, gate; threshold, maxdim, iscanonical, renormalize)` (src/Ansatz/Chain.jl:543-584).

    ``info`` (a dict), if given, receives ``kept`` and ``discarded`` of a 2-site update."""
    lanes = _lanes_checked(psi, gate)
    lib = L.load()
    g = gate.canonical()
    gp = g.ctypes.data_as(L._dp)
    if len(lanes) == 1:
        psi.ctx.check(lib.qh_evolve1(psi.ctx.handle, psi.handle, lanes[0], gp))
        return psi
    thr, md = _trunc_args(threshold, maxdim)
    flags = _PREC[precision]
    if iscanonical:
        flags |= L.QH_CANONICAL
    if renormalize:
        flags |= L.QH_RENORMALIZE
    kept = C.c_int()
    disc = C.c_double()
    want = info is not None
    psi.ctx.check(
        lib.qh_evolve2(psi.ctx.handle, psi.handle, lanes[0], gp, thr, md, flags,
                       C.byref(kept) if want else None, C.byref(disc) if want else None),
        bond=(lanes[0], lanes[0] + 1),
    )
    if want:
        info["kept"] = kept.value
        info["discarded"] = disc.value
    return psi


def simple_update_(psi: Chain, gate: Dense, threshold=None, maxdim=None, renormalize=False, precision="fp64"):
    """Vidal-form TEBD step: `evolve!(…; iscanonical=true)` (src/Ansatz/Chain.jl:606-722)."""
    return evolve_(psi, gate, threshold=threshold, maxdim=maxdim, iscanonical=True, renormalize=renormalize,
                   precision=precision)


def evolve_layer_(updates, threshold=None, maxdim=None, iscanonical=False, renormalize=False, precision="fp64",
                  context: Context | None = None):
    """A brickwork layer: ``updates`` is a list of (Chain, 2-site Dense) on disjoint
    sites, applied in one batched pipeline (`qh_evolve2_batch`).  ``threshold`` and
    ``maxdim`` are either one value for every update or a list with one entry per
    update (None = nothing).  Returns (kept, discarded)."""
    if not updates:
        return [], []
    ctx = context or updates[0][0].ctx
    n = len(updates)
    handles = (C.c_void_p * n)()
    bonds = (C.c_int * n)()
    gates = []
    for i, (psi, g) in enumerate(updates):
        lanes = _lanes_checked(psi, g)
        if len(lanes) != 2:
            raise ArgumentError("layer updates must be 2-site gates")
        handles[i] = psi.handle
        bonds[i] = lanes[0]
        gates.append(g.canonical())
    gbuf = np.ascontiguousarray(np.concatenate(gates))
    ths = threshold if isinstance(threshold, (list, tuple, np.ndarray)) else [threshold] * n
    mds = maxdim if isinstance(maxdim, (list, tuple, np.ndarray)) else [maxdim] * n
    if len(ths) != n or len(mds) != n:
        raise ArgumentError("one threshold / maxdim per update")
    tv = (C.c_double * n)()
    mv = (C.c_int * n)()
    for i in range(n):
        tv[i], mv[i] = _trunc_args(ths[i], mds[i])
    flags = _PREC[precision] | (L.QH_CANONICAL if iscanonical else 0) | (L.QH_RENORMALIZE if renormalize else 0)
    kept = (C.c_int * n)()
    disc = (C.c_double * n)()
    ctx.check(L.load().qh_evolve2_batch(ctx.handle, n, handles, bonds, gbuf.ctypes.data_as(L._dp), tv, mv, flags,
                                        kept, disc))
    return list(kept), list(disc)


def truncate_(psi: Chain, bond, threshold=None, maxdim=None) -> Chain:
    """`truncate!(qtn, bond; threshold, maxdim)` (src/Ansatz/Chain.jl:390-422)."""
    b1, b2 = _sid(bond[0]), _sid(bond[1])
    if not (1 <= b1 < psi.n and b2 == b1 + 1):
        raise ArgumentError(f"Invalid bond ({b1}, {b2})")
    thr, md = _trunc_args(threshold, maxdim)
    psi.ctx.check(L.load().qh_truncate(psi.ctx.handle, psi.handle, b1, thr, md), bond=(b1, b2))
    return psi


def truncate(psi: Chain, bond, **kw) -> Chain:
    return truncate_(psi.copy(), bond, **kw)


def contract_between_(psi: Chain, site1, site2, direction="left", delete_lambda=True) -> Chain:
    """`contract!(tn, :between, site1, site2; direction, delete_Λ)` (src/Ansatz/Chain.jl:309-333)."""
    dcode = _DIR.get(direction, -1)
    psi.ctx.check(L.load().qh_contract_between(psi.ctx.handle, psi.handle, _sid(site1), _sid(site2), dcode,
                                               1 if delete_lambda else 0))
    return psi


def contract_between(psi: Chain, site1, site2, **kw) -> Chain:
    return contract_between_(psi.copy(), site1, site2, **kw)


def canonize_site_(psi: Chain, s, direction, method="qr") -> Chain:
    """`canonize_site!(tn, site; direction, method)` (src/Ansatz/Chain.jl:339-376)."""
    psi.ctx.check(L.load().qh_canonize_site(psi.ctx.handle, psi.handle, _sid(s), _DIR.get(direction, -1),
                                            _METHOD.get(method, -1)))
    return psi


def canonize_site(psi: Chain, s, **kw) -> Chain:
    return canonize_site_(psi.copy(), s, **kw)


def canonize_(psi: Chain) -> Chain:
    """`canonize!(tn)` (src/Ansatz/Chain.jl:469-497): Vidal form."""
    psi.ctx.check(L.load().qh_canonize(psi.ctx.handle, psi.handle))
    return psi


def canonize(psi: Chain) -> Chain:
    return canonize_(psi.copy())


def mixed_canonize_(psi: Chain, center) -> Chain:
    """`mixed_canonize!(tn, center)` (src/Ansatz/Chain.jl:509-524)."""
    psi.ctx.check(L.load().qh_mixed_canonize(psi.ctx.handle, psi.handle, _sid(center)))
    return psi


def mixed_canonize(psi: Chain, center) -> Chain:
    return mixed_canonize_(psi.copy(), center)


def normalize_(psi: Chain, root) -> Chain:
    """`normalize!(tn, root)` (src/Ansatz/Chain.jl:532-536)."""
    psi.ctx.check(L.load().qh_normalize(psi.ctx.handle, psi.handle, _sid(root)))
    return psi


def _canon(psi, s, atol, fn):
    res = C.c_int()
    dev = C.c_double()
    psi.ctx.check(fn(psi.ctx.handle, psi.handle, _sid(s), float(atol), C.byref(res), C.byref(dev)))
    return bool(res.value)


def isleftcanonical(psi: Chain, s, atol=1e-12) -> bool:
    """`isleftcanonical(qtn, site; atol)` (src/Ansatz/Chain.jl:424-440)."""
    return _canon(psi, s, atol, L.load().qh_is_left_canonical)


def isrightcanonical(psi: Chain, s, atol=1e-12) -> bool:
    """`isrightcanonical(qtn, site; atol)` (src/Ansatz/Chain.jl:442-458)."""
    return _canon(psi, s, atol, L.load().qh_is_right_canonical)


def _observables(psi: Chain, observables):
    keep = []
    arr = (L.QHObservable * max(len(observables), 1))()
    for k, o in enumerate(observables):
        lanes = _lanes_checked(psi, o)
        buf = o.canonical()
        keep.append(buf)
        arr[k].nsites = len(lanes)
        arr[k].site = lanes[0]
        arr[k].op = buf.ctypes.data_as(L._dp)
    return arr, keep


def expect(psi: Chain, observables) -> complex:
    """`expect(ψ, observables)` (src/Ansatz/Chain.jl:724-735): unnormalised ⟨ψ|O…|ψ⟩."""
    arr, keep = _observables(psi, list(observables))
    out = (C.c_double * 2)()
    psi.ctx.check(L.load().qh_expect(psi.ctx.handle, psi.handle, len(keep), arr, out))
    return complex(out[0], out[1])


def expect_many(psi: Chain, terms) -> np.ndarray:
    """``[expect(ψ, t) for t in terms]`` from one pair of environment sweeps."""
    flat = []
    offsets = [0]
    for t in terms:
        flat.extend(t)
        offsets.append(len(flat))
    arr, keep = _observables(psi, flat)
    offs = (C.c_int * len(offsets))(*offsets)
    out = np.zeros(2 * max(len(terms), 1), dtype=np.float64)
    psi.ctx.check(L.load().qh_expect_many(psi.ctx.handle, psi.handle, len(terms), offs, arr,
                                          out.ctypes.data_as(L._dp)))
    return out[0::2][: len(terms)] + 1j * out[1::2][: len(terms)]


def overlap(a: Chain, b: Chain) -> complex:
    """`overlap(a, b)` (src/Ansatz/Chain.jl:737-748): Σ a·conj(b) = ⟨b|a⟩."""
    out = (C.c_double * 2)()
    a.ctx.check(L.load().qh_overlap(a.ctx.handle, a.handle, b.handle, out))
    return complex(out[0], out[1])


def overlap_noalloc(a: Chain, b: Chain) -> complex:
    """`overlap(a, b)` with no device allocation (`qh_overlap_noalloc`: the
    environments live in the library's fixed per-GPU scratch; open chains only)."""
    out = (C.c_double * 2)()
    a.ctx.check(L.load().qh_overlap_noalloc(a.ctx.handle, a.handle, b.handle, out))
    return complex(out[0], out[1])


def norm(psi: Chain, p=2) -> float:
    """`norm(ψ, p=2)` (src/Ansatz.jl:101-109)."""
    if p != 2:
        raise ArgumentError("only L2-norm is implemented yet")
    out = C.c_double()
    psi.ctx.check(L.load().qh_norm(psi.ctx.handle, psi.handle, C.byref(out)))
    return out.value
