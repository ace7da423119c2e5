"""Gate-level circuits on a device MPS.

The reference ingests Quac gates by turning each into a `Dense` operator on
`[Site.(lanes)..., Site.(lanes; dual=true)...]` (ext/QrochetQuacExt.jl:7-15) and
applies it with `evolve!` (ext/QrochetQuacExt.jl:15); Yao circuits become a
`Quantum` network gate by gate (ext/QrochetYaoExt.jl:15-48), `from_yao` below.
Here a circuit is a list of `Op(name, lanes, params)` (or raw matrices) and
`run_` applies it to a device MPS, scheduling the gates into layers of disjoint
lanes so that every layer of two-site gates is ONE batched update
(`qh_evolve2_batch`).

Matrix convention (the reference's, `Dense.jl:21-34`): a k-lane gate is a
d^k × d^k matrix whose row/column index has the FIRST lane as the least
significant digit, i.e. reshape(M, (d,)*2k) in column-major order gives
(out_1 … out_k, in_1 … in_k).  `CX(a, b)` has control a and target b.
"""
from __future__ import annotations

import dataclasses

import numpy as np

from .chain import Chain, evolve_, evolve_layer_, gate
from .errors import ArgumentError

_I = np.eye(2, dtype=np.complex128)
_X = np.array([[0, 1], [1, 0]], dtype=np.complex128)
_Y = np.array([[0, -1j], [1j, 0]], dtype=np.complex128)
_Z = np.diag([1.0, -1.0]).astype(np.complex128)
_H = np.array([[1, 1], [1, -1]], dtype=np.complex128) / np.sqrt(2.0)


def _controlled(u):
    """Control on the first lane (least significant), u on the second."""
    m = np.zeros((4, 4), dtype=np.complex128)
    for t_in in range(2):
        for t_out in range(2):
            m[0 + 2 * t_out, 0 + 2 * t_in] = _I[t_out, t_in]
            m[1 + 2 * t_out, 1 + 2 * t_in] = u[t_out, t_in]
    return m


def _rot(p, theta):
    return np.cos(theta / 2) * _I - 1j * np.sin(theta / 2) * p


GATES_1 = {
    "I": lambda: _I, "X": lambda: _X, "Y": lambda: _Y, "Z": lambda: _Z, "H": lambda: _H,
    "S": lambda: np.diag([1, 1j]), "Sd": lambda: np.diag([1, -1j]),
    "T": lambda: np.diag([1, np.exp(1j * np.pi / 4)]), "Td": lambda: np.diag([1, np.exp(-1j * np.pi / 4)]),
    "Rx": lambda t: _rot(_X, t), "Ry": lambda t: _rot(_Y, t), "Rz": lambda t: _rot(_Z, t),
    "U2": lambda phi, lam: np.array([[1, -np.exp(1j * lam)], [np.exp(1j * phi), np.exp(1j * (phi + lam))]])
    / np.sqrt(2.0),
}
GATES_2 = {
    "CX": lambda: _controlled(_X), "CY": lambda: _controlled(_Y), "CZ": lambda: _controlled(_Z),
    "Swap": lambda: np.array([[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]], dtype=np.complex128),
    "iSwap": lambda: np.array([[1, 0, 0, 0], [0, 0, 1j, 0], [0, 1j, 0, 0], [0, 0, 0, 1]], dtype=np.complex128),
    "Rxx": lambda t: np.cos(t / 2) * np.eye(4) - 1j * np.sin(t / 2) * np.kron(_X, _X),
    "Ryy": lambda t: np.cos(t / 2) * np.eye(4) - 1j * np.sin(t / 2) * np.kron(_Y, _Y),
    "Rzz": lambda t: np.cos(t / 2) * np.eye(4) - 1j * np.sin(t / 2) * np.kron(_Z, _Z),
}


@dataclasses.dataclass
class Op:
    name: str
    lanes: tuple
    params: tuple = ()
    matrix: np.ndarray | None = None  # explicit matrix instead of a named gate

    def array(self) -> np.ndarray:
        if self.matrix is not None:
            return np.asarray(self.matrix, dtype=np.complex128)
        table = GATES_1 if len(self.lanes) == 1 else GATES_2
        if self.name not in table:
            raise ArgumentError(f"unknown {len(self.lanes)}-lane gate {self.name!r}")
        return np.asarray(table[self.name](*self.params), dtype=np.complex128)


def _oriented(op: Op):
    """(first site, 4×4 matrix on (site, site+1)) for a two-lane gate on adjacent lanes."""
    a, b = op.lanes
    m = op.array()
    if abs(a - b) != 1:
        raise ArgumentError("Gate lanes must be contiguous")
    if a < b:
        return a, m
    sw = GATES_2["Swap"]()
    return b, sw @ m @ sw  # lanes given as (b+1, b): swap the digit order


def layers(circuit) -> list:
    """Greedy as-soon-as-possible schedule: each layer holds gates on disjoint lanes,
    in an order that preserves the circuit's order on every lane."""
    free = {}  # lane -> first layer where the lane is free
    out = []
    for op in circuit:
        ln = tuple(op.lanes)
        k = max(free.get(l, 0) for l in ln)
        if k == len(out):
            out.append([])
        out[k].append(op)
        for l in ln:
            free[l] = k + 1
    return out


# ---------------------------------------------------------------------------
# Yao circuits (ext/QrochetYaoExt.jl:15-48)
# ---------------------------------------------------------------------------
@dataclasses.dataclass
class Put:
    """Yao's `put(n, locs => block)`: a dense 2^k × 2^k block on `locs` (1-based), the
    first location the least significant digit of the matrix index (`mat(block)`)."""

    locs: tuple
    matrix: np.ndarray


@dataclasses.dataclass
class Swap:
    """Yao's `swap(n, a, b)`."""

    a: int
    b: int


def flatten(blocks) -> list:
    """Nested chains of blocks (Yao `ChainBlock`s) as one flat list, in order
    (`flatten_circuit` / `Yao.Optimise.eliminate_nested`, ext/QrochetYaoExt.jl:7-13)."""
    out = []
    for b in blocks:
        if isinstance(b, (list, tuple)):
            out.extend(flatten(b))
        else:
            out.append(b)
    return out


def from_yao(blocks, n: int) -> list:
    """The gates of a Yao circuit on `n` qubits as `Op`s on MPS sites, with the
    semantics of `Qrochet.Quantum(circuit::AbstractBlock)` (ext/QrochetYaoExt.jl:15-48):
    a `Swap` adds no tensor, it exchanges the two wires; every other block acts on
    the wires currently at its locations; at the end the input and output sites of
    lane l are the two ends of the wire that ended at l (Site(l; dual) => first,
    Site(l) => last, :44-47).  A wire is identified with the MPS site where it ends,
    so a gate applied at a time when lane l carries wire w acts on site
    end_slot(w), and no SWAP has to be applied to the state.  Only 2-level blocks
    (`@assert nlevel(circuit) == 2`, :16)."""
    flat = flatten(blocks)
    lane_wire = list(range(n))  # lane -> wire (0-based); wires are named by their start lane
    for b in flat:
        if isinstance(b, Swap):
            if not (1 <= b.a <= n and 1 <= b.b <= n):
                raise ArgumentError("swap locations outside the circuit")
            lane_wire[b.a - 1], lane_wire[b.b - 1] = lane_wire[b.b - 1], lane_wire[b.a - 1]
        elif not isinstance(b, Put):
            raise ArgumentError(f"unsupported Yao block {type(b).__name__} (put and swap blocks only)")
    end_slot = [0] * n
    for lane, w in enumerate(lane_wire):
        end_slot[w] = lane + 1
    lane_wire = list(range(n))
    ops = []
    for b in flat:
        if isinstance(b, Swap):
            lane_wire[b.a - 1], lane_wire[b.b - 1] = lane_wire[b.b - 1], lane_wire[b.a - 1]
            continue
        m = np.asarray(b.matrix, dtype=np.complex128)
        k = len(b.locs)
        if m.shape != (2**k, 2**k):
            raise ArgumentError("Only support 2-level qubits")
        if not all(1 <= l <= n for l in b.locs):
            raise ArgumentError("block locations outside the circuit")
        sites = tuple(end_slot[lane_wire[l - 1]] for l in b.locs)
        ops.append(Op("yao", sites, matrix=m))
    return ops


def run_(psi: Chain, circuit, threshold=None, maxdim=None, precision="fp64") -> Chain:
    """Apply ``circuit`` (list of Op) to ``psi`` in place: per layer, the one-site
    gates with `evolve!` (`qh_evolve1`) and all two-site gates in one batched call."""
    for layer in layers(circuit):
        two = []
        for op in layer:
            if len(op.lanes) == 1:
                evolve_(psi, gate(op.array(), op.lanes[0]))
            elif len(op.lanes) == 2:
                s, m = _oriented(op)
                two.append((psi, gate(m, s, s + 1)))
            else:
                raise ArgumentError(f"Invalid number of lanes {len(op.lanes)}, maximum is 2")
        if two:
            evolve_layer_(two, threshold=threshold, maxdim=maxdim, precision=precision)
    return psi
