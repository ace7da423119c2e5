"""Seeded synthetic workloads for the MPS update path (inputs only).

TEST / BENCH INFRASTRUCTURE.  These generate the inputs of the five workload
configurations (random MPS, Haar gates, Trotter gates, brickwork schedules)
with numpy's PCG64 so the GPU path and the CPU oracle see identical inputs.
The generators are plain numpy and carry no reference semantics of their
own; the product library receives their output as host arrays.
"""
from __future__ import annotations

import numpy as np

X = np.array([[0, 1], [1, 0]], dtype=np.complex128)
Y = np.array([[0, -1j], [1j, 0]], dtype=np.complex128)
Z = np.array([[1, 0], [0, -1]], dtype=np.complex128)
I2 = np.eye(2, dtype=np.complex128)


def haar_unitary(rng, n):
    """Haar-random U(n): QR of a complex Ginibre matrix with the phase fix of
    Mezzadri (2007), 'How to generate random matrices from the classical
    compact groups', Notices AMS 54(5)."""
    z = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(2.0)
    q, r = np.linalg.qr(z)
    ph = np.diag(r) / np.abs(np.diag(r))
    return q * ph[None, :]


def as_gate4(m, d=2):
    """(d²×d²) matrix → (out_l, out_r, in_l, in_r) array, site l = fastest
    index of the row/column (Julia column-major `reshape`)."""
    return np.asarray(m, dtype=np.complex128).reshape((d, d, d, d), order="F")


def haar_gate4(rng, d=2):
    return as_gate4(haar_unitary(rng, d * d), d)


def random_mps_arrays(rng, n, d, chis):
    """Internal-layout arrays (χl, d, χr) with i.i.d. (N(0,1)+iN(0,1))/√2 entries.
    ``chis`` has n-1 bond dimensions."""
    dims = [1] + list(chis) + [1]
    out = []
    for i in range(n):
        shp = (dims[i], d, dims[i + 1])
        a = (rng.standard_normal(shp) + 1j * rng.standard_normal(shp)) / np.sqrt(2.0)
        out.append(a)
    return out


def left_canonize_normalize(arrays):
    """Left-canonizing Householder QR sweep, then normalise the last site."""
    arrays = [a.copy() for a in arrays]
    n = len(arrays)
    for i in range(n - 1):
        cl, d, cr = arrays[i].shape
        q, r = np.linalg.qr(arrays[i].reshape(cl * d, cr))
        k = q.shape[1]
        arrays[i] = q.reshape(cl, d, k)
        arrays[i + 1] = np.tensordot(r, arrays[i + 1], axes=(1, 0))
    arrays[-1] = arrays[-1] / np.linalg.norm(arrays[-1])
    return arrays


def max_profile(n, d, chi):
    """χ_b = min(d^b, d^(n-b), χ) for b = 1..n-1."""
    return [min(d**b, d ** (n - b), chi) for b in range(1, n)]


def brickwork_bonds(n, layer):
    """Layer parity 0: bonds (1,2),(3,4),…; parity 1: (2,3),(4,5),…  (1-based left sites)."""
    start = 1 if layer % 2 == 0 else 2
    return list(range(start, n, 2))


def isometric_environment_state(rng, n=8, bond=4, chi=8):
    """Internal arrays and Λ's of a state whose environment is isometric on both
    sides of `bond` (sites < bond left-isometric including the Λ on bond-1, sites
    > bond+1 right-isometric including the Λ on bond+1), with non-trivial Λ
    vectors on bond-1, bond and bond+1."""
    arrays = random_mps_arrays(rng, n, 2, max_profile(n, 2, chi))
    for i in range(bond - 1):  # left-canonize sites 1 .. bond-1
        cl, d, cr = arrays[i].shape
        q, r = np.linalg.qr(arrays[i].reshape(cl * d, cr))
        arrays[i] = q.reshape(cl, d, q.shape[1])
        arrays[i + 1] = np.tensordot(r, arrays[i + 1], axes=(1, 0))
    for i in range(n - 1, bond, -1):  # right-canonize sites n .. bond+2
        cl, d, cr = arrays[i].shape
        q, r = np.linalg.qr(arrays[i].reshape(cl, d * cr).T)
        arrays[i] = q.T.reshape(q.shape[1], d, cr)
        arrays[i - 1] = np.tensordot(arrays[i - 1], r.T, axes=(2, 0))
    lams = [None] * (n - 1)
    lo = rng.random(arrays[bond - 2].shape[2]) + 0.2
    hi = rng.random(arrays[bond].shape[2]) + 0.2
    arrays[bond - 2] = arrays[bond - 2] / lo[None, None, :]   # A Λ stays isometric
    arrays[bond + 1] = arrays[bond + 1] / hi[:, None, None]
    lams[bond - 2], lams[bond] = lo, hi
    lams[bond - 1] = rng.random(arrays[bond - 1].shape[2]) + 0.2
    return arrays, lams


# ---------------------------------------------------------------------------
# workload 1: n=16, χ=32, 8 Haar brickwork layers, maxdim=32
# ---------------------------------------------------------------------------
def config1(seed_state=1234, seed_gates=5678, n=16, d=2, chi=32, layers=8):
    rng = np.random.default_rng(seed_state)
    arrays = left_canonize_normalize(random_mps_arrays(rng, n, d, max_profile(n, d, chi)))
    grng = np.random.default_rng(seed_gates)
    schedule = []
    for layer in range(layers):
        for b in brickwork_bonds(n, layer):
            schedule.append((b, haar_gate4(grng, d)))
    return arrays, schedule, dict(maxdim=chi, threshold=None)


# ---------------------------------------------------------------------------
# workload 2: n=64, χmax=256, Ising-type Trotter evolution (Strang splitting)
# ---------------------------------------------------------------------------
def ising_gate(J, h, tau):
    """exp(-i τ h_b), h_b = J Z⊗Z + (h/2)(X⊗I + I⊗X), from a Hermitian eigendecomposition."""
    hb = J * np.kron(Z, Z) + 0.5 * h * (np.kron(X, I2) + np.kron(I2, X))
    w, v = np.linalg.eigh(hb)
    u = (v * np.exp(-1j * tau * w)[None, :]) @ v.conj().T
    return as_gate4(u)


def config2(seed=2024, n=64, chi=256, steps=20, dt=0.05, J=None, h=None):
    """Returns (product-state arrays, schedule of (bond, gate) per layer, trunc)."""
    rng = np.random.default_rng(seed)
    arrays = []
    for _ in range(n):
        v = rng.standard_normal(2) + 1j * rng.standard_normal(2)
        v = v / np.linalg.norm(v)
        arrays.append(v.reshape(1, 2, 1))
    Js = rng.uniform(0.5, 1.5, n - 1) if J is None else np.full(n - 1, J)
    hs = rng.uniform(0.5, 1.5, n - 1) if h is None else np.full(n - 1, h)
    layers = []
    for _ in range(steps):
        for parity, tau in ((0, dt / 2), (1, dt), (0, dt / 2)):
            layers.append([(b, ising_gate(Js[b - 1], hs[b - 1], tau)) for b in brickwork_bonds(n, parity)])
    return arrays, layers, dict(maxdim=chi, threshold=1e-12), (Js, hs)


# ---------------------------------------------------------------------------
# workload 3: batch of uneven MPS, Haar brickwork, depth 10
# ---------------------------------------------------------------------------
def config3(seed=3, count=512, depth=10, nmin=8, nmax=64, chis=(16, 32, 64, 128, 256)):
    rng = np.random.default_rng(seed)
    items = []
    for _ in range(count):
        n = int(rng.integers(nmin, nmax + 1))
        chi = int(chis[rng.integers(0, len(chis))])
        arrays = left_canonize_normalize(random_mps_arrays(rng, n, 2, max_profile(n, 2, chi)))
        items.append(dict(n=n, chi=chi, arrays=arrays))
    grng = np.random.default_rng(seed + 1)
    for it in items:
        it["layers"] = [[(b, haar_gate4(grng)) for b in brickwork_bonds(it["n"], layer)] for layer in range(depth)]
    return items
