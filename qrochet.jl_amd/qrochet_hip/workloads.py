"""Seeded synthetic inputs of the benchmark workloads (host side).

These build the host arrays that `bench.py` uploads: random MPS (i.i.d. complex
Gaussian entries, left-canonized), Haar-random two-site gates (Mezzadri's QR
construction), Ising Trotter gates and brickwork schedules, all from numpy's
PCG64.  The parity tests use generators with the same seeds and algorithms on
their side and check (tests/test_workloads.py) that both produce identical
arrays, so the benchmarked inputs are the ones the tests validate.  Nothing here
computes a result of the update path.
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


# ---------------------------------------------------------------------------
# workload 4: random Hermitian nearest-neighbour MPO applied to a random MPS
# ---------------------------------------------------------------------------
def random_hermitian_mpo(rng, n, d, D):
    """Reference-shaped MPO arrays (o, i, l, r): complex Gaussian entries × 1/sqrt(D·d),
    each d×d block W[w, :, :, v] replaced by its Hermitian part (so H is Hermitian)."""
    dims = [1] + [D] * (n - 1) + [1]
    out = []
    for k in range(n):
        shp = (dims[k], d, d, dims[k + 1])
        w = (rng.standard_normal(shp) + 1j * rng.standard_normal(shp)) / np.sqrt(2.0 * D * d)
        w = 0.5 * (w + np.conj(np.transpose(w, (0, 2, 1, 3))))
        a = np.transpose(w, (1, 2, 0, 3))  # (o, i, l, r)
        if k == 0:
            a = a[:, :, 0, :]
        if k == n - 1:
            a = a[..., 0]
        out.append(np.ascontiguousarray(a))
    return out


def config4(seed=4, n=32, d=2, chi=512, D=8):
    """Random normalised MPS at the χ profile and a random Hermitian MPO of bond D."""
    rng = np.random.default_rng(seed)
    arrays = left_canonize_normalize(random_mps_arrays(rng, n, d, max_profile(n, d, chi)))
    mpo = random_hermitian_mpo(rng, n, d, D)
    return arrays, mpo
