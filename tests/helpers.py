"""Shared test helpers: seeded models and independent (numpy) reference
routes used to pin the oracle -- dense Hamiltonians from Kronecker products of
Pauli matrices, closed-form log psi, basis enumeration."""
from __future__ import annotations

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402

# Single-spin matrices in the local basis (index 0 = bit 0 = spin -1,
# index 1 = bit 1 = spin +1).  In the usual (up, down) basis these are the
# textbook Pauli matrices; here the order is (down, up).
SX = np.array([[0, 1], [1, 0]], dtype=complex)
SY = np.array([[0, 1j], [-1j, 0]], dtype=complex)
SZ = np.array([[-1, 0], [0, 1]], dtype=complex)
ID2 = np.eye(2, dtype=complex)
PAULI = {0: ID2, 1: SX, 2: SY, 3: SZ}


def pauli_string_matrix(N, ops):
    """Dense 2^N matrix of prod_k sigma^{type_k}_{index_k} in the basis where
    basis index bit i is spin i (kron with spin N-1 leftmost)."""
    mats = [ID2] * N
    mats = list(mats)
    for t, i in ops:
        mats[i] = PAULI[t] @ mats[i] if mats[i] is not ID2 else PAULI[t]
    out = np.array([[1.0 + 0j]])
    for i in reversed(range(N)):
        out = np.kron(out, mats[i])
    return out


def dense_hamiltonian(N, coeffs, types, idx):
    """Dense matrix of an operator table whose strings have distinct indices."""
    H = np.zeros((2 ** N, 2 ** N), dtype=complex)
    for n in range(len(coeffs)):
        ops = []
        for t in range(types.shape[1]):
            if idx[n, t] == -1:
                break
            ops.append((int(types[n, t]), int(idx[n, t])))
        H += coeffs[n] * pauli_string_matrix(N, ops)
    return H


def basis_spins(N):
    """(2^N, N) array of spins in {-1, +1}, row = basis index."""
    idx = np.arange(2 ** N)
    return np.where((idx[:, None] >> np.arange(N)[None, :]) & 1, 1.0, -1.0)


def basis_words(N):
    return O.spins_words(np.arange(2 ** N))


def random_rbm(N, M, scale=0.1, seed=0, visible=False, fp32_exact=False):
    rng = np.random.default_rng(seed)
    b = scale * (rng.normal(size=M) + 1j * rng.normal(size=M))
    W = scale * (rng.normal(size=(N, M)) + 1j * rng.normal(size=(N, M)))
    a = scale * (rng.normal(size=N) + 1j * rng.normal(size=N)) if visible else None
    if fp32_exact:
        b, W = b.astype(np.complex64).astype(complex), W.astype(np.complex64).astype(complex)
        if a is not None:
            a = a.astype(np.complex64).astype(complex)
    return b, W, a


def random_operator(N, S, L, seed=0, hermitian_coeffs=False):
    """Random Pauli strings with distinct indices per string."""
    rng = np.random.default_rng(seed)
    coeffs = rng.normal(size=S) + (0 if hermitian_coeffs else 1j) * rng.normal(size=S)
    types = np.zeros((S, L + 1), dtype=np.int32)
    idx = -np.ones((S, L + 1), dtype=np.int32)
    for n in range(S):
        k = int(rng.integers(1, L + 1))
        pos = rng.choice(N, size=k, replace=False)
        for t in range(k):
            types[n, t] = int(rng.integers(1, 4))
            idx[n, t] = int(pos[t])
    return coeffs, types, idx


def deep_layers(N, M, C, seed=0, scale=0.3):
    """Random complex layers with the reference's connection tables."""
    rng = np.random.default_rng(seed)
    from rbmongpu_amd.new_neural_network import deep_connections
    conns = deep_connections(N, M, C)
    biases = [scale * (rng.normal(size=m) + 1j * rng.normal(size=m)) for m in M]
    weights = [scale * (rng.normal(size=(c, m)) + 1j * rng.normal(size=(c, m))) for c, m in zip(C, M)]
    final = 1.0 + 0.5 * (rng.normal(size=M[-1]) + 1j * rng.normal(size=M[-1]))
    return biases, conns, weights, final


def rel_err(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))
