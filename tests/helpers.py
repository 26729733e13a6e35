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


def bf16(x):
    """Round to the nearest bfloat16 through float32 (ties to even), as float64."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def bf16c(z):
    z = np.asarray(z)
    return bf16(z.real) + 1j * bf16(z.imag)


def dense_layers(N, conns, weights):
    """Dense (inputs x units) complex matrices of sparse PsiDeep layers."""
    out, prev = [], N
    for c, w in zip(conns, weights):
        C, size = w.shape
        D = np.zeros((prev, size), complex)
        for k in range(C):
            np.add.at(D, (c[k], np.arange(size)), w[k])
        out.append(D)
        prev = size
    return out


def deep_bf16_log_psi(N, biases, conns, weights, final, words):
    """log psi of the BF16 MFMA forward pass in float64 arithmetic: weights and
    the inputs of every layer after the first rounded to bfloat16, biases and
    final weights rounded to float32, my_logcosh from the FP64 oracle."""
    w = np.asarray(words, np.uint64)
    x = np.where((w[:, :1] >> np.arange(N, dtype=np.uint64)) & np.uint64(1), 1.0, -1.0).astype(complex)
    dense = dense_layers(N, conns, weights)
    for l, (D, b) in enumerate(zip(dense, biases)):
        a = x @ bf16c(D) + np.asarray(b, np.complex64).astype(complex)
        y = O.logcosh(a)
        x = bf16c(y) if l + 1 < len(dense) else y
    return x @ np.asarray(final, np.complex64).astype(complex)
