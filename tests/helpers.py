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


# ------------------------------------------------------------------ branch disagreements
# my_logcosh is sign-branched on Re theta > 0 (psi_functions.hpp:49-57) and jumps
# across Re theta = 0; a complex64 angle on the other side of 0 than the FP64
# angle is evaluated on the other branch.  These helpers replay the kernels'
# FP32 angle arithmetic to find such configurations exactly.
def spins_pm(words, N):
    s = np.where((words[:, :1] >> np.arange(min(N, 64), dtype=np.uint64)) & np.uint64(1), 1.0, -1.0)
    if N > 64:
        s = np.concatenate([s, np.where((words[:, 1:] >> np.arange(N - 64, dtype=np.uint64)) & np.uint64(1), 1.0, -1.0)],
                           axis=1)
    return s


def flip_groups(tab):
    """Distinct flip sets of the table's strings (X and Y positions, sorted
    ascending like the canonical operator form) with the sum of |coefficient|
    of the strings that share each set."""
    c, t, idx = tab
    groups = {}
    for n in range(len(c)):
        fl = tuple(sorted({int(idx[n, k]) for k in range(t.shape[1]) if idx[n, k] != -1 and t[n, k] in (1, 2)}))
        if fl:
            groups[fl] = groups.get(fl, 0.0) + abs(c[n])
    return groups


def _jump(y):
    """|my_logcosh(+0 + i y) - my_logcosh(-0 + i y)|, the jump across Re = 0."""
    return np.abs(O.logcosh(1e-300 + 1j * y) - O.logcosh(-1e-300 + 1j * y))


def branch_analysis(b, W, s, tab):
    """Replays the kernel's FP32 angles.  Returns per configuration:
    own   -- an angle of s itself is on the other side of Re = 0 than in FP64
             (log psi and every ratio move by the jump);
    any   -- own, or an angle of a flipped neighbour s' disagrees (that ratio
             moves);
    bound -- for flagged configurations, an upper bound of the E_loc change
             the disagreements can cause: sum over flip sets of
             |coefficients| |psi(s')/psi(s)| x (the largest jump involved)."""
    W32 = W.astype(np.complex64)
    th32 = np.tile(b.astype(np.complex64), (len(s), 1))
    for i in range(W.shape[0]):
        th32 = th32 + s[:, i:i + 1].astype(np.float32) * W32[i][None, :]
    th64 = b[None, :] + s @ W

    def worst_jump(v32, v64):
        dis = (v32.real > 0) != (v64.real > 0)
        out = np.zeros(len(s))
        if dis.any():
            rows, cols = np.nonzero(dis)
            np.maximum.at(out, rows, _jump(np.abs(v64.imag[rows, cols])))
        return out

    j_own = worst_jump(th32, th64)
    groups = flip_groups(tab)
    j_term = np.zeros((len(s), len(groups)))

    def neighbour(th, fl, rows, dtype):
        v = th[rows].copy()
        for p in fl:
            v = v + (-2.0 * s[rows, p:p + 1]).astype(dtype) * W.astype(v.dtype)[p][None, :]
        return v

    every = np.arange(len(s))
    for g, fl in enumerate(groups):
        j_term[:, g] = worst_jump(neighbour(th32, fl, every, np.float32), neighbour(th64, fl, every, np.float64))
    own = j_own > 0
    anyd = own | np.any(j_term > 0, axis=1)
    bound = np.zeros(len(s))
    rows = np.nonzero(anyd)[0]
    if len(rows):
        lc0 = O.logcosh(th64[rows]).sum(axis=1)
        for g, (fl, coef) in enumerate(groups.items()):
            ratio = np.abs(np.exp(O.logcosh(neighbour(th64, fl, rows, np.float64)).sum(axis=1) - lc0))
            bound[rows] += coef * ratio * np.maximum(j_own[rows], j_term[rows, g])
    return own, anyd, bound


def check_local_energies(label, el, El, lp, lpr, b, W, s, tab):
    own, anyd, bound = branch_analysis(b, W, s, tab)
    good = ~anyd
    scale = np.abs(El).max()
    err_good = float(np.abs(el[good] - El[good]).max() / scale)
    err_lp = float(np.abs(lp[~own] - lpr[~own]).max() / np.abs(lpr).max())
    print(f"[{label}] n={len(El)}  branch-disagreeing configurations: {int(anyd.sum())} ({int(own.sum())} at s "
          f"itself)  E_loc rel err (others) {err_good:.2e}  log psi rel err {err_lp:.2e}", flush=True)
    assert err_good < 1e-5
    assert err_lp < 1e-5
    if anyd.any():
        dev = np.abs(el[anyd] - El[anyd])
        print(f"  disagreeing: |E_loc error| {dev.max():.2e} against the jump bound {bound[anyd].min():.2e}.."
              f"{bound[anyd].max():.2e}", flush=True)
        assert np.all(dev <= 1.05 * bound[anyd] + 1e-5 * scale)
        assert anyd.mean() < 0.02
    return good
