"""Model constructors with the reference's initialisation
(pyRBMonGPU/new_neural_network.py:21-153).  An optional numpy Generator makes
the draws reproducible; without one, numpy's global state is used as in the
reference."""
import math

import numpy as np


def real_noise(shape, rng=None):
    r = rng.random(shape) if rng is not None else np.random.random_sample(shape)
    return 2 * r - 1


def complex_noise(shape, rng=None):
    return real_noise(shape, rng) + 1j * real_noise(shape, rng)


def new_neural_network(N, M, initial_value=(0.01 + 1j * math.pi / 4), alpha=0, beta=0, free_quantum_axis=False,
                       noise=1e-6, gpu=True, rng=None, **kwargs):
    """RBM: W = noise * U(-1,1)(1+i) + initial_value on M // N diagonal blocks
    (new_neural_network.py:29-50)."""
    from . import Psi
    if isinstance(alpha, (int, float)):
        alpha = alpha * np.ones(N)
    if isinstance(beta, (int, float)):
        beta = beta * np.ones(N)
    b = noise * complex_noise(M, rng)
    W = noise * complex_noise((N, M), rng)
    for r in range(M // N):
        W[:, r * N:(r + 1) * N] += initial_value * np.diag(np.ones(N))
    return Psi(alpha, beta, b, W, 1, free_quantum_axis, gpu, **kwargs)


def _stride(n, m, c):
    """Input stride between neighbouring units of a layer of m units over n
    inputs with c connections each (new_neural_network.py:108-114)."""
    if m > n:
        return 1
    if n % m == 0:
        return n // m
    return c


def deep_connections(N, M, C):
    """Connection tables of the reference's deep constructor
    (new_neural_network.py:116-145).  1-D: unit j of a layer reads inputs
    (j * stride + i) % n for i < c.  2-D (N = (n0, n1), layers and
    connection counts as pairs): unit (j0, j1) reads the c0 x c1 window of
    inputs ((j0 * stride0 + i0) % n0, (j1 * stride1 + i1) % n1), units and
    window positions in row-major order, inputs numbered row * n1 + column.
    (The reference numbers them row * n0 + column, which is the same on the
    square lattices it is used with and addresses past the layer otherwise.)"""
    tables = []
    if isinstance(N, (list, tuple)):
        for n, m, c in zip([tuple(N)] + [tuple(x) for x in M[:-1]], [tuple(x) for x in M], [tuple(x) for x in C]):
            d0, d1 = _stride(n[0], m[0], c[0]), _stride(n[1], m[1], c[1])
            rows = []
            for i0 in range(c[0]):
                for i1 in range(c[1]):
                    rows.append([((j0 * d0 + i0) % n[0]) * n[1] + (j1 * d1 + i1) % n[1]
                                 for j0 in range(m[0]) for j1 in range(m[1])])
            tables.append(np.array(rows, dtype=np.uint32))
        return tables
    for n, m, c in zip([N] + M[:-1], M, C):
        d = _stride(n, m, c)
        tables.append(np.array([[(j * d + i) % n for j in range(m)] for i in range(c)], dtype=np.uint32))
    return tables


def new_deep_neural_network(N, M, C, initial_value=(0.01 + 1j * math.pi / 4), alpha=0, beta=0,
                            free_quantum_axis=False, noise=1e-4, gpu=True, noise_modulation="auto", final_weights=10,
                            rng=None, **kwargs):
    """PsiDeep with the reference's initialisation (new_neural_network.py:53-153):
    1-D layouts (N, M and C integers) or 2-D ones (N a pair, M and C lists of
    pairs; the layer sizes are the products)."""
    from . import PsiDeep
    two_d = isinstance(N, (list, tuple))
    for n, m, c in zip([N] + list(M[:-1]), M, C):
        if two_d:
            assert (m[0] * c[0]) % n[0] == 0 and (m[1] * c[1]) % n[1] == 0
            assert c[0] <= n[0] and c[1] <= n[1]
        else:
            assert (m * c) % n == 0
            assert c <= n
    connections = deep_connections(N, list(M), list(C))
    if two_d:
        N, M, C = N[0] * N[1], [m[0] * m[1] for m in M], [c[0] * c[1] for c in C]
    if isinstance(alpha, (float, int)):
        alpha = alpha * np.ones(N)
    if isinstance(beta, (float, int)):
        beta = beta * np.ones(N)
    b = [noise * complex_noise(m, rng) for m in M]
    w = noise * complex_noise((C[0], M[0]), rng)
    w[C[0] // 2, :] += initial_value
    W = [w]
    if noise_modulation == "auto":
        noise_modulation = [math.sqrt(6 / (c + next_c)) for c, m, next_c in zip(C[1:], M[1:], C[2:] + [1])]
    for c, m, next_c, nm in zip(C[1:], M[1:], C[2:] + [1], noise_modulation):
        W.append(nm * math.sqrt(6 / (c + next_c)) * real_noise((c, m), rng) + noise * complex_noise((c, m), rng))
    if isinstance(final_weights, (float, int)):
        final_weights = final_weights * np.ones(M[-1])
    assert len(final_weights) == M[-1]
    return PsiDeep(alpha, beta, b, connections, W, np.asarray(final_weights, dtype=complex), 1.0, free_quantum_axis, gpu,
                   **kwargs)
