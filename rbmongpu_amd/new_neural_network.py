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


def deep_connections(N, M, C):
    """Connection tables (j * delta + i) % n of new_neural_network.py:116-127 (1-D)."""
    def delta_func(n, m, c):
        if m > n:
            return 1
        elif n % m == 0:
            return n // m
        return c

    tables = []
    for n, m, c in zip([N] + M[:-1], M, C):
        d = delta_func(n, m, c)
        tables.append(np.array([[(j * d + i) % n for j in range(m)] for i in range(c)], dtype=np.uint32))
    return tables


def new_deep_neural_network(N, M, C, initial_value=(0.01 + 1j * math.pi / 4), alpha=0, beta=0,
                            free_quantum_axis=False, noise=1e-4, gpu=True, noise_modulation="auto", final_weights=10,
                            rng=None, **kwargs):
    """PsiDeep with the reference's initialisation (new_neural_network.py:53-153), 1-D layouts."""
    from . import PsiDeep
    for n, m, c in zip([N] + M[:-1], M, C):
        assert (m * c) % n == 0
        assert c <= n
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
    connections = deep_connections(N, M, C)
    if isinstance(final_weights, (float, int)):
        final_weights = final_weights * np.ones(M[-1])
    assert len(final_weights) == M[-1]
    return PsiDeep(alpha, beta, b, connections, W, np.asarray(final_weights, dtype=complex), 1.0, free_quantum_axis, gpu,
                   **kwargs)
