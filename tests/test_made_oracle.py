"""Pins of the MADE / independent-MH restatement (oracle/made.py) by routes
that do not retype its formulas: the C oracle's Philox, exact normalisation of
q over the whole basis, the autoregressive dependency structure, central
finite differences of the loss, Adam's constant-gradient fixed step, known
bfloat16 roundings, and the stationary distribution of the exact transition
matrix."""
import numpy as np
import pytest

from helpers import O
from oracle import made as MD


def rand_params(N, H, seed, scale=0.7):
    rng = np.random.default_rng(seed)
    M1, M2 = MD.masks(N, H)
    return MD.pack(scale * rng.normal(size=(H, N)) * M1, scale * rng.normal(size=H),
                   scale * rng.normal(size=(N, H)) * M2, scale * rng.normal(size=N))


def test_vectorised_philox_matches_c_oracle():
    rng = np.random.default_rng(1)
    for _ in range(20):
        ctr = rng.integers(0, 2 ** 32, size=4, dtype=np.uint64)
        key = rng.integers(0, 2 ** 32, size=2, dtype=np.uint64)
        got = [int(v) for v in MD.philox_np(*ctr, *key)]
        assert got == O.philox(ctr.tolist(), key.tolist()).tolist()


@pytest.mark.parametrize("N,H", [(1, 4), (2, 3), (6, 10), (9, 24)])
def test_q_is_normalised(N, H):
    p = rand_params(N, H, seed=N)
    X = MD.all_configs(N)
    assert abs(np.exp(MD.log_q(p, X, N, H)).sum() - 1.0) < 1e-12


def test_autoregressive_dependencies():
    N, H = 7, 30
    p = rand_params(N, H, seed=4)
    rng = np.random.default_rng(0)
    X = (rng.random((50, N)) < 0.5).astype(float)
    _, _, l0 = MD.forward(p, X, N, H)
    for j in range(N):
        Y = X.copy()
        Y[:, j] = 1 - Y[:, j]
        _, _, l1 = MD.forward(p, Y, N, H)
        assert np.array_equal(l0[:, :j + 1], l1[:, :j + 1])       # outputs 0..j ignore x_j
        if j < N - 1:
            assert np.abs(l0[:, j + 1:] - l1[:, j + 1:]).max() > 1e-6  # and later outputs see it


def test_gradient_matches_finite_differences():
    N, H, B = 6, 11, 9
    p = rand_params(N, H, seed=7, scale=0.5)
    M1, M2 = MD.masks(N, H)
    allowed = MD.pack(M1, np.ones(H), M2, np.ones(N)) > 0
    rng = np.random.default_rng(2)
    X = (rng.random((B, N)) < 0.5).astype(float)
    g = MD.grad(p, X, N, H)
    assert np.all(g[~allowed] == 0)
    eps = 1e-6
    for k in np.flatnonzero(allowed):
        d = np.zeros_like(p)
        d[k] = eps
        fd = (MD.loss(p + d, X, N, H) - MD.loss(p - d, X, N, H)) / (2 * eps)
        assert abs(fd - g[k]) < 1e-8 + 1e-6 * abs(fd)


def test_adam_constant_gradient_moves_by_lr():
    """With a constant gradient the bias-corrected moments are g and g^2 at
    every step, so each step moves by lr * g / (|g| + eps) exactly."""
    g = np.array([0.3, -2.0, 1e-3])
    p, m, v = np.zeros(3), np.zeros(3), np.zeros(3)
    for t in range(1, 6):
        p_new, m, v = MD.adam(p, g, m, v, t, lr=0.01)
        assert np.allclose(p_new - p, -0.01 * g / (np.abs(g) + 1e-8), rtol=1e-12, atol=0)
        p = p_new


def test_bf16_rounding_known_values():
    x = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, 1 + 2 ** -7, -3.140625, 1 + 2 ** -8 + 2 ** -20], dtype=np.float32)
    want = [1.0, 1.0, 1 + 2 ** -6, 1 + 2 ** -7, -3.140625, 1 + 2 ** -7]
    assert MD.bf16_round(x).tolist() == want


def test_sampler_draws_q():
    N, H = 5, 12
    p = rand_params(N, H, seed=9)
    rng = np.random.default_rng(5)
    n = 200_000
    X, lq = MD.sample(p, rng.random((n, N)), N, H)
    assert np.allclose(lq, MD.log_q(p, X, N, H), atol=1e-12)
    idx = (X * (1 << np.arange(N))).sum(axis=1).astype(int)
    counts = np.bincount(idx, minlength=2 ** N)
    q = np.exp(MD.log_q(p, MD.all_configs(N), N, H))
    z = (counts - n * q) / np.sqrt(n * q * (1 - q))
    assert np.abs(z).max() < 4.5


def test_imh_transition_matrix_is_stationary_at_target():
    rng = np.random.default_rng(11)
    n = 16
    prob = rng.random(n) ** 3
    prob /= prob.sum()
    q = rng.random(n) + 0.05
    q /= q.sum()
    T = MD.imh_transition_matrix(prob, q)
    assert np.allclose(T.sum(axis=1), 1.0, atol=1e-14)
    assert np.allclose(prob @ T, prob, atol=1e-14)
    # detailed balance
    F = prob[:, None] * T
    assert np.allclose(F, F.T, atol=1e-15)


def test_imh_chain_follows_the_transition_matrix():
    """A long chain driven by imh_chain visits states with frequencies |psi|^2."""
    N, H = 4, 8
    p = rand_params(N, H, seed=3, scale=0.4)
    rng = np.random.default_rng(8)
    lpr = rng.normal(size=2 ** N) * 0.6          # Re log psi per state
    prob = np.exp(2 * lpr)
    prob /= prob.sum()
    steps = 400_000
    X, lq = MD.sample(p, rng.random((steps + 1, N)), N, H)
    idx = (X * (1 << np.arange(N))).sum(axis=1).astype(int)
    held, acc = MD.imh_chain(lpr[idx], lq, rng.random(steps + 1), 0, steps)
    counts = np.bincount(idx[held], minlength=2 ** N)
    # autocorrelated chain: allow a wide band on the relative deviation
    assert np.abs(counts / steps - prob).max() < 0.01
    assert 0 < acc < steps


@pytest.mark.parametrize("N,H", [(5, 8), (9, 40)])
def test_inference_arithmetic_is_a_normalised_autoregressive_model(N, H):
    """bfloat16 weights and activations change q but keep it a distribution
    whose conditionals are those the sampler draws from."""
    p = rand_params(N, H, seed=N * H)
    X = MD.all_configs(N)
    lq = MD.log_q_inference(p, X, N, H)
    assert abs(np.exp(lq).sum() - 1.0) < 1e-12
    assert np.abs(lq - MD.log_q(p, X, N, H)).max() < 0.1
