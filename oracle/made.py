"""FP64 numpy restatement of the masked autoregressive proposal (MADE) and of
the independent Metropolis-Hastings chain that uses it.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker.  The product (rbmongpu_amd) never
imports this module.

The reference has no learned proposal; this component is defined here (after
Germain et al., "MADE: Masked Autoencoder for Distribution Estimation", ICML
2015) and every GPU kernel of it is checked against these functions.

Model (N visible spins, H hidden units):
  x_i = 1 if spin i is up (bit i set) else 0            (include/Spins.h:125-129)
  a = W1 x + b1,  h = relu(a),  l = W2 h + b2
  q(x) = prod_j sigma(l_j)^x_j (1 - sigma(l_j))^(1 - x_j)
  log q(x) = sum_j x_j l_j - softplus(l_j)
Masks from degrees: input i has degree i + 1, hidden unit k has
m_k = 1 + (k mod (N - 1)); W1[k, i] is allowed iff i < m_k and W2[j, k] iff
m_k <= j, so l_j depends on x_0 .. x_{j-1} only.
Flat parameter layout (FP32 on the device): [W1 (H x N), b1 (H), W2 (N x H), b2 (N)].
"""
from __future__ import annotations

import numpy as np

MASK32 = np.uint64(0xFFFFFFFF)


def philox_np(c0, c1, c2, c3, k0, k1):
    """Vectorised Philox4x32-10 (Salmon et al., SC'11) on uint64 arrays holding
    32-bit words; the same function as oracle.philox, batched."""
    c = [np.asarray(v, dtype=np.uint64) & MASK32 for v in (c0, c1, c2, c3)]
    c = np.broadcast_arrays(*c)
    c = [v.copy() for v in c]
    k0 = np.uint64(k0) & MASK32
    k1 = np.uint64(k1) & MASK32
    for r in range(10):
        if r:
            k0 = (k0 + np.uint64(0x9E3779B9)) & MASK32
            k1 = (k1 + np.uint64(0xBB67AE85)) & MASK32
        p0 = np.uint64(0xD2511F53) * c[0]
        p1 = np.uint64(0xCD9E8D57) * c[2]
        n0 = (p1 >> np.uint64(32)) ^ c[1] ^ k0
        n2 = (p0 >> np.uint64(32)) ^ c[3] ^ k1
        c = [n0 & MASK32, p1 & MASK32, n2 & MASK32, p0 & MASK32]
    return c


def degrees(N, H):
    if N <= 1:
        return np.zeros(H, dtype=np.int64)
    return 1 + (np.arange(H) % (N - 1))


def masks(N, H):
    m = degrees(N, H)
    M1 = (np.arange(N)[None, :] < m[:, None]).astype(np.float64)      # H x N
    M2 = (m[None, :] <= np.arange(N)[:, None]).astype(np.float64)     # N x H
    return M1, M2


def num_params(N, H):
    return 2 * N * H + H + N


def unpack(p, N, H):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    W1 = p[o:o + H * N].reshape(H, N); o += H * N
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + N * H].reshape(N, H); o += N * H
    b2 = p[o:o + N]
    return W1, b1, W2, b2


def pack(W1, b1, W2, b2):
    return np.concatenate([np.ravel(W1), b1, np.ravel(W2), b2])


def init_params(N, H, seed):
    """Seeded initialisation: parameter k (flat index) draws Philox(key = seed;
    ctr = (k, 0, 0, 3)) word x, u = (x + 0.5) 2^-32; weights are (2u - 1)/sqrt(fan_in)
    times the mask (fan_in N for W1, H for W2); biases are zero."""
    P = num_params(N, H)
    k = np.arange(P, dtype=np.uint64)
    x = philox_np(k, 0, 0, 3, seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF)[0]
    u = (x.astype(np.float64) + 0.5) * 2.0 ** -32
    W1, b1, W2, b2 = unpack(2.0 * u - 1.0, N, H)
    M1, M2 = masks(N, H)
    W1 = W1 * M1 / np.sqrt(N)
    W2 = W2 * M2 / np.sqrt(H)
    p = pack(W1, np.zeros(H), W2, np.zeros(N))
    return p.astype(np.float32).astype(np.float64)


def bits(spins, N):
    """(n, 2) uint64 words -> (n, N) float 0/1."""
    w = np.asarray(spins, dtype=np.uint64).reshape(-1, 2)
    i = np.arange(N)
    word = np.where(i < 64, w[:, [0]], w[:, [1]])
    shift = (i % 64).astype(np.uint64)
    return ((word >> shift) & np.uint64(1)).astype(np.float64)


def words(X):
    X = np.asarray(X).astype(np.uint64)
    n, N = X.shape
    out = np.zeros((n, 2), dtype=np.uint64)
    for i in range(N):
        out[:, i // 64] |= X[:, i] << np.uint64(i % 64)
    return out


def softplus(l):
    return np.maximum(l, 0.0) + np.log1p(np.exp(-np.abs(l)))


def forward(p, X, N, H):
    W1, b1, W2, b2 = unpack(p, N, H)
    a = X @ W1.T + b1
    h = np.maximum(a, 0.0)
    l = h @ W2.T + b2
    return a, h, l


def log_q(p, X, N, H):
    _, _, l = forward(p, X, N, H)
    return np.sum(X * l - softplus(l), axis=1)


def loss(p, X, N, H):
    return -float(np.mean(log_q(p, X, N, H)))


def grad(p, X, N, H):
    """Gradient of loss = -mean_s log q(x_s) by back-propagation."""
    W1, b1, W2, b2 = unpack(p, N, H)
    B = X.shape[0]
    a, h, l = forward(p, X, N, H)
    dl = (1.0 / (1.0 + np.exp(-l)) - X) / B
    gW2 = dl.T @ h
    gb2 = dl.sum(axis=0)
    dh = (dl @ W2) * (a > 0)
    gW1 = dh.T @ X
    gb1 = dh.sum(axis=0)
    M1, M2 = masks(N, H)
    return pack(gW1 * M1, gb1, gW2 * M2, gb2)


def adam(p, g, m, v, t, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (Kingma & Ba 2015, Algorithm 1); t is the 1-based step."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v


def sample(p, U, N, H):
    """Autoregressive sampling with given uniforms U (n x N): x_j = 1 iff
    U_j < sigma(l_j).  Returns X and log q(X)."""
    W1, b1, W2, b2 = unpack(p, N, H)
    n = U.shape[0]
    X = np.zeros((n, N))
    a = np.tile(b1, (n, 1))
    lq = np.zeros(n)
    for j in range(N):
        l = np.maximum(a, 0.0) @ W2[j] + b2[j]
        x = (U[:, j] < 1.0 / (1.0 + np.exp(-l))).astype(np.float64)
        X[:, j] = x
        lq += x * l - softplus(l)
        a += np.outer(x, W1[:, j])
    return X, lq


def bf16_round(x):
    """Round FP32 values to bfloat16 (nearest, ties to even) and back."""
    f = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    f = (f + np.uint64(0x7FFF) + ((f >> np.uint64(16)) & np.uint64(1))) & np.uint64(0xFFFF0000)
    return f.astype(np.uint32).view(np.float32).astype(np.float64)


def inference_params(p, N, H):
    """The parameters as the sampling kernels use them: weights in bfloat16,
    biases in FP32."""
    W1, b1, W2, b2 = unpack(p, N, H)
    return pack(bf16_round(W1), np.float32(b1).astype(np.float64), bf16_round(W2), np.float32(b2).astype(np.float64))


def all_configs(N):
    return ((np.arange(2 ** N)[:, None] >> np.arange(N)[None, :]) & 1).astype(np.float64)


def imh_chain(logpsi_re, lq, U, therm, steps):
    """Independent Metropolis-Hastings over a precomputed proposal stream of
    one chain: proposal t has Re log psi logpsi_re[t], log q lq[t] and uniform
    U[t].  Proposal 0 is the start; proposal t >= 1 is accepted iff
    U[t] <= exp(2 (Re log psi' - Re log psi) + log q - log q').  Returns the
    proposal index held after each of the `steps` measured proposals (taken
    after `therm` unmeasured ones) and the number of acceptances."""
    cur = 0
    out = []
    acc = 0
    for t in range(1, 1 + therm + steps):
        r = 2.0 * (logpsi_re[t] - logpsi_re[cur]) + lq[cur] - lq[t]
        if U[t] <= np.exp(r):
            cur = t
            acc += 1
        if t > therm:
            out.append(cur)
    return np.array(out, dtype=np.int64), acc


def imh_transition_matrix(prob, q):
    """Exact transition matrix of independent MH with target prob and proposal q."""
    n = len(prob)
    T = np.zeros((n, n))
    for s in range(n):
        for t in range(n):
            if s != t and q[t] > 0:
                T[s, t] = q[t] * min(1.0, prob[t] * q[s] / (prob[s] * q[t]))
        T[s, s] = 1.0 - T[s].sum()
    return T


def log_q_inference(p, X, N, H):
    """log q with the device sampler's arithmetic: bfloat16 weights, hidden
    pre-activations accumulated in FP32 in position order (a += W1[:, j] for
    x_j = 1), relu(a) rounded to bfloat16 before the logit dot products, which
    (and everything after them) are evaluated here in FP64."""
    W1, b1, W2, b2 = unpack(inference_params(p, N, H), N, H)
    X = np.asarray(X, dtype=np.float64)
    n = X.shape[0]
    a = np.tile(b1.astype(np.float32), (n, 1))
    W1f = W1.astype(np.float32)
    lq = np.zeros(n)
    for j in range(N):
        h = bf16_round(np.maximum(a, np.float32(0)))
        l = h @ W2[j] + b2[j]
        lq += X[:, j] * l - softplus(l)
        on = X[:, j] > 0
        a[on] = (a[on] + W1f[:, j][None, :]).astype(np.float32)
    return lq
