"""Pins of the FP64 CPU oracle (oracle/rbm_oracle.c) against routes that do
not restate its formulas: Random123 known-answer vectors, calculus identities
of the log-cosh approximant, closed-form log psi with numpy angles (the
reference's own test_Psi / test_PsiDeep checks), dense Hamiltonians from
Kronecker products of Pauli matrices, central finite differences and the
stationary distribution of the Metropolis chain."""
import numpy as np
import pytest

from helpers import (O, basis_spins, basis_words, dense_hamiltonian, deep_layers, random_operator, random_rbm,
                     rel_err)
from rbmongpu_amd import hamiltonians as HAM


# ---------------------------------------------------------------- Philox
@pytest.mark.parametrize("ctr,key,expected", [
    ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
])
def test_philox_known_answers(ctr, key, expected):
    """Random123 kat_vectors for philox4x32-10."""
    assert O.philox(ctr, key).tolist() == expected


# ---------------------------------------------------------------- activation
def test_logcosh_approximates_log_cosh_on_real_axis():
    x = np.linspace(-1, 1, 2001)
    err = np.abs(O.logcosh(x).real - np.log(np.cosh(x)))
    assert err.max() < 3e-2            # a rational fit of log cosh near 0, not log cosh itself
    assert np.abs(O.logcosh(x).imag).max() == 0.0
    f0, f1 = O.logcosh(0.0)[0].real, O.logcosh(0.01)[0].real
    assert abs((f1 - f0) / 0.01 ** 2 - 0.5) < 0.05   # log cosh x = x^2 / 2 + O(x^4)
    assert abs(O.logcosh(0.0)[0].real - (-1.306188e-7)) < 1e-12  # my_logcosh(0) of the reference approximant


def test_logcosh_is_even():
    rng = np.random.default_rng(3)
    z = rng.normal(size=200) + 1j * rng.normal(size=200)
    z = z[np.abs(z.real) > 1e-3]
    assert np.allclose(O.logcosh(z), O.logcosh(-z), rtol=0, atol=1e-14)


def test_tanh_is_derivative_of_logcosh():
    """my_tanh = d my_logcosh / dz (holomorphic on each half plane): a dropped
    or altered term in either function breaks this."""
    rng = np.random.default_rng(4)
    z = rng.normal(size=300) + 1j * rng.normal(size=300)
    z = z[np.abs(z.real) > 1e-2]
    h = 1e-6
    for d in (h, 1j * h):
        fd = (O.logcosh(z + d) - O.logcosh(z - d)) / (2 * d)
        assert np.allclose(fd, O.tanh(z), rtol=1e-7, atol=1e-8)


def test_logcosh_branch_jump_is_in_imaginary_part():
    y = 0.05
    hi, lo = O.logcosh(1e-14 + 1j * y)[0], O.logcosh(-1e-14 + 1j * y)[0]
    assert abs(hi.real - lo.real) < 1e-12
    assert 1e-6 < abs(hi.imag - lo.imag) < 1e-4


# ---------------------------------------------------------------- RBM
@pytest.mark.parametrize("N,M,visible", [(3, 9, False), (6, 5, True), (10, 20, False)])
def test_rbm_log_psi_closed_form(N, M, visible):
    """psi(s) = exp(sum_j activation(W[:, j] . s + b_j)) (+ a . s), test_Psi.py:21-30."""
    b, W, a = random_rbm(N, M, seed=N, visible=visible)
    psi = O.RBM(b, W, a)
    S = basis_spins(N)
    theta = S @ W + b
    ref = O.logcosh(theta.ravel()).reshape(theta.shape).sum(axis=1)
    if visible:
        ref = ref + S @ a
    got = O.log_psi(psi, basis_words(N))
    assert rel_err(got, ref) < 1e-13
    assert rel_err(O.psi_vector(psi), np.exp(ref)) < 1e-13


def _fd_logpsi(model_from_params, params, words, k, h=1e-6):
    p1, p2 = params.copy(), params.copy()
    p1[k] += h
    p2[k] -= h
    return (O.log_psi(model_from_params(p1), words) - O.log_psi(model_from_params(p2), words)) / (2 * h)


def test_rbm_o_k_finite_differences():
    N, M = 4, 6
    b, W, a = random_rbm(N, M, seed=7, visible=True)
    params = np.concatenate([b, W.ravel(), a])

    def model(p):
        return O.RBM(p[:M], p[M:M + N * M].reshape(N, M), p[M + N * M:])

    words = O.spins_words([5, 9])
    for w in words:
        ok = O.o_k(model(params), w)
        assert np.all(ok[:2 * N] == 0)
        for k in range(len(params)):
            fd = _fd_logpsi(model, params, w[None, :], k)[0]
            assert abs(fd - ok[2 * N + k]) < 1e-7


@pytest.mark.parametrize("ham", ["tfim", "heis", "random"])
def test_local_energy_dense(ham):
    N, M = 6, 7
    b, W, _ = random_rbm(N, M, seed=11)
    psi = O.RBM(b, W)
    if ham == "tfim":
        tab = HAM.tfim_1d(N, 1.0, 0.7)
    elif ham == "heis":
        tab = HAM.heisenberg([(i, (i + 1) % N) for i in range(N)])
    else:
        tab = random_operator(N, 12, 3, seed=5)
    op = O.Operator(*tab)
    H = dense_hamiltonian(N, *tab)
    v = O.psi_vector(psi)
    ref = (H @ v) / v
    got = O.local_energy(psi, op, basis_words(N))
    assert rel_err(got, ref) < 1e-12


def test_exact_expectation_fluctuation_norm_dense():
    N, M = 6, 4
    b, W, _ = random_rbm(N, M, seed=12)
    tab = HAM.tfim_1d(N, 1.0, 1.3)
    H = dense_hamiltonian(N, *tab)
    psi = O.RBM(b, W)
    v = O.psi_vector(psi)
    ex = O.Exact(N)
    assert abs(O.psi_norm(psi, ex) - np.linalg.norm(v)) < 1e-12 * np.linalg.norm(v)
    # unnormalised: <E> = sum_s |psi|^2 E_loc = psi^H H psi
    assert abs(O.expectation(psi, O.Operator(*tab), ex) - np.vdot(v, H @ v)) < 1e-10 * abs(np.vdot(v, H @ v))
    psi.prefactor = 1.0 / np.linalg.norm(v)
    vn = v / np.linalg.norm(v)
    f, E = O.fluctuation(psi, O.Operator(*tab), ex)
    E_ref = np.vdot(vn, H @ vn).real
    assert abs(E - E_ref) < 1e-12
    assert abs(f - np.sqrt(np.vdot(H @ vn, H @ vn).real - E_ref ** 2)) < 1e-9


def test_exact_sector_enumeration():
    N = 6
    ex = O.Exact(N, sector=0)
    assert ex.num_configs == 20
    b, W, _ = random_rbm(N, 3, seed=2)
    psi = O.RBM(b, W)
    v = O.psi_vector(psi)
    sel = [i for i in range(2 ** N) if 2 * bin(i).count("1") - N == 0]
    assert abs(O.psi_norm(psi, ex) - np.linalg.norm(v[sel])) < 1e-12


def test_gradient_is_derivative_of_energy():
    """For Hermitian H and normalised psi the reference estimator equals
    d<H>/dRe p + i d<H>/dIm p of the dense normalised energy."""
    N, M = 4, 3
    b, W, _ = random_rbm(N, M, seed=21, scale=0.3)
    tab = HAM.heisenberg([(i, (i + 1) % N) for i in range(N)], 1.0)
    tab = (np.concatenate([tab[0], [0.4]]), np.concatenate([tab[1], [[1, 0, 0]]]),
           np.concatenate([tab[2], [[2, -1, -1]]]))
    H = dense_hamiltonian(N, *tab)
    params = np.concatenate([b, W.ravel()])

    def model(p):
        return O.RBM(p[:M], p[M:].reshape(N, M))

    def energy(p):
        v = O.psi_vector(model(p))
        return (np.vdot(v, H @ v) / np.vdot(v, v)).real

    psi = model(params)
    psi.prefactor = 1.0 / np.linalg.norm(O.psi_vector(psi))
    g0, E0 = O.gradient(psi, O.Operator(*tab), O.Exact(N), mode=0)
    g1, _ = O.gradient(psi, O.Operator(*tab), O.Exact(N), mode=1)
    assert abs(E0 - energy(params)) < 1e-12
    assert np.allclose(g0, g1, rtol=0, atol=1e-12)
    h = 1e-6
    for k in range(len(params)):
        d = np.zeros_like(params)
        d[k] = h
        fd = (energy(params + d) - energy(params - d)) / (2 * h)
        fd += 1j * (energy(params + 1j * d) - energy(params - 1j * d)) / (2 * h)
        assert abs(g0[2 * N + k] - fd) < 1e-7
    assert np.all(g0[:2 * N] == 0)


def test_fluctuation_gradient_is_derivative_of_sigma():
    """For Hermitian H and normalised psi the reference's fluctuation gradient
    equals d sigma/dRe p + i d sigma/dIm p of sigma = sqrt(<H^2> - <H>^2) from
    dense matrices; PsiDeep included."""
    N, M = 4, 3
    b, W, _ = random_rbm(N, M, seed=22, scale=0.3)
    tab = HAM.tfim_1d(N, 1.0, 0.8)
    tab = (np.concatenate([tab[0], [0.3]]), np.concatenate([tab[1], [[2, 2, 0]]]),
           np.concatenate([tab[2], [[0, 1, -1]]]))
    H = dense_hamiltonian(N, *tab)
    assert np.allclose(H, H.conj().T)
    params = np.concatenate([b, W.ravel()])

    def model(p):
        return O.RBM(p[:M], p[M:].reshape(N, M))

    def sigma(p):
        v = O.psi_vector(model(p))
        v = v / np.linalg.norm(v)
        e = np.vdot(v, H @ v).real
        return np.sqrt(np.vdot(H @ v, H @ v).real - e * e)

    psi = model(params)
    psi.prefactor = 1.0 / np.linalg.norm(O.psi_vector(psi))
    g = O.fluctuation_gradient(psi, O.Operator(*tab), O.Exact(N))
    f, _ = O.fluctuation(psi, O.Operator(*tab), O.Exact(N))
    assert abs(f - sigma(params)) < 1e-12
    h = 1e-6
    for k in range(len(params)):
        d = np.zeros_like(params)
        d[k] = h
        fd = (sigma(params + d) - sigma(params - d)) / (2 * h)
        fd += 1j * (sigma(params + 1j * d) - sigma(params - 1j * d)) / (2 * h)
        assert abs(g[2 * N + k] - fd) < 1e-7
    assert np.all(g[:2 * N] == 0)
    # PsiDeep: the same identity on one weight and one bias per layer
    Nd, Md, Cd = 4, [6, 3], [2, 6]
    biases, conns, weights, final = deep_layers(Nd, Md, Cd, seed=3, scale=0.3)
    deep = O.Deep(Nd, biases, conns, weights, final, ref_compat_final=False)
    deep.prefactor = 1.0 / np.linalg.norm(O.psi_vector(deep))
    gd = O.fluctuation_gradient(deep, O.Operator(*tab), O.Exact(Nd))

    def sigma_d(wl, dz):
        w2 = [x.copy() for x in weights]
        w2[wl][0, 0] += dz
        v = O.psi_vector(O.Deep(Nd, biases, conns, w2, final))
        v = v / np.linalg.norm(v)
        e = np.vdot(v, H @ v).real
        return np.sqrt(np.vdot(H @ v, H @ v).real - e * e)

    for wl, k in ((0, 2 * Nd + Md[0]), (1, 2 * Nd + Md[0] + Cd[0] * Md[0] + Md[1])):
        fd = (sigma_d(wl, h) - sigma_d(wl, -h)) / (2 * h) + 1j * (sigma_d(wl, 1j * h) - sigma_d(wl, -1j * h)) / (2 * h)
        assert abs(gd[k] - fd) < 1e-7


def test_difference_invariants():
    """difference(psi, psi', ops): zero for psi' = psi (any operator, any
    prefactor of psi'); zero for the identity and any psi' on a normalised psi
    (a = <w> = 1 and b / <r> = 1); local energies are pinned separately."""
    N, M = 5, 4
    b, W, _ = random_rbm(N, M, seed=41, scale=0.3)
    psi = O.RBM(b, W)
    psi.prefactor = 1.0 / np.linalg.norm(O.psi_vector(psi))
    ops = [O.Operator(*HAM.tfim_1d(N)), O.Operator(*random_operator(N, 6, 3, seed=2))]
    d = O.difference(psi, O.RBM(b, W), ops, O.Exact(N))
    assert np.max(np.abs(d)) < 1e-12
    other = O.RBM(b, W, prefactor=3.0)       # the prefactor does not enter log psi'
    assert np.max(np.abs(O.difference(psi, other, ops, O.Exact(N)))) < 1e-12
    b2 = b + 0.7 - 0.2j
    ident = O.Operator(np.ones(1, complex), np.zeros((1, 1), np.int32), -np.ones((1, 1), np.int32))
    d2 = O.difference(psi, O.RBM(b2, W), [ident], O.Exact(N))
    assert abs(d2[0]) < 1e-12
    # a genuinely different psi' gives a non-zero result for a non-trivial operator
    assert abs(O.difference(psi, O.RBM(b2, W), ops, O.Exact(N))[0]) > 1e-3


def test_corrected_identity_and_closed_form():
    """corrected with the identity: E = 1 so X = 1 and the value is 1; for a
    diagonal operator E is the classical energy and X = exp(E^2 - E)."""
    N, M = 4, 3
    b, W, _ = random_rbm(N, M, seed=43, scale=0.3)
    psi = O.RBM(b, W)
    psi.prefactor = 1.0 / np.linalg.norm(O.psi_vector(psi))
    ident = O.Operator(np.ones(1, complex), np.zeros((1, 1), np.int32), -np.ones((1, 1), np.int32))
    v, x = O.corrected(psi, ident, O.Exact(N))
    assert abs(v - 1.0) < 1e-12 and abs(x - 1.0) < 1e-12
    zz = HAM.table([(0.5, [(3, i), (3, (i + 1) % N)]) for i in range(N)])
    v, x = O.corrected(psi, O.Operator(*zz), O.Exact(N))
    p = np.abs(O.psi_vector(psi)) ** 2
    s = basis_spins(N)
    E = 0.5 * (s * np.roll(s, -1, axis=1)).sum(axis=1)
    X = (p * np.exp(E * E - E)).sum()
    assert abs(x - X) < 1e-12 * X and abs(v - (p * E).sum() / X) < 1e-12


def _dense_local_energies(N, tab, b, W):
    """psi vector without prefactor and E(s) = (H psi)(s) / psi(s) from the
    dense matrix of the table."""
    v = O.psi_vector(O.RBM(b, W))
    H = dense_hamiltonian(N, *tab)
    return v, (H @ v) / v


def test_difference_dense_closed_form():
    """ExpectationValue::difference (ExpectationValue.cu:334-446) on the exact
    basis: with p = |prefactor psi|^2, r = |psi / psi'|^2 (log psi without
    prefactor) and E, E' the local energies of psi and psi', the result is
    sum p E' - (sum p r E) / (sum p r).  The same sums with r inverted, or with
    E and E' swapped, differ by more than 1e-3, so either mistake fails."""
    N, M = 5, 4
    b, W, _ = random_rbm(N, M, seed=51, scale=0.3)
    b2, W2, _ = random_rbm(N, M, seed=52, scale=0.3)
    tabs = [HAM.tfim_1d(N, 1.0, 0.8), random_operator(N, 6, 3, seed=53)]
    psi = O.RBM(b, W)
    psi.prefactor = 1.0 / np.linalg.norm(O.psi_vector(psi))
    got = O.difference(psi, O.RBM(b2, W2, prefactor=2.5), [O.Operator(*t) for t in tabs], O.Exact(N))
    for k, tab in enumerate(tabs):
        v, E = _dense_local_energies(N, tab, b, W)
        v2, E2 = _dense_local_energies(N, tab, b2, W2)
        p = np.abs(psi.prefactor * v) ** 2
        r = np.abs(v) ** 2 / np.abs(v2) ** 2

        def form(rr, e_own, e_other):
            return (p * e_other).sum() - (p * rr * e_own).sum() / (p * rr).sum()

        want = form(r, E, E2)
        assert abs(got[k] - want) < 1e-12 * max(1.0, abs(want)), (k, got[k], want)
        assert abs(form(1 / r, E, E2) - want) > 1e-3
        assert abs(form(r, E2, E) - want) > 1e-3


def test_corrected_dense_closed_form_with_complex_local_energy():
    """ExpectationValue::corrected (ExpectationValue.cu:74-127) on the exact
    basis with a complex psi, so E = (H psi)/psi has an imaginary part:
    X = exp(2 Im E + Re(E^2 - E)) (both local energies from the first
    operator), value = sum p E / sum p X, norm = sum p X.  Dropping the 2 Im E
    term changes the norm by more than 1 %."""
    N, M = 5, 4
    b, W, _ = random_rbm(N, M, seed=55, scale=0.3)
    psi = O.RBM(b, W)
    psi.prefactor = 1.0 / np.linalg.norm(O.psi_vector(psi))
    for tab in (HAM.tfim_1d(N, 1.0, 0.8), random_operator(N, 5, 2, seed=56)):
        v, E = _dense_local_energies(N, tab, b, W)
        assert np.max(np.abs(E.imag)) > 0.1
        p = np.abs(psi.prefactor * v) ** 2
        X = (p * np.exp(2 * E.imag + (E * E - E).real)).sum()
        value, norm = O.corrected(psi, O.Operator(*tab), O.Exact(N))
        assert abs(norm - X) < 1e-12 * X
        assert abs(value - (p * E).sum() / X) < 1e-12 * max(1.0, abs(value))
        X_no_im = (p * np.exp((E * E - E).real)).sum()
        assert abs(X_no_im - X) > 1e-2 * X


def test_s_matrix_outer_products():
    N, M = 3, 4
    b, W, _ = random_rbm(N, M, seed=31)
    psi = O.RBM(b, W)
    S = O.s_matrix(psi, O.Exact(N))
    ref = np.zeros_like(S)
    for w in basis_words(N):
        ok = O.o_k(psi, w)
        ref += np.outer(ok, ok.conj())
    assert rel_err(S, ref / 2 ** N) < 1e-13


def test_hs_distance_identity_closed_form_and_gradient():
    """Identity operator, unitary: d = sqrt(1 - |<psi|psi'>|^2 / (<psi|psi><psi'|psi'>))
    (test_HilbertSpaceDistance.py:182-242 uses the same setting)."""
    N, M = 4, 5
    b, W, _ = random_rbm(N, M, seed=41, scale=0.3)
    b2, W2, _ = random_rbm(N, M, seed=42, scale=0.3)
    ident = O.Operator(np.array([1.0 + 0j]), np.zeros((1, 1), np.int32), -np.ones((1, 1), np.int32))
    psi = O.RBM(b, W)
    params = np.concatenate([b2, W2.ravel()])

    def model(p):
        return O.RBM(p[:M], p[M:].reshape(N, M))

    def dist(p):
        v, v2 = O.psi_vector(psi), O.psi_vector(model(p))
        return np.sqrt(1 - abs(np.vdot(v2, v)) ** 2 / (np.vdot(v, v).real * np.vdot(v2, v2).real))

    g, d = O.hs_distance(psi, model(params), ident, True, O.Exact(N), want_grad=True, nu=1.0)
    assert abs(d - dist(params)) < 1e-12
    h = 1e-6
    for k in range(len(params)):
        dd = np.zeros_like(params)
        dd[k] = h
        fd = (dist(params + dd) - dist(params - dd)) / (2 * h) + 1j * (dist(params + 1j * dd) - dist(params - 1j * dd)) / (2 * h)
        assert abs(g[2 * N + k] - fd) < 1e-6


def _site_op(A, i, N):
    """Single-site operator A on spin i of N as a dense matrix over basis index
    x = sum_j bit_j 2^j (spin j = +1 iff bit j is set)."""
    out = np.ones((1, 1))
    for j in reversed(range(N)):
        out = np.kron(out, A if j == i else np.eye(2))
    return out


def _axis_closed_form(v, vp, N, alpha, beta, alpha_p, beta_p):
    """Identity-operator, unitary distance with psi'_c = psi' exp((K psi') / psi'),
    K = sum_i [da_i Z_i X_i + i db_i (cos_i X_i - sin_i Z_i)] built from dense
    Pauli matrices (angles may be complex)."""
    Z = np.diag([-1.0, 1.0])
    X = np.array([[0.0, 1.0], [1.0, 0.0]])
    K = np.zeros((2 ** N, 2 ** N), complex)
    for i in range(N):
        da, db = alpha_p[i] - alpha[i], beta_p[i] - beta[i]
        c, s = np.cos(alpha_p[i] + alpha[i]), np.sin(alpha_p[i] + alpha[i])
        K += da * _site_op(Z @ X, i, N) + 1j * db * (c * _site_op(X, i, N) - s * _site_op(Z, i, N))
    vc = vp * np.exp((K @ vp) / vp)
    return np.sqrt(1 - abs(np.vdot(vc, v)) ** 2 / (np.vdot(v, v).real * np.vdot(vc, vc).real))


def test_hs_distance_free_axis_closed_form_and_gradient():
    """free_quantum_axis (HilbertSpaceDistance.cu:47-80, 109-122): the distance
    against the dense-operator closed form above, the alpha / beta gradient
    entries against complex central differences of that closed form (beta at
    any delta beta; alpha at delta beta = 0, where the reference's O_alpha is
    the exact derivative), and zero axis change giving the plain distance."""
    N, M = 4, 5
    b, W, _ = random_rbm(N, M, seed=61, scale=0.3)
    b2, W2, _ = random_rbm(N, M, seed=62, scale=0.3)
    ident = O.Operator(np.array([1.0 + 0j]), np.zeros((1, 1), np.int32), -np.ones((1, 1), np.int32))
    psi, psi2 = O.RBM(b, W), O.RBM(b2, W2)
    v, vp = O.psi_vector(psi), O.psi_vector(psi2)
    rng = np.random.default_rng(63)
    alpha, beta = rng.uniform(-0.4, 0.4, N), rng.uniform(-0.4, 0.4, N)
    alpha_p, beta_p = alpha + rng.uniform(-0.1, 0.1, N), beta + rng.uniform(-0.1, 0.1, N)
    g, d = O.hs_distance(psi, psi2, ident, True, O.Exact(N), want_grad=True, nu=1.0,
                         axis=(alpha, beta, alpha_p, beta_p))
    assert abs(d - _axis_closed_form(v, vp, N, alpha, beta, alpha_p, beta_p)) < 1e-12
    h = 1e-6

    def fd(f, x, i):
        e = np.zeros(N)
        e[i] = h
        return (f(x + e) - f(x - e)) / (2 * h) + 1j * (f(x + 1j * e) - f(x - 1j * e)) / (2 * h)

    for i in range(N):
        gb = fd(lambda x: _axis_closed_form(v, vp, N, alpha, beta, alpha_p, x), beta_p.astype(complex), i)
        assert abs(g[N + i] - gb) < 1e-6, (i, g[N + i], gb)
    g0, d0 = O.hs_distance(psi, psi2, ident, True, O.Exact(N), want_grad=True, nu=1.0,
                           axis=(alpha, beta, alpha_p, beta))
    for i in range(N):
        ga = fd(lambda x: _axis_closed_form(v, vp, N, alpha, beta, x, beta), alpha_p.astype(complex), i)
        assert abs(g0[i] - ga) < 1e-6, (i, g0[i], ga)
    # no change of axis: the plain distance and the plain parameter gradient
    gz, dz = O.hs_distance(psi, psi2, ident, True, O.Exact(N), want_grad=True, nu=1.0,
                           axis=(alpha, beta, alpha, beta))
    gp, dp = O.hs_distance(psi, psi2, ident, True, O.Exact(N), want_grad=True, nu=1.0)
    assert abs(dz - dp) < 1e-14 and np.allclose(gz[2 * N:], gp[2 * N:], rtol=0, atol=1e-14)


def test_hs_distance_small_time_step():
    """Non-unitary operator i t H against sqrt(1 - |<psi|exp(iHt)|psi>|^2),
    the reference's test_distance1 (test_HilbertSpaceDistance.py:10-29)."""
    N, M = 5, 4
    b, W, _ = random_rbm(N, M, seed=51, scale=0.3)
    psi = O.RBM(b, W)
    v = O.psi_vector(psi)
    psi.prefactor = 1 / np.linalg.norm(v)
    tab = HAM.heisenberg([(i, (i + 1) % N) for i in range(N)])
    t = 1e-3
    H = dense_hamiltonian(N, *tab)
    e, U = np.linalg.eigh(H)
    Ut = U @ np.diag(np.exp(1j * e * t)) @ U.conj().T
    vn = v / np.linalg.norm(v)
    ref = np.sqrt(1 - abs(np.vdot(vn, Ut @ vn)) ** 2)
    op = O.Operator(1j * t * tab[0], tab[1], tab[2])
    d = O.hs_distance(psi, psi, op, False, O.Exact(N))
    assert d == pytest.approx(ref, rel=1e-3, abs=1e-8)


def test_metropolis_stationary_distribution():
    """Visit frequencies of independent chains against |psi|^2: per-bin
    z-scores with the standard error taken across chains (chains are
    independent, samples within a chain are not)."""
    N, M = 4, 3
    b, W, _ = random_rbm(N, M, seed=61, scale=0.6)
    psi = O.RBM(b, W)
    p = np.abs(O.psi_vector(psi)) ** 2
    p /= p.sum()
    chains, per = 256, 300
    mc = O.MonteCarlo(N, chains * per, 2, 10, chains, seed=7)
    spins, _, _, _ = O.records(psi, None, mc)
    idx = spins[:, 0].astype(np.int64).reshape(per, chains)
    freq = np.stack([np.bincount(idx[:, c], minlength=2 ** N) / per for c in range(chains)])
    mean, se = freq.mean(axis=0), freq.std(axis=0, ddof=1) / np.sqrt(chains)
    z = np.abs(mean - p) / np.maximum(se, 1e-12)
    assert z.max() < 4.5, z
    assert mc.acceptances + mc.rejections == chains * (10 + per * 2) * N


@pytest.mark.parametrize("move", ["exchange", "fast_sweep"])
def test_metropolis_moves_stationary_distribution(move):
    """The exchange move keeps every sample in its total S_z sector and visits
    the sector's configurations with frequencies |psi|^2 / Z_sector; the fast
    sweep visits the whole basis with |psi|^2.  Same z-score test as above."""
    N, M = 6, 4
    b, W, _ = random_rbm(N, M, seed=63, scale=0.6)
    psi = O.RBM(b, W)
    p = np.abs(O.psi_vector(psi)) ** 2
    chains, per = 256, 200
    if move == "exchange":
        sector = 0
        up = np.array([bin(i).count("1") for i in range(2 ** N)])
        p = np.where(2 * up - N == sector, p, 0.0)
        mc = O.MonteCarlo(N, chains * per, 1, 10, chains, seed=8, total_z=sector)
    else:
        mc = O.MonteCarlo(N, chains * per, 3, 10, chains, seed=8, fast_sweep=2)
    p /= p.sum()
    spins, _, _, _ = O.records(psi, None, mc)
    idx = spins[:, 0].astype(np.int64).reshape(per, chains)
    if move == "exchange":
        assert np.all(p[idx] > 0)               # never leaves the sector
    freq = np.stack([np.bincount(idx[:, c], minlength=2 ** N) / per for c in range(chains)])
    mean, se = freq.mean(axis=0), freq.std(axis=0, ddof=1) / np.sqrt(chains)
    # rare configurations that no chain visited: floor with the binomial error
    se = np.maximum(se, np.sqrt(p * (1 - p) / (chains * per)))
    z = np.abs(mean - p) / np.maximum(se, 1e-12)
    assert z.max() < 4.5, z
    per_sweep = N if move == "exchange" else 1
    sweeps = 1 if move == "exchange" else 3
    assert mc.acceptances + mc.rejections == chains * (10 * N + per * sweeps * per_sweep)


# ---------------------------------------------------------------- PsiDeep
@pytest.mark.parametrize("ti", [False, True])
def test_deep_log_psi_closed_form(ti):
    """test_PsiDeep.py:20-49: layer j reads activations[(j * delta + i) % n],
    translational invariance averages np.roll shifts."""
    N, M, C = 6, [12, 4], [3, 6]
    biases, conns, weights, final = deep_layers(N, M, C, seed=3)
    psi = O.Deep(N, biases, conns, weights, final, translational_invariance=ti)
    S = basis_spins(N)
    ref = np.zeros(2 ** N, dtype=complex)
    for s_idx in range(2 ** N):
        acc = 0
        for shift in range(N if ti else 1):
            x = np.roll(S[s_idx], shift)
            for w, b, cn in zip(weights, biases, conns):
                a = np.array([sum(w[i, j] * x[cn[i, j]] for i in range(w.shape[0])) + b[j] for j in range(len(b))])
                x = O.logcosh(a)
            acc += np.sum(x * final)
        ref[s_idx] = acc / (N if ti else 1)
    assert rel_err(O.log_psi(psi, basis_words(N)), ref) < 1e-13


def test_deep_o_k_finite_differences_and_final_weight_convention():
    N, M, C = 5, [10, 5], [2, 4]
    biases, conns, weights, final = deep_layers(N, M, C, seed=8)

    def split(p):
        out_b, out_w, off = [], [], 0
        for m, c in zip(M, C):
            out_b.append(p[off:off + m]); off += m
            out_w.append(p[off:off + c * m].reshape(c, m)); off += c * m
        return out_b, out_w, p[off:]

    params = np.concatenate([np.concatenate([b, w.ravel()]) for b, w in zip(biases, weights)] + [final])

    def model(p, compat=False):
        bb, ww, ff = split(p)
        return O.Deep(N, bb, conns, ww, ff, ref_compat_final=compat)

    w = O.spins_words([13])[0]
    ok = O.o_k(model(params), w)
    for k in range(len(params)):
        fd = _fd_logpsi(model, params, w[None, :], k)[0]
        assert abs(fd - ok[2 * N + k]) < 1e-6
    ok_ref = O.o_k(model(params, compat=True), w)
    lp = O.log_psi(model(params), w[None, :])[0]
    nf = len(final)
    assert np.allclose(ok_ref[:-nf], ok[:-nf], rtol=0, atol=0)
    assert np.allclose(ok_ref[-nf:], lp * ok[-nf:], rtol=1e-14, atol=0)


def test_deep_local_energy_dense():
    N, M, C = 5, [10, 5], [5, 10]
    biases, conns, weights, final = deep_layers(N, M, C, seed=9)
    psi = O.Deep(N, biases, conns, weights, final)
    tab = random_operator(N, 10, 3, seed=9)
    H = dense_hamiltonian(N, *tab)
    v = O.psi_vector(psi)
    got = O.local_energy(psi, O.Operator(*tab), basis_words(N))
    assert rel_err(got, (H @ v) / v) < 1e-12
