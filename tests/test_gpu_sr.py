"""GPU parity of the stochastic-reconfiguration pieces: FP64 Cholesky solve
against numpy.linalg.solve, the centred S-matrix against numpy's weighted
covariance of oracle O_k vectors (ExactSummation), and the SR update against a
dense solve of the same system.

Tolerances: the Cholesky solve is FP64 end to end (1e-10 relative); S is built
from FP32 MFMA products accumulated in FP64 (2e-5 relative to max|S|); the SR
update inherits S's error through (S + eps I)^{-1} (1e-3 relative at eps = 1e-2)."""
import numpy as np
import pytest

from helpers import O, basis_words, random_rbm, rel_err

import rbmongpu_amd as R
from rbmongpu_amd import hamiltonians as HAM

pytestmark = pytest.mark.gpu


def gpu_psi(b, W, a=None, precision=64):
    N = W.shape[0]
    return R.Psi(np.zeros(N), np.zeros(N), b, W, 1.0, False, True, a=a, precision=precision)


def hpd(n, seed):
    rng = np.random.default_rng(seed)
    B = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    return B @ B.conj().T / n + 0.5 * np.eye(n), rng.normal(size=n) + 1j * rng.normal(size=n)


@pytest.mark.parametrize("n", [1, 5, 31, 32, 33, 64, 100, 129, 160, 257, 600, 1000])
def test_cholesky_solve_matches_numpy(n):
    print(f"[cholesky] n={n}", flush=True)
    A, b = hpd(n, n)
    x = R.cholesky_solve(A, b)
    ref = np.linalg.solve(A, b)
    assert rel_err(x, ref) < 1e-10
    assert rel_err(A @ x, b) < 1e-10


def test_cholesky_rejects_indefinite_matrix():
    A, b = hpd(40, 3)
    A[17, 17] = -5.0
    with pytest.raises(ValueError):
        R.cholesky_solve(A, b)


def exact_O_and_prob(b, W, a):
    N = W.shape[0]
    words = basis_words(N)
    model = O.RBM(b, W, a)
    lp = O.log_psi(model, words)
    prob = np.exp(2 * (lp.real - lp.real.max()))
    prob /= prob.sum()
    Ok = np.array([O.o_k(model, w)[2 * N:] for w in words])
    return Ok, prob


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("visible", [False, True])
def test_centred_s_matrix_is_the_weighted_covariance(precision, visible):
    print(f"[S centred] precision={precision} visible={visible}", flush=True)
    N, M = 8, 10
    b, W, a = random_rbm(N, M, seed=17, scale=0.3, visible=visible)
    psi = gpu_psi(b, W, a, precision=precision)
    ex = R.ExactSummation(N, True)
    psi.normalize(ex)
    S = R.get_S_matrix_centred(psi, ex)
    Ok, prob = exact_O_and_prob(b, W, a)
    ref = np.cov(Ok.T, aweights=prob, bias=True).conj()
    assert S.shape == ref.shape
    assert rel_err(S, ref) < 2e-5
    assert np.allclose(S, S.conj().T, atol=1e-12 * np.abs(S).max())


def test_sr_update_solves_the_shifted_system():
    N, M, eps = 8, 12, 1e-2
    b, W, _ = random_rbm(N, M, seed=23, scale=0.3)
    tab = HAM.tfim_1d(N, 1.0, 0.9)
    psi = gpu_psi(b, W, precision=64)
    ex = R.ExactSummation(N, True)
    psi.normalize(ex)
    upd, E = R.sr_update(psi, R.Operator(*tab, True), ex, eps)
    Ok, prob = exact_O_and_prob(b, W, None)
    S = np.cov(Ok.T, aweights=prob, bias=True).conj()
    g, Er = O.gradient(O.RBM(b, W, prefactor=psi.prefactor), O.Operator(*tab), O.Exact(N), mode=1)
    ref = np.linalg.solve(S + eps * np.eye(S.shape[0]), g[2 * N:])
    assert abs(E - Er) < 1e-10 * abs(Er)
    assert np.all(upd[:2 * N] == 0)
    assert rel_err(upd[2 * N:], ref) < 1e-3
    # the update lowers the energy of a small step along -update
    g_gpu, _ = R.ExpectationValue(True).gradient(psi, R.Operator(*tab, True), ex)
    assert np.real(np.vdot(g_gpu, upd)) > 0


def test_sr_update_on_sampled_ensembles():
    N, M = 12, 24
    b, W, _ = random_rbm(N, M, seed=29, scale=0.1)
    tab = HAM.tfim_1d(N)
    psi = gpu_psi(b, W, precision=32)
    op = R.Operator(*tab, True)
    mc = R.MonteCarloLoop(512 * 40, 1, 20, 512, True, seed=2)
    upd, E = R.sr_update(psi, op, mc, 1e-3)
    assert np.all(np.isfinite(upd)) and np.abs(upd[2 * N:]).max() > 0
    md = R.MADE(N, 64, seed=1)
    lrn = R.LearnedMonteCarloLoop(md, 512 * 40, 10, 512, True, seed=3)
    upd2, E2 = R.sr_update(psi, op, lrn, 1e-3)
    assert np.all(np.isfinite(upd2))
    assert abs(E.real - E2.real) < 0.05 * abs(E.real)
