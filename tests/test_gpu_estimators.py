"""GPU parity of the E_k gradient estimator (any operator), the fluctuation
gradient, `corrected` and `difference` against the FP64 oracle, on the exact
basis and on Monte Carlo chains that the complex128 kernels reproduce bit for
bit (tolerance 1e-10 there).  complex64 kernels feed FP64 sums here, so their
error is the FP32 rounding of log psi differences and tanh records, held to
1e-5.  The E_k estimator is the difference of three sums,
<E_loc conj O_k> + conj<E_k(s')> - 2 <E_loc> conj<O_k> (ExpectationValue.cu:
246-250), which is not centred: its complex64 error is held to 1e-5 of the
three sums' magnitudes, computed from the oracle's E_loc and O_k on the exact
basis (a dropped or doubled term exceeds that by orders of magnitude)."""
import numpy as np
import pytest

from helpers import O, deep_layers, dense_hamiltonian, random_operator, random_rbm, rel_err

import rbmongpu_amd as R
from rbmongpu_amd import hamiltonians as HAM

pytestmark = pytest.mark.gpu
TOL = {64: 1e-10, 32: 1e-5}


def gpu_psi(b, W, a=None, precision=64, prefactor=1.0):
    N = W.shape[0]
    return R.Psi(np.zeros(N), np.zeros(N), b, W, prefactor, False, True, a=a, precision=precision)


def gpu_deep(N, biases, conns, weights, final, precision):
    return R.PsiDeep(np.zeros(N), np.zeros(N), list(biases), [np.asarray(c, np.uint32) for c in conns], list(weights),
                     np.asarray(final, complex), 1.0, False, True, precision=precision, ref_compat_final=False)


def non_hermitian(N, seed):
    tab = random_operator(N, 8, 3, seed=seed)
    H = dense_hamiltonian(N, *tab)
    assert not np.allclose(H, H.conj().T)
    return tab


def ek_term_scale(ref, op, N, g_ref):
    """|<E conj O>| + |conj<E_k>| + 2 |<E>| |<O>| per parameter on the exact
    basis: the magnitudes of the estimator's three sums (averages are weighted
    sums over the 2^N states divided by 2^N, as ExactSummation's).  <E_k> is
    what g_ref leaves after the other two."""
    spins, _, el, w = O.records(ref, op, O.Exact(N))
    w = w / len(w)
    Ok = np.array([O.o_k(ref, s) for s in spins])
    e_o = (w * el) @ Ok.conj()
    two_e_o = 2 * (w @ el) * (w @ Ok).conj()
    return np.abs(e_o) + np.abs(g_ref - e_o + two_e_o) + np.abs(two_e_o)


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("model", ["rbm", "rbm_visible", "deep"])
def test_ek_gradient_exact(precision, model):
    print(f"[E_k gradient exact] precision={precision} model={model}", flush=True)
    N = 7
    if model == "deep":
        biases, conns, weights, final = deep_layers(N, [14, 7], [4, 14], seed=5, scale=0.25)
        psi = gpu_deep(N, biases, conns, weights, final, precision)
        ref = O.Deep(N, biases, conns, weights, final, ref_compat_final=False)
    else:
        b, W, a = random_rbm(N, 9, seed=13, scale=0.3, visible=model == "rbm_visible")
        psi = gpu_psi(b, W, a, precision)
        ref = O.RBM(b, W, a)
    ex = R.ExactSummation(N, True)
    psi.normalize(ex)
    ref.prefactor = psi.prefactor
    ev = R.ExpectationValue(True)
    for tab, est in ((HAM.tfim_1d(N, 1.0, 0.7), "E_k"), (non_hermitian(N, 3), "auto")):
        g, E = ev.gradient(psi, R.Operator(*tab, True), ex, estimator=est)
        gr, Er = O.gradient(ref, O.Operator(*tab), O.Exact(N), mode=0)
        scale = ek_term_scale(ref, O.Operator(*tab), N, gr)
        err = np.linalg.norm(g - gr) / np.linalg.norm(scale)
        print(f"  {est}: |g - g_ref| / |terms| = {err:.2e} (relative to |g_ref|: {rel_err(g, gr):.2e})", flush=True)
        assert abs(E - Er) < TOL[precision] * abs(Er)
        assert err < TOL[precision], (est, err)
    # a non-Hermitian operator cannot use the Hermitian estimator
    with pytest.raises(ValueError):
        ev.gradient(psi, R.Operator(*non_hermitian(N, 3), True), ex, estimator="hermitian")


def test_ek_gradient_monte_carlo_bit_identical_chains():
    N, M = 8, 10
    b, W, _ = random_rbm(N, M, seed=17, scale=0.3)
    tab = non_hermitian(N, 5)
    psi = gpu_psi(b, W, precision=64)
    chains, per = 64, 20
    mc = R.MonteCarloLoop(chains * per, 1, 5, chains, True, seed=23)
    g, E = R.ExpectationValue(True).gradient(psi, R.Operator(*tab, True), mc)
    gr, Er = O.gradient(O.RBM(b, W), O.Operator(*tab), O.MonteCarlo(N, chains * per, 1, 5, chains, seed=23), mode=0)
    assert abs(E - Er) < 1e-10 * abs(Er)
    assert rel_err(g, gr) < 1e-10


@pytest.mark.parametrize("precision", [64, 32])
def test_fluctuation_gradient(precision):
    print(f"[fluctuation gradient] precision={precision}", flush=True)
    N, M = 7, 9
    b, W, _ = random_rbm(N, M, seed=19, scale=0.3)
    tab = HAM.tfim_1d(N, 1.0, 0.9)
    psi = gpu_psi(b, W, precision=precision)
    ex = R.ExactSummation(N, True)
    psi.normalize(ex)
    g = R.ExpectationValue(True).fluctuation_gradient(psi, R.Operator(*tab, True), ex)
    gr = O.fluctuation_gradient(O.RBM(b, W, prefactor=psi.prefactor), O.Operator(*tab), O.Exact(N))
    assert rel_err(g, gr) < TOL[precision]
    # PsiDeep and Monte Carlo (complex128 chains are bit-identical to the oracle's)
    biases, conns, weights, final = deep_layers(N, [14, 7], [4, 14], seed=6, scale=0.25)
    deep = gpu_deep(N, biases, conns, weights, final, 64)
    chains, per = 32, 10
    mc = R.MonteCarloLoop(chains * per, 1, 4, chains, True, seed=29)
    g = R.ExpectationValue(True).fluctuation_gradient(deep, R.Operator(*tab, True), mc)
    gr = O.fluctuation_gradient(O.Deep(N, biases, conns, weights, final, ref_compat_final=False), O.Operator(*tab),
                                O.MonteCarlo(N, chains * per, 1, 4, chains, seed=29))
    assert rel_err(g, gr) < 1e-10


@pytest.mark.parametrize("precision", [64, 32])
def test_corrected(precision):
    N = 7
    biases, conns, weights, final = deep_layers(N, [14, 7], [4, 14], seed=7, scale=0.2)
    psi = gpu_deep(N, biases, conns, weights, final, precision)
    ex = R.ExactSummation(N, True)
    psi.normalize(ex)
    tab = HAM.heisenberg([(i, (i + 1) % N) for i in range(N)], 0.3)
    v, x = R.ExpectationValue(True).corrected(psi, R.Operator(*tab, True), R.Operator(*HAM.tfim_1d(N), True), ex)
    ref = O.Deep(N, biases, conns, weights, final, prefactor=psi.prefactor, ref_compat_final=False)
    vr, xr = O.corrected(ref, O.Operator(*tab), O.Exact(N))
    assert abs(x - xr) < TOL[precision] * abs(xr)
    assert abs(v - vr) < TOL[precision] * abs(vr)


@pytest.mark.parametrize("precision", [64, 32])
def test_difference(precision):
    N, M = 7, 9
    b, W, _ = random_rbm(N, M, seed=31, scale=0.3)
    b2, W2, _ = random_rbm(N, M, seed=32, scale=0.3)
    psi, psi2 = gpu_psi(b, W, precision=precision), gpu_psi(b2, W2, precision=precision)
    ex = R.ExactSummation(N, True)
    psi.normalize(ex)
    tabs = [HAM.tfim_1d(N), random_operator(N, 6, 3, seed=8)]
    ops = [R.Operator(*t, True) for t in tabs]
    d = R.ExpectationValue(True).difference(psi, psi2, ops, ex)
    dr = O.difference(O.RBM(b, W, prefactor=psi.prefactor), O.RBM(b2, W2), [O.Operator(*t) for t in tabs], O.Exact(N))
    assert rel_err(np.array(d), np.array(dr)) < TOL[precision]
    if precision == 64:
        chains, per = 32, 10
        mc = R.MonteCarloLoop(chains * per, 1, 4, chains, True, seed=37)
        d = R.ExpectationValue(True).difference(psi, psi2, ops, mc)
        dr = O.difference(O.RBM(b, W, prefactor=psi.prefactor), O.RBM(b2, W2), [O.Operator(*t) for t in tabs],
                          O.MonteCarlo(N, chains * per, 1, 4, chains, seed=37))
        assert rel_err(np.array(d), np.array(dr)) < 1e-10
