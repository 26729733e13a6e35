"""GPU parity of the PsiDeep path against the FP64 oracle (tolerances as in
test_gpu_rbm.py: 1e-10 for complex128 kernels, 1e-5 for complex64)."""
import os

import numpy as np
import pytest

from helpers import O, basis_words, deep_layers, random_operator, rel_err

import rbmongpu_amd as R
from rbmongpu_amd import hamiltonians as HAM

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TOL = {64: 1e-10, 32: 1e-5}


def gpu_deep(N, biases, conns, weights, final, precision, prefactor=1.0, ti=False, compat=True):
    psi = R.PsiDeep(np.zeros(N), np.zeros(N), list(biases), [np.asarray(c, np.uint32) for c in conns], list(weights),
                    np.asarray(final, complex), prefactor, False, True, precision=precision, ref_compat_final=compat)
    psi.translational_invariance = ti
    return psi


@pytest.mark.parametrize("precision", [64, 32])
def test_deep_golden(precision):
    print(f"[deep golden] precision={precision}", flush=True)
    d = np.load(os.path.join(GOLDEN, "deep_small.npz"))
    L, N = 2, 6
    psi = gpu_deep(N, [d[f"b{l}"] for l in range(L)], [d[f"c{l}"] for l in range(L)],
                   [d[f"w{l}"] for l in range(L)], d["final"], precision)
    tol = TOL[precision]
    assert rel_err(R.log_psi_batch(psi, basis_words(N)), d["log_psi"]) < tol
    op = R.Operator(d["coeffs"], d["types"], d["idx"], True)
    assert rel_err(R.local_energy_batch(psi, op, basis_words(N)), d["e_loc"]) < tol
    for c, ok in zip(d["ok_configs"], d["o_k"]):
        assert rel_err(R.psi_O_k_vector(psi, R.Spins(int(c), N)), ok[2 * N:]) < tol
    psi.prefactor = float(d["prefactor"])
    g, E = R.ExpectationValue(True).gradient(psi, op, R.ExactSummation(N, True))
    assert abs(E - d["energy"]) < tol * abs(d["energy"])
    assert rel_err(g, d["gradient"]) < tol


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("layout", ["sparse3", "dense2", "wide"])
def test_deep_layouts_and_translational_invariance(precision, layout):
    print(f"[deep layouts] precision={precision} layout={layout}", flush=True)
    if layout == "sparse3":
        N, M, C = 8, [16, 8, 4], [4, 2, 4]
    elif layout == "dense2":
        N, M, C = 7, [14, 7], [7, 14]
    else:
        N, M, C = 9, [36, 18], [9, 36]
    biases, conns, weights, final = deep_layers(N, M, C, seed=len(M) * 10 + N, scale=0.2)
    words = basis_words(N)
    for ti in (False, True):
        ref = O.Deep(N, biases, conns, weights, final, translational_invariance=ti)
        psi = gpu_deep(N, biases, conns, weights, final, precision, ti=ti)
        assert rel_err(R.log_psi_batch(psi, words), O.log_psi(ref, words)) < TOL[precision]
    tab = HAM.heisenberg([(i, (i + 1) % N) for i in range(N)])
    psi = gpu_deep(N, biases, conns, weights, final, precision, compat=False)
    ref = O.Deep(N, biases, conns, weights, final, ref_compat_final=False)
    ex, oex = R.ExactSummation(N, True), O.Exact(N)
    g, E = R.ExpectationValue(True).gradient(psi, R.Operator(*tab, True), ex)
    gr, Er = O.gradient(ref, O.Operator(*tab), oex, mode=0)
    assert abs(E - Er) < TOL[precision] * abs(Er)
    assert rel_err(g, gr) < TOL[precision]
    assert abs(psi.norm(ex) - O.psi_norm(ref, oex)) < TOL[precision] * O.psi_norm(ref, oex)


def test_deep_mc_chains_bit_identical_fp64():
    N, M, C = 8, [16, 8], [4, 16]
    biases, conns, weights, final = deep_layers(N, M, C, seed=4, scale=0.25)
    tab = random_operator(N, 10, 2, seed=4, hermitian_coeffs=True)
    psi = gpu_deep(N, biases, conns, weights, final, 64)
    chains, per = 32, 10
    mc = R.MonteCarloLoop(chains * per, 1, 3, chains, True, seed=17)
    E = R.ExpectationValue(True)(psi, R.Operator(*tab, True), mc)
    spins, lp, el = mc.last_samples()
    rs, rlp, rel, _ = O.records(O.Deep(N, biases, conns, weights, final), O.Operator(*tab),
                                O.MonteCarlo(N, chains * per, 1, 3, chains, seed=17))
    assert np.array_equal(spins, rs)
    assert rel_err(el, rel) < 1e-10 and rel_err(lp, rlp) < 1e-12
    assert abs(E - rel.mean()) < 1e-10 * abs(E)


def test_deep_hs_distance():
    N, M, C = 6, [12, 6], [3, 12]
    b1, c1, w1, f1 = deep_layers(N, M, C, seed=5, scale=0.2)
    b2, _, w2, f2 = deep_layers(N, M, C, seed=6, scale=0.2)
    ident = (np.ones(1, complex), np.zeros((1, 1), np.int32), -np.ones((1, 1), np.int32))
    for precision in (64, 32):
        psi, psi2 = gpu_deep(N, b1, c1, w1, f1, precision), gpu_deep(N, b2, c1, w2, f2, precision)
        hs = R.HilbertSpaceDistance(N, psi2.num_params, True)
        g, d = hs.gradient(psi, psi2, R.Operator(*ident, True), True, R.ExactSummation(N, True), 1.0)
        gr, dr = O.hs_distance(O.Deep(N, b1, c1, w1, f1), O.Deep(N, b2, c1, w2, f2), O.Operator(*ident), True,
                               O.Exact(N), want_grad=True, nu=1.0)
        assert abs(d - dr) < TOL[precision] * dr
        assert rel_err(g, gr) < TOL[precision] * 10


def test_deep_hs_distance_free_quantum_axis():
    """free_quantum_axis with PsiDeep on both sides: distance and gradient,
    alpha / beta entries included, against the oracle on the exact sum."""
    N, M, C = 6, [12, 6], [3, 12]
    b1, c1, w1, f1 = deep_layers(N, M, C, seed=5, scale=0.2)
    b2, _, w2, f2 = deep_layers(N, M, C, seed=6, scale=0.2)
    rng = np.random.default_rng(72)
    alpha, beta = rng.uniform(-0.4, 0.4, N), rng.uniform(-0.4, 0.4, N)
    alpha_p, beta_p = alpha + rng.uniform(-0.1, 0.1, N), beta + rng.uniform(-0.1, 0.1, N)
    ident = (np.ones(1, complex), np.zeros((1, 1), np.int32), -np.ones((1, 1), np.int32))
    cs = [np.asarray(c, np.uint32) for c in c1]
    for precision in (64, 32):
        print(f"[deep hs free axis] precision={precision}", flush=True)
        psi = R.PsiDeep(alpha, beta, list(b1), cs, list(w1), np.asarray(f1, complex), 1.0, True, True, precision=precision)
        psi2 = R.PsiDeep(alpha_p, beta_p, list(b2), cs, list(w2), np.asarray(f2, complex), 1.0, False, True,
                         precision=precision)
        hs = R.HilbertSpaceDistance(N, psi2.num_params, True)
        g, d = hs.gradient(psi, psi2, R.Operator(*ident, True), True, R.ExactSummation(N, True), 1.0)
        gr, dr = O.hs_distance(O.Deep(N, b1, c1, w1, f1), O.Deep(N, b2, c1, w2, f2), O.Operator(*ident), True,
                               O.Exact(N), want_grad=True, nu=1.0, axis=(alpha, beta, alpha_p, beta_p))
        assert abs(d - dr) < TOL[precision] * dr
        assert rel_err(g, gr) < TOL[precision] * 10
        assert rel_err(g[:2 * N], gr[:2 * N]) < TOL[precision] * 10
        if precision == 64:
            # Metropolis samples of psi: the oracle's own chains, bit for bit
            mc = R.MonteCarloLoop(32 * 20, 1, 4, 32, True, seed=23)
            gm, dm = hs.gradient(psi, psi2, R.Operator(*ident, True), True, mc, 1.0)
            omc = O.MonteCarlo(N, 32 * 20, 1, 4, 32, seed=23, call_index=0)
            grm, drm = O.hs_distance(O.Deep(N, b1, c1, w1, f1), O.Deep(N, b2, c1, w2, f2), O.Operator(*ident), True,
                                     omc, want_grad=True, nu=1.0, axis=(alpha, beta, alpha_p, beta_p))
            assert abs(dm - drm) < 1e-10 * drm
            assert rel_err(gm, grm) < 1e-9
