"""GPU checks of PsiDeep with precision=16: every layer as a BF16 MFMA GEMM with
FP32 accumulation and FP32 activations.

* The forward pass is compared with a float64 evaluation of the same rounded
  model (bfloat16 weights and hidden-layer inputs, float32 biases and final
  weights): median |d log psi| < 2e-5, max < 1e-3 (the max covers hidden
  activations that sit on a bfloat16 rounding boundary, where FP32 and FP64
  round to neighbouring values).
* E_loc is compared with the dense Hamiltonian applied to the kernel's own
  log psi over the whole basis (1e-5 relative): this isolates the local-energy
  logic (skipped zero terms, flips, coefficients) from the rounding.
* Sampling: visit frequencies and energy against the exact |psi|^2 of the same
  model.
* Against the unrounded FP64 model (the oracle) the error is stated and
  bounded loosely.
"""
import numpy as np
import pytest

from helpers import O, basis_words, bf16c, deep_bf16_log_psi, deep_layers, dense_hamiltonian, rel_err

import rbmongpu_amd as R
from rbmongpu_amd import hamiltonians as HAM

pytestmark = pytest.mark.gpu


def gpu_deep(N, biases, conns, weights, final, precision=16, compat=True):
    return R.PsiDeep(np.zeros(N), np.zeros(N), list(biases), [np.asarray(c, np.uint32) for c in conns], list(weights),
                     np.asarray(final, complex), 1.0, False, True, precision=precision, ref_compat_final=compat)


def layers(N, M, C, seed, re=0.8, im=0.2, bias=0.5):
    """Layers whose pre-activations are O(1) with small imaginary parts, away
    from the poles of my_logcosh near Im z = pi/2 where the output is
    ill-conditioned and any rounding of the inputs is amplified."""
    rng = np.random.default_rng(seed)
    from rbmongpu_amd.new_neural_network import deep_connections
    conns = deep_connections(N, M, C)
    biases = [bias * rng.normal(size=m) + 0.2j * bias * rng.normal(size=m) for m in M]
    weights = [(re * rng.normal(size=(c, m)) + 1j * im * rng.normal(size=(c, m))) / np.sqrt(c) for c, m in zip(C, M)]
    final = 1.0 + 0.5 * (rng.normal(size=M[-1]) + 1j * rng.normal(size=M[-1]))
    return biases, conns, weights, final


def random_words(N, n, seed):
    rng = np.random.default_rng(seed)
    w = np.zeros((n, 2), np.uint64)
    w[:, 0] = rng.integers(0, 2 ** min(N, 63), size=n, dtype=np.uint64)
    if N == 64:
        w[:, 0] |= rng.integers(0, 2, size=n, dtype=np.uint64) << np.uint64(63)
    return w


SHAPES = {
    # name: N, layer sizes, connections per unit
    "dense40-160-80": (40, [160, 80], [40, 160]),
    "sparse3": (12, [24, 12, 6], [6, 12, 4]),
    "odd": (7, [13, 5], [7, 13]),          # padding of every dimension
    "wide": (64, [256, 48], [64, 256]),
}


@pytest.mark.parametrize("shape", list(SHAPES))
def test_bf16_forward_matches_rounded_model(shape):
    print(f"[deep bf16 forward] {shape}", flush=True)
    N, M, C = SHAPES[shape]
    biases, conns, weights, final = layers(N, M, C, seed=N + len(M))
    psi = gpu_deep(N, biases, conns, weights, final)
    words = basis_words(N) if N <= 12 else random_words(N, 4096, seed=N)
    lp = R.log_psi_batch(psi, words)
    ref = deep_bf16_log_psi(N, biases, conns, weights, final, words)
    scale = 1.0 + np.abs(ref).max()
    d = np.abs(lp - ref)
    print(f"  kernel vs rounded model: max {d.max():.3e} median {np.median(d):.3e}", flush=True)
    assert np.median(d) < 1e-5 * scale and d.max() < 1e-3 * scale, (np.median(d), d.max(), scale)
    # against the unrounded FP64 model
    exact = O.log_psi(O.Deep(N, biases, conns, weights, final), words)
    print(f"  bf16 vs fp64: max {np.abs(lp - exact).max():.3e} median {np.median(np.abs(lp - exact)):.3e} "
          f"(max |log psi| {np.abs(exact).max():.2f})", flush=True)
    assert np.abs(lp - exact).max() < 0.02 * scale


def test_bf16_local_energy_against_dense_hamiltonian():
    N, M, C = 10, [20, 10], [10, 20]
    biases, conns, weights, final = layers(N, M, C, seed=3)
    tab = HAM.j1j2_chain(N)
    psi = gpu_deep(N, biases, conns, weights, final)
    words = basis_words(N)
    lp = R.log_psi_batch(psi, words)
    H = dense_hamiltonian(N, *tab)
    psi_v = np.exp(lp - lp.real.max())
    ref = (H @ psi_v) / psi_v
    el = R.local_energy_batch(psi, R.Operator(*tab, True), words)
    assert rel_err(el, ref) < 1e-5


def test_bf16_exact_energy_and_gradient():
    N, M, C = 10, [20, 10], [10, 20]
    biases, conns, weights, final = layers(N, M, C, seed=4)
    tab = HAM.j1j2_chain(N)
    psi = gpu_deep(N, biases, conns, weights, final, compat=False)
    ex = R.ExactSummation(N, True)
    psi.normalize(ex)
    g, E = R.ExpectationValue(True).gradient(psi, R.Operator(*tab, True), ex)
    # the FP64 oracle on the bfloat16-rounded weights: the remaining difference
    # is the rounding of the hidden activations
    rw = [bf16c(w) for w in weights]
    ref = O.Deep(N, biases, conns, rw, final, prefactor=psi.prefactor, ref_compat_final=False)
    gr, Er = O.gradient(ref, O.Operator(*tab), O.Exact(N), mode=0)
    assert abs(E - Er) < 2e-3 * abs(Er), (E, Er)
    assert rel_err(g, gr) < 2e-2


@pytest.mark.parametrize("shape,compat,estimator", [
    ("two", False, "auto"), ("two", True, "auto"), ("three", False, "E_k"), ("odd", True, "hermitian")])
def test_bf16_gradient_from_mfma_backward(shape, compat, estimator):
    """The gradient contraction from MFMA back-propagation records and per-layer
    GEMMs over the samples against the FP64 oracle on the bfloat16-rounded
    weights (the remaining difference is the rounding of activations and of
    delta in the backward products)."""
    print(f"[bf16 gradient] {shape} compat={compat} {estimator}", flush=True)
    N, M, C = {"two": (10, [20, 10], [10, 20]), "three": (9, [18, 12, 6], [9, 6, 12]),
               "odd": (7, [13, 5], [7, 13])}[shape]
    biases, conns, weights, final = layers(N, M, C, seed=len(M) + N)
    tab = HAM.j1j2_chain(N)
    psi = gpu_deep(N, biases, conns, weights, final, compat=compat)
    ex = R.ExactSummation(N, True)
    psi.normalize(ex)
    g, E = R.ExpectationValue(True).gradient(psi, R.Operator(*tab, True), ex, estimator=estimator)
    rw = [bf16c(w) for w in weights]
    ref = O.Deep(N, biases, conns, rw, final, prefactor=psi.prefactor, ref_compat_final=compat)
    gr, Er = O.gradient(ref, O.Operator(*tab), O.Exact(N), mode=0)
    print(f"  rel err {rel_err(g, gr):.2e}", flush=True)
    assert abs(E - Er) < 2e-3 * abs(Er)
    assert rel_err(g, gr) < 2e-2
    assert np.all(g[:2 * N] == 0)


def test_bf16_sampler_visits_psi_squared():
    N, M, C = 8, [16, 8], [8, 16]
    biases, conns, weights, final = layers(N, M, C, seed=5, re=0.6, bias=0.3)
    psi = gpu_deep(N, biases, conns, weights, final)
    words = basis_words(N)
    lp = R.log_psi_batch(psi, words)
    p = np.exp(2 * (lp.real - lp.real.max()))
    p /= p.sum()
    chains, per = 1024, 200
    mc = R.MonteCarloLoop(chains * per, 1, 20, chains, True, seed=9)
    tab = HAM.heisenberg([(i, (i + 1) % N) for i in range(N)])
    R.ExpectationValue(True)(psi, R.Operator(*tab, True), mc)
    spins, slp, _ = mc.last_samples()
    counts = np.bincount(spins[:, 0].astype(np.int64), minlength=2 ** N)
    n = counts.sum()
    f = counts / n
    # Sokal-free bound: chains are independent, samples one sweep apart
    se = np.sqrt(np.maximum(p * (1 - p), 1.0 / n) / n) * 3.0
    big = p > 1e-3
    z = np.abs(f - p)[big] / se[big]
    assert np.mean(z < 4.5) > 0.97, np.sort(z)[-5:]
    assert rel_err(slp, lp[spins[:, 0].astype(np.int64)]) < 1e-6
    assert 0.05 < mc.acceptance_rate < 1.0


def test_bf16_mc_energy_within_three_standard_errors():
    N, M, C = 12, [24, 12], [12, 24]
    biases, conns, weights, final = layers(N, M, C, seed=6, re=0.5, bias=0.3)
    tab = HAM.j1j2_chain(N)
    psi = gpu_deep(N, biases, conns, weights, final)
    ev = R.ExpectationValue(True)
    ex = R.ExactSummation(N, True)
    psi.normalize(ex)
    E_exact = ev(psi, R.Operator(*tab, True), ex).real
    chains, per = 512, 100
    mc = R.MonteCarloLoop(chains * per, 1, 30, chains, True, seed=11)
    E_mc = ev(psi, R.Operator(*tab, True), mc).real
    _, _, el = mc.last_samples()
    per_chain = el.real.reshape(per, chains).mean(axis=0)
    se = per_chain.std(ddof=1) / np.sqrt(chains)
    assert abs(E_mc - E_exact) < 3 * se, (E_mc, E_exact, se)


def test_bf16_deep_tier_shape_sampler_agrees_with_fp32():
    """The 40 -> 160 -> 80 network runs a compile-time-shaped, software-pipelined
    forward pass: its chain's energy agrees with the complex64 kernels' chain
    on the same model within 3 combined standard errors (the models differ by
    the bfloat16 rounding, well below the statistical error here)."""
    N, M, C = 40, [160, 80], [40, 160]
    biases, conns, weights, final = layers(N, M, C, seed=44, re=0.25, im=0.1, bias=0.2)
    tab = HAM.j1j2_chain(N)
    ev = R.ExpectationValue(True)
    res = []
    for prec in (16, 32):
        psi = gpu_deep(N, biases, conns, weights, final, precision=prec)
        chains, per = 2048, 20
        mc = R.MonteCarloLoop(chains * per, 2, 50, chains, True, seed=5 + prec)
        E = ev(psi, R.Operator(*tab, True), mc).real
        _, _, el = mc.last_samples()
        per_chain = el.real.reshape(per, chains).mean(axis=0)
        res.append((E, per_chain.std(ddof=1) / np.sqrt(chains), mc.acceptance_rate))
    (e16, s16, a16), (e32, s32, a32) = res
    print(f"  bf16 {e16:.5f} +- {s16:.5f} (acc {a16:.3f}), fp32 {e32:.5f} +- {s32:.5f} (acc {a32:.3f}): "
          f"{abs(e16 - e32) / np.hypot(s16, s32):.2f} combined standard errors apart", flush=True)
    assert abs(e16 - e32) < 3 * np.hypot(s16, s32)
    assert abs(a16 - a32) < 0.02


def test_bf16_rejects_unsupported_models():
    N, M, C = 8, [256, 256], [8, 256]           # 2 x 256 x 256 bfloat16 weights exceed LDS
    biases, conns, weights, final = deep_layers(N, M, C, seed=1, scale=0.01)
    psi = gpu_deep(N, biases, conns, weights, final)
    with pytest.raises(ValueError):
        R.log_psi_batch(psi, basis_words(N))
    N, M, C = 6, [12, 6], [6, 12]
    biases, conns, weights, final = deep_layers(N, M, C, seed=1, scale=0.1)
    psi = gpu_deep(N, biases, conns, weights, final)
    psi.translational_invariance = True
    with pytest.raises(NotImplementedError):
        R.log_psi_batch(psi, basis_words(N))


def test_bf16_hs_distance_and_learned_ensemble():
    """The BF16 PsiDeep as psi' of the Hilbert-space distance (its gradient goes
    through the MFMA back-propagation) against the FP64 oracle on the rounded
    weights, and as the wavefunction of a learned-proposal ensemble against
    ExactSummation."""
    N, M, C = 8, [16, 8], [8, 16]
    b1, c1, w1, f1 = layers(N, M, C, seed=31)
    b2, _, w2, f2 = layers(N, M, C, seed=32)
    ident = (np.ones(1, complex), np.zeros((1, 1), np.int32), -np.ones((1, 1), np.int32))
    psi = gpu_deep(N, b1, c1, w1, f1, precision=16, compat=False)
    psi2 = gpu_deep(N, b2, c1, w2, f2, precision=16, compat=False)
    hs = R.HilbertSpaceDistance(N, psi2.num_params, True)
    g, d = hs.gradient(psi, psi2, R.Operator(*ident, True), True, R.ExactSummation(N, True), 1.0)
    r1 = O.Deep(N, b1, c1, [bf16c(w) for w in w1], f1, ref_compat_final=False)
    r2 = O.Deep(N, b2, c1, [bf16c(w) for w in w2], f2, ref_compat_final=False)
    gr, dr = O.hs_distance(r1, r2, O.Operator(*ident), True, O.Exact(N), want_grad=True, nu=1.0)
    assert abs(d - dr) < 2e-3 * dr
    assert rel_err(g, gr) < 3e-2
    # free quantum axis: psi' on the single-flip neighbours through the MFMA forward pass
    rng = np.random.default_rng(73)
    alpha, beta = rng.uniform(-0.4, 0.4, N), rng.uniform(-0.4, 0.4, N)
    alpha_p, beta_p = alpha + rng.uniform(-0.1, 0.1, N), beta + rng.uniform(-0.1, 0.1, N)
    cs = [np.asarray(c, np.uint32) for c in c1]
    psia = R.PsiDeep(alpha, beta, list(b1), cs, list(w1), np.asarray(f1, complex), 1.0, True, True, precision=16,
                     ref_compat_final=False)
    psia2 = R.PsiDeep(alpha_p, beta_p, list(b2), cs, list(w2), np.asarray(f2, complex), 1.0, False, True, precision=16,
                      ref_compat_final=False)
    g, d = hs.gradient(psia, psia2, R.Operator(*ident, True), True, R.ExactSummation(N, True), 1.0)
    gr, dr = O.hs_distance(r1, r2, O.Operator(*ident), True, O.Exact(N), want_grad=True, nu=1.0,
                           axis=(alpha, beta, alpha_p, beta_p))
    assert abs(d - dr) < 2e-3 * dr
    assert rel_err(g, gr) < 3e-2
    # learned-proposal ensemble on the BF16 model
    tab = HAM.j1j2_chain(N)
    ex = R.ExactSummation(N, True)
    psi.normalize(ex)
    ev = R.ExpectationValue(True)
    E_ex = ev(psi, R.Operator(*tab, True), ex).real
    # an untrained proposal network: independent-MH acceptance is low, so the
    # chains thermalise over 300 proposals (the per-chain standard error
    # covers the autocorrelation of the kept samples, not a start-up bias)
    chains, per = 512, 60
    md = R.MADE(N, 32, seed=2)
    ens = R.LearnedMonteCarloLoop(md, chains * per, 300, chains, True, seed=7)
    E = ev(psi, R.Operator(*tab, True), ens).real
    _, _, el = ens.last_samples()
    per_chain = el.real.reshape(per, chains).mean(axis=0)
    se = per_chain.std(ddof=1) / np.sqrt(chains)
    print(f"  learned-proposal chain on the bfloat16 model: E {E:.6f} vs exact {E_ex:.6f}, "
          f"{abs(E - E_ex) / se:.2f} standard errors, acceptance {ens.acceptance_rate:.3f}", flush=True)
    assert abs(E - E_ex) < 3 * se, (E, E_ex, se)
