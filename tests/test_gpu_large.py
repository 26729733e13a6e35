"""GPU parity at the benchmark shapes against the FP64 oracle on the same
samples.

* 6x6 Heisenberg antiferromagnet, RBM M = 144 (216 Pauli strings, 72 bonds):
  E_loc, the complex64 gradient, a block of the centred S-matrix and the SR
  update on more samples than parameters (ExpectationValue.cu:177-256,
  S_matrix.cu:10-53).
* 10x10 TFIM at h = 3.04, RBM M = 200: E_loc (Operator.hpp:90-115).

Branch disagreements.  The reference's my_logcosh is sign-branched on
Re theta > 0 (psi_functions.hpp:49-57) and jumps across Re theta = 0 by
2i Im L(i Im theta) (0.026i at |Im theta| = 0.7).  Where the complex64 angle
and the FP64 angle of a configuration (or of one of its flipped neighbours)
fall on different sides of Re theta = 0, the two evaluations legitimately
differ by that jump.  The tests find those configurations exactly by
replaying the kernel's FP32 angle arithmetic (theta = b, then theta += s_i W_i
for i = 0..N-1, theta' = theta + (-2 s_p) W_p per flip in ascending p) in
numpy float32, report how many there are, bound their error by the jump, and
hold every other configuration to the 1e-5 complex64 bar."""
import numpy as np
import pytest

from helpers import O, check_local_energies, rel_err, spins_pm

import rbmongpu_amd as R
from rbmongpu_amd import hamiltonians as HAM

pytestmark = pytest.mark.gpu
BAR32 = 1e-5


def seeded_rbm(N, M, seed, scale):
    rng = np.random.default_rng(seed)
    b = scale * (rng.normal(size=M) + 1j * rng.normal(size=M))
    W = scale * (rng.normal(size=(N, M)) + 1j * rng.normal(size=(N, M)))
    return b, W


# ------------------------------------------------------------------ 6x6 Heisenberg, M = 144
@pytest.fixture(scope="module")
def heisenberg6x6():
    N, M = 36, 144
    b, W = seeded_rbm(N, M, 1235, 0.01)
    tab = HAM.heisenberg_2d(6, 6)
    psi = R.Psi(np.zeros(N), np.zeros(N), b, W, 1.0, False, True, precision=32)
    op = R.Operator(*tab, True)
    chains, per = 2048, 4
    mc = R.MonteCarloLoop(chains * per, 1, 100, chains, True, seed=6)
    mc_g, mc_s = R.MonteCarloLoop(mc), R.MonteCarloLoop(mc)
    g, E = R.ExpectationValue(True).gradient(psi, op, mc_g)
    spins, lp, el = mc_g.last_samples()
    s = spins_pm(spins, N)
    El = O.local_energy(O.RBM(b, W), O.Operator(*tab), spins)
    lpr = O.log_psi(O.RBM(b, W), spins)
    T = O.tanh(b[None, :] + s @ W)
    Ok = np.concatenate([T, (s[:, :, None] * T[:, None, :]).reshape(len(s), N * M)], axis=1)
    return dict(N=N, M=M, b=b, W=W, tab=tab, psi=psi, op=op, mc=mc, mc_s=mc_s, g=g, E=E, spins=spins, s=s,
                lp=lp, el=el, El=El, lpr=lpr, Ok=Ok)


def test_heisenberg6x6_local_energies(heisenberg6x6):
    d = heisenberg6x6
    assert d["op"].num_strings == 216
    check_local_energies("6x6 Heisenberg M=144", d["el"], d["El"], d["lp"], d["lpr"], d["b"], d["W"], d["s"], d["tab"])


def test_heisenberg6x6_gradient(heisenberg6x6):
    """The complex64 gradient of the recorded samples against the FP64
    estimator 2/n sum_s (E_s - <E>) conj(O_k(s)) with the oracle's E_loc and
    O_k on the same samples, and the same estimator with the kernel's E_loc
    (which separates the contraction's error from E_loc's)."""
    d = heisenberg6x6
    N, Ok, El, el = d["N"], d["Ok"], d["El"], d["el"]
    n = len(El)
    ref = (2.0 / n) * ((El - El.mean())[:, None] * Ok.conj()).sum(axis=0)
    with_gpu_eloc = (2.0 / n) * ((el - el.mean())[:, None] * Ok.conj()).sum(axis=0)
    e_all = rel_err(d["g"][2 * N:], ref)
    e_contraction = rel_err(d["g"][2 * N:], with_gpu_eloc)
    e_eloc = rel_err(with_gpu_eloc, ref)
    print(f"[6x6 Heisenberg gradient] vs FP64: {e_all:.2e} (E_loc part {e_eloc:.2e}, contraction part "
          f"{e_contraction:.2e}); max|g| {np.abs(ref).max():.3e}", flush=True)
    assert abs(d["E"] - El.mean()) < BAR32 * abs(El.mean())
    assert e_contraction < BAR32
    assert e_all < BAR32


def test_heisenberg6x6_centred_s_block_and_sr_update(heisenberg6x6):
    """Centred S on FP32 MFMA with FP64 accumulation: a 512 x 512 block
    against numpy.cov of the oracle's O_k (uniform weights) and the whole
    matrix against the FP64 product, both at 1e-5.  Then the SR update
    x = (S + eps I)^{-1} g on 8192 samples (more than the P = 5328 parameters
    with an O_k) against numpy.linalg.solve of the FP64 system: the FP64
    Cholesky solve adds no error of its own, so x_gpu - x_ref equals the
    first-order propagation of the inputs' errors,
    A^{-1} ((g_gpu - g_ref) - (S_gpu - S_ref) x_ref), to within 5 %; the
    propagation is amplified by up to |S| / eps (about 70 here), which is why
    x itself carries ~3e-4 relative error at eps = 1e-4."""
    d = heisenberg6x6
    N, Ok = d["N"], d["Ok"]
    P = Ok.shape[1]
    assert len(Ok) > P
    S = R.get_S_matrix_centred(d["psi"], d["mc_s"])
    spins_s, _, _ = d["mc_s"].last_samples()
    assert np.array_equal(spins_s, d["spins"])
    rows = np.r_[0:256, d["M"]:d["M"] + 256]          # b entries and the first W entries
    cols = np.r_[0:128, P - 384:P]
    ref_block = np.cov(Ok[:, rows].T, Ok[:, cols].T, bias=True)[:len(rows), len(rows):].conj()
    e_block = rel_err(S[np.ix_(rows, cols)], ref_block)
    # the full FP64 S for the solve
    Oc = Ok - Ok.mean(axis=0)
    S_ref = (Oc.conj().T @ Oc) / len(Ok)
    e_s = rel_err(S, S_ref)
    eps = 1e-4
    mc_u = R.MonteCarloLoop(d["mc"])
    upd, E = R.sr_update(d["psi"], d["op"], mc_u, eps)
    spins_u, _, _ = mc_u.last_samples()
    assert np.array_equal(spins_u, d["spins"])
    El = d["El"]
    g_ref = (2.0 / len(El)) * ((El - El.mean())[:, None] * Ok.conj()).sum(axis=0)
    A = S_ref + eps * np.eye(P)
    x_ref = np.linalg.solve(A, g_ref)
    x = upd[2 * N:]
    predicted = np.linalg.solve(A, (d["g"][2 * N:] - g_ref) - (S - S_ref) @ x_ref)
    mismatch = np.linalg.norm((x - x_ref) - predicted) / np.linalg.norm(predicted)
    e_upd = rel_err(x, x_ref)
    print(f"[6x6 Heisenberg SR] S block rel err {e_block:.2e}, full S {e_s:.2e}, update {e_upd:.2e} "
          f"(first-order propagation of the S and g errors explains it to {mismatch:.1%}; {len(Ok)} samples, "
          f"P = {P}, |S| / eps = {np.abs(S_ref).max() / eps:.0f})", flush=True)
    assert e_block < BAR32
    assert e_s < BAR32
    assert mismatch < 0.05


# ------------------------------------------------------------------ 10x10 TFIM, M = 200
@pytest.mark.parametrize("scale", [0.01, 0.05])
def test_tfim10x10_local_energies(scale):
    """E_loc of sampled 10x10 TFIM configurations (N = 100, two spin words,
    M = 200, 300 strings) at the tier's N(0, 0.01^2) initialisation and at a
    5x larger scale (|theta| ~ 0.7, where the my_logcosh jump is 0.026)."""
    N, M = 100, 200
    b, W = seeded_rbm(N, M, 1238, scale)
    tab = HAM.tfim_2d(10, 10, 1.0, 3.04)
    psi = R.Psi(np.zeros(N), np.zeros(N), b, W, 1.0, False, True, precision=32)
    op = R.Operator(*tab, True)
    chains, per = 1024, 4
    mc = R.MonteCarloLoop(chains * per, 1, 50, chains, True, seed=13)
    E = R.ExpectationValue(True)(psi, op, mc)
    spins, lp, el = mc.last_samples()
    El = O.local_energy(O.RBM(b, W), O.Operator(*tab), spins)
    lpr = O.log_psi(O.RBM(b, W), spins)
    good = check_local_energies(f"10x10 TFIM M=200 scale {scale}", el, El, lp, lpr, b, W, spins_pm(spins, N), tab)
    assert abs(E - El.mean()) < BAR32 * abs(El.mean()) + np.abs(el[~good] - El[~good]).sum() / len(El)


# ------------------------------------------------------------------ J1-J2 N = 40, PsiDeep 40 -> 160 -> 80
def j1j2_deep_tier_model(scale=1.0, seed=1237):
    """The deep-network tier's model: dense 40 -> 160 -> 80 connections, complex
    weights N(0, (0.05 scale)^2 / fan_in) in real and imaginary part, zero
    biases, final weights 1."""
    from rbmongpu_amd.new_neural_network import deep_connections
    N, M, C = 40, [160, 80], [40, 160]
    rng = np.random.default_rng(seed)
    conns = [np.asarray(c, np.uint32) for c in deep_connections(N, M, C)]
    weights = [(0.05 * scale / np.sqrt(c)) * (rng.normal(size=(c, m)) + 1j * rng.normal(size=(c, m)))
               for c, m in zip(C, M)]
    biases = [np.zeros(m, complex) for m in M]
    final = np.ones(M[-1], complex)
    return N, biases, conns, weights, final


def random_words40(n, seed):
    rng = np.random.default_rng(seed)
    w = np.zeros((n, 2), np.uint64)
    w[:, 0] = rng.integers(0, 2 ** 40, size=n, dtype=np.uint64)
    return w


@pytest.mark.parametrize("precision", [64, 32])
def test_j1j2_deep_tier_shape_against_oracle(precision):
    """log psi, E_loc and O_k of the generic complex64 / complex128 PsiDeep
    kernels at the deep-network tier's shape (N = 40, 40 -> 160 -> 80, 240
    Pauli strings) on 512 random configurations, against the FP64 oracle
    (PsiDeep.hpp:89-169, 214-293; Operator.hpp:90-115).  At the specified
    initialisation log psi is ~1e-5, so its error is held in absolute terms
    (the relative error of psi): 1e-5 in complex64, 1e-12 in complex128."""
    N, biases, conns, weights, final = j1j2_deep_tier_model()
    tab = HAM.j1j2_chain(N)
    assert len(tab[0]) == 240
    psi = R.PsiDeep(np.zeros(N), np.zeros(N), biases, conns, weights, final, 1.0, False, True, precision=precision)
    ref = O.Deep(N, biases, conns, weights, final)
    words = random_words40(512, seed=40)
    lp = R.log_psi_batch(psi, words)
    lpr = O.log_psi(ref, words)
    el = R.local_energy_batch(psi, R.Operator(*tab, True), words)
    elr = O.local_energy(ref, O.Operator(*tab), words)
    e_lp, e_el = np.abs(lp - lpr).max(), rel_err(el, elr)
    oks = [(R.psi_O_k_vector(psi, R.Spins(int(w), N)), O.o_k(ref, np.array([w, 0], np.uint64))[2 * N:])
           for w in words[:8, 0]]
    e_ok = max(rel_err(a, b) for a, b in oks)
    print(f"[J1-J2 deep tier, precision {precision}] |d log psi| max {e_lp:.2e} (max |log psi| "
          f"{np.abs(lpr).max():.2e}); E_loc rel {e_el:.2e}; O_k rel {e_ok:.2e}", flush=True)
    bar = 1e-10 if precision == 64 else 1e-5
    assert e_lp < (1e-12 if precision == 64 else 1e-5)
    assert e_el < bar
    assert e_ok < bar


def test_j1j2_deep_tier_shape_fp64_chains_and_gradient():
    """complex128 Metropolis chains of the deep-network tier's model (weights
    x 12: acceptance ~0.5) replay the oracle's chains bit for bit, and the
    Monte Carlo gradient matches the oracle's estimator on them to 1e-10 (the
    centred estimator 2<E conj O> - 2<E> conj<O> that both use for a
    Hermitian operator)."""
    N, biases, conns, weights, final = j1j2_deep_tier_model(scale=12.0)
    tab = HAM.j1j2_chain(N)
    psi = R.PsiDeep(np.zeros(N), np.zeros(N), biases, conns, weights, final, 1.0, False, True, precision=64)
    ref = O.Deep(N, biases, conns, weights, final)
    chains, per, therm = 32, 4, 5
    mc = R.MonteCarloLoop(chains * per, 1, therm, chains, True, seed=19)
    mc_g = R.MonteCarloLoop(mc)
    R.ExpectationValue(True)(psi, R.Operator(*tab, True), mc)
    spins, lp, el = mc.last_samples()
    rs, rlp, rel, _ = O.records(ref, O.Operator(*tab), O.MonteCarlo(N, chains * per, 1, therm, chains, seed=19))
    assert np.array_equal(spins, rs)
    assert rel_err(el, rel) < 1e-10
    print(f"[J1-J2 deep tier chains] acceptance {mc.acceptance_rate:.3f}", flush=True)
    assert 0.05 < mc.acceptance_rate < 0.95
    g, E = R.ExpectationValue(True).gradient(psi, R.Operator(*tab, True), mc_g)
    gr, Er = O.gradient(ref, O.Operator(*tab), O.MonteCarlo(N, chains * per, 1, therm, chains, seed=19), mode=1)
    assert abs(E - Er) < 1e-10 * abs(Er)
    assert rel_err(g, gr) < 1e-10
