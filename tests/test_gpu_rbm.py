"""GPU parity of the RBM path against the FP64 oracle.

Tolerances: complex128 kernels (precision=64) agree with the oracle to 1e-10
relative; complex64 kernels (precision=32, FP64 reductions) to 1e-5 relative,
measured as max|gpu - ref| / max|ref| over each output.  Basis enumeration,
spin indexing and, in FP64, whole Markov chains are bit-exact."""
import os

import numpy as np
import pytest

from helpers import O, basis_words, check_local_energies, dense_hamiltonian, random_operator, random_rbm, rel_err, spins_pm

import rbmongpu_amd as R
from rbmongpu_amd import hamiltonians as HAM

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TOL = {64: 1e-10, 32: 1e-5}
PRECISIONS = [64, 32]


def gpu_psi(b, W, a=None, prefactor=1.0, precision=64):
    N = W.shape[0]
    return R.Psi(np.zeros(N), np.zeros(N), b, W, prefactor, False, True, a=a, precision=precision)


def gpu_op(tab):
    return R.Operator(tab[0], tab[1], tab[2], True)


def words_to_int(w):
    return [int(x[0]) | (int(x[1]) << 64) for x in w]


def fluctuation_bar(tol, f, E):
    """The fluctuation is sqrt(<|E_loc|^2> - |<E_loc>|^2) (ExpectationValue.cu:
    130-175).  With <|E_loc|^2> = f^2 + |E|^2 held to tol and <E_loc> to tol
    (so |<E_loc>|^2 to 2 tol), f^2 carries at most tol (f^2 + 3 |E|^2) and f
    half of that relative to f^2: the bar is tol (f^2 + 3 |E|^2) / (2 f),
    about tol f / 2 where the variance does not cancel and growing as
    |E|^2 / f^2 where it does."""
    return tol * (f * f + 3 * abs(E) ** 2) / (2 * f)


# ------------------------------------------------------------------ golden
@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("pre", ["", "vis_"])
def test_c1_golden(precision, pre):
    print(f"[c1 golden] precision={precision} visible={bool(pre)}", flush=True)
    d = np.load(os.path.join(GOLDEN, "c1_tfim_rbm.npz"))
    a = d["vis_a"] if pre else None
    tol = TOL[precision]
    psi = gpu_psi(d[pre + "b"], d[pre + "W"], a, precision=precision)
    op = gpu_op(HAM.tfim_1d(10))
    ex = R.ExactSummation(10, True)
    assert rel_err(psi.vector, d[pre + "psi"]) < tol
    el = R.local_energy_batch(psi, op, basis_words(10))
    assert rel_err(el, d[pre + "e_loc"]) < tol                 # per state, all 1024 of them
    assert np.abs(el - d[pre + "e_loc"]).max() < tol * np.abs(d[pre + "e_loc"]).max()
    psi.normalize(ex)
    assert abs(psi.prefactor - d[pre + "prefactor"]) < tol * d[pre + "prefactor"]
    psi.prefactor = float(d[pre + "prefactor"])
    ev = R.ExpectationValue(True)
    g, E = ev.gradient(psi, op, ex)
    assert abs(E - d[pre + "energy"]) < tol * abs(d[pre + "energy"])
    assert rel_err(g, d[pre + "gradient"]) < tol
    f, E2 = ev.fluctuation(psi, op, ex)
    fr = float(d[pre + "fluctuation"])
    assert abs(f - fr) < fluctuation_bar(tol, fr, d[pre + "energy"]), (f, fr)
    psi2 = gpu_psi(d[pre + "b"], d[pre + "W2"], a, precision=precision)
    hs = R.HilbertSpaceDistance(10, psi2.num_params, True)
    hg, hd = hs.gradient(psi, psi2, op, True, ex, 1.0)
    assert abs(hd - d[pre + "hs_distance"]) < tol * d[pre + "hs_distance"]
    assert rel_err(hg, d[pre + "hs_gradient"]) < tol
    assert abs(hs(psi, psi2, op, True, ex) - hd) < 1e-12


# ------------------------------------------------------------------ shapes
@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("N,M", [(1, 1), (5, 16), (7, 17), (20, 80), (64, 40), (65, 33), (100, 200), (128, 256), (3, 250)])
def test_log_psi_shapes(precision, N, M):
    print(f"[log_psi shapes] precision={precision} N={N} M={M}", flush=True)
    b, W, a = random_rbm(N, M, seed=N * 1000 + M, scale=0.05, visible=True)
    rng = np.random.default_rng(N + M)
    configs = [int.from_bytes(rng.bytes(16), "little") & ((1 << N) - 1) for _ in range(300)]
    words = O.spins_words(configs)
    ref = O.log_psi(O.RBM(b, W, a), words)
    psi = gpu_psi(b, W, a, precision=precision)
    got = R.log_psi_batch(psi, words)
    assert rel_err(got, ref) < TOL[precision]


@pytest.mark.parametrize("precision", PRECISIONS)
def test_local_energy_ratios_beyond_fp32_range(precision):
    """psi(s')/psi(s) up to e^400: complex64 kernels still return finite E_loc
    (the ratio's exponent is applied in FP64)."""
    N, M = 6, 3
    rng = np.random.default_rng(5)
    W = 40.0 * rng.normal(size=(N, M)) + 1j * rng.normal(size=(N, M))
    b = rng.normal(size=M) + 1j * rng.normal(size=M)
    tab = HAM.tfim_1d(N, 1.0, 0.5)
    ref = O.local_energy(O.RBM(b, W), O.Operator(*tab), basis_words(N))
    assert np.abs(ref).max() > np.exp(90.0)
    got = R.local_energy_batch(gpu_psi(b, W, precision=precision), gpu_op(tab), basis_words(N))
    assert np.all(np.isfinite(got))
    assert np.max(np.abs(got - ref) / np.abs(ref)) < TOL[precision] * 10


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("N,M", [(31, 20), (33, 17), (64, 40), (65, 33), (100, 48), (128, 16)])
def test_local_energy_wide(precision, N, M):
    """Flip sites and Z/Y factors in every 32-bit quarter of the two spin words."""
    print(f"[local energy wide] precision={precision} N={N} M={M}", flush=True)
    b, W, a = random_rbm(N, M, seed=N + 7 * M, scale=0.05, visible=True)
    coeffs, types, idx = random_operator(N, 40, 5, seed=N)
    tfim = HAM.tfim_1d(N, 1.0, 0.7)
    tab = (np.concatenate([coeffs, tfim[0]]),
           np.concatenate([types, np.pad(tfim[1], ((0, 0), (0, types.shape[1] - tfim[1].shape[1])))]),
           np.concatenate([idx, np.pad(tfim[2], ((0, 0), (0, idx.shape[1] - tfim[2].shape[1])), constant_values=-1)]))
    rng = np.random.default_rng(N)
    configs = [int.from_bytes(rng.bytes(16), "little") & ((1 << N) - 1) for _ in range(200)]
    words = O.spins_words(configs)
    ref = O.local_energy(O.RBM(b, W, a), O.Operator(*tab), words)
    got = R.local_energy_batch(gpu_psi(b, W, a, precision=precision), gpu_op(tab), words)
    assert rel_err(got, ref) < TOL[precision]


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("kind", ["tfim", "heis", "random", "repeated", "continuation", "visible"])
def test_local_energy_operators(precision, kind):
    print(f"[local energy] precision={precision} kind={kind}", flush=True)
    N, M = 8, 21
    b, W, a = random_rbm(N, M, seed=3, scale=0.1, visible=kind == "visible")
    if kind == "tfim":
        tab = HAM.tfim_1d(N, 1.0, 0.9)
    elif kind in ("heis", "visible"):
        tab = HAM.heisenberg([(i, (i + 1) % N) for i in range(N)])
    elif kind == "random":
        tab = random_operator(N, 20, 4, seed=8)
    elif kind == "repeated":
        # repeated indices follow the reference's literal rule (factors use the original spins)
        tab = HAM.table([(1.0, [(1, 0), (1, 0)]), (0.5, [(2, 1), (2, 1)]), (0.3j, [(1, 2), (3, 2)]),
                         (0.7, [(3, 3), (2, 3), (1, 4)])])
    else:
        # a row without -1 continues into the next row (Operator.hpp:148-152)
        tab = (np.array([1.0, 0.5 + 0.1j, 0.2]), np.array([[1, 3], [1, 2], [3, 0]], np.int32),
               np.array([[0, 1], [2, 5], [4, -1]], np.int32))
    psi = gpu_psi(b, W, a, precision=precision)
    ref = O.local_energy(O.RBM(b, W, a), O.Operator(*tab), basis_words(N))
    got = R.local_energy_batch(psi, gpu_op(tab), basis_words(N))
    assert rel_err(got, ref) < TOL[precision]


def test_empty_operator():
    N, M = 4, 5
    b, W, _ = random_rbm(N, M, seed=1)
    psi = gpu_psi(b, W)
    op = R.Operator(np.zeros(0, complex), np.zeros((0, 1), np.int32), np.zeros((0, 1), np.int32), True)
    assert np.all(R.local_energy_batch(psi, op, basis_words(N)) == 0)
    assert R.ExpectationValue(True)(psi, op, R.ExactSummation(N, True)) == 0


# ------------------------------------------------------------------ exact summation
@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("sector", [None, 0, 2])
def test_exact_functions(precision, sector):
    print(f"[exact functions] precision={precision} sector={sector}", flush=True)
    N, M = 8, 13
    b, W, a = random_rbm(N, M, seed=17, scale=0.2, visible=True)
    tab = HAM.heisenberg([(i, (i + 1) % N) for i in range(N)])
    tab2 = HAM.tfim_1d(N, 0.5, 1.5)
    ref_psi = O.RBM(b, W, a, prefactor=0.3)
    oex = O.Exact(N, sector)
    psi = gpu_psi(b, W, a, prefactor=0.3, precision=precision)
    ex = R.ExactSummation(N, True)
    if sector is not None:
        ex.set_total_z_symmetry(sector)
    assert ex.num_steps == oex.num_configs
    tol = TOL[precision]
    ev = R.ExpectationValue(True)
    E = ev(psi, gpu_op(tab), ex)
    assert abs(E - O.expectation(ref_psi, O.Operator(*tab), oex)) < tol * abs(E)
    El = ev(psi, [gpu_op(tab), gpu_op(tab2)], ex)
    assert abs(El[1] - O.expectation(ref_psi, O.Operator(*tab2), oex)) < tol * abs(El[1])
    assert abs(El[0] - E) < 1e-12 * abs(E)
    f, E2 = ev.fluctuation(psi, gpu_op(tab), ex)
    fr, Er = O.fluctuation(ref_psi, O.Operator(*tab), oex)
    if np.isfinite(fr):
        assert abs(f - fr) < fluctuation_bar(tol, fr, Er), (f, fr)
    g, E3 = ev.gradient(psi, gpu_op(tab), ex)
    gr, _ = O.gradient(ref_psi, O.Operator(*tab), oex, mode=0)
    assert rel_err(g, gr) < tol
    assert abs(psi.norm(ex) - O.psi_norm(ref_psi, oex)) < tol * O.psi_norm(ref_psi, oex)


@pytest.mark.parametrize("precision", PRECISIONS)
def test_s_matrix_and_o_k_vector(precision):
    N, M = 5, 6
    b, W, _ = random_rbm(N, M, seed=23, scale=0.2)
    psi = gpu_psi(b, W, precision=precision)
    ref = O.RBM(b, W)
    S = R.get_S_matrix(psi, R.ExactSummation(N, True))
    assert rel_err(S, O.s_matrix(ref, O.Exact(N))) < TOL[precision]
    for c in (0, 11, 31):
        ok = R.psi_O_k_vector(psi, R.Spins(c, N))
        okr = O.o_k(ref, O.spins_words([c])[0])
        assert len(ok) == M + N * M
        assert rel_err(ok, okr[2 * N:]) < TOL[precision]


@pytest.mark.parametrize("precision", PRECISIONS)
def test_hs_distance_nonunitary_and_mc(precision):
    N, M = 7, 9
    b, W, _ = random_rbm(N, M, seed=5, scale=0.2)
    b2, W2, _ = random_rbm(N, M, seed=6, scale=0.2)
    tab = HAM.heisenberg([(i, (i + 1) % N) for i in range(N)])
    op_t = (1e-2j * tab[0], tab[1], tab[2])
    psi, psi2 = gpu_psi(b, W, precision=precision), gpu_psi(b2, W2, precision=precision)
    hs = R.HilbertSpaceDistance(N, psi2.num_params, True)
    g, d = hs.gradient(psi, psi2, gpu_op(op_t), False, R.ExactSummation(N, True), 0.5)
    gr, dr = O.hs_distance(O.RBM(b, W), O.RBM(b2, W2), O.Operator(*op_t), False, O.Exact(N), want_grad=True, nu=0.5)
    assert abs(d - dr) < TOL[precision] * dr
    assert rel_err(g, gr) < TOL[precision] * 10
    # Monte Carlo: same samples as the oracle's chains
    mc = R.MonteCarloLoop(32 * 40, 1, 5, 32, True, seed=11)
    gm, dm = hs.gradient(psi, psi2, gpu_op(op_t), False, mc, 0.5)
    omc = O.MonteCarlo(N, 32 * 40, 1, 5, 32, seed=11, call_index=0)
    grm, drm = O.hs_distance(O.RBM(b, W), O.RBM(b2, W2), O.Operator(*op_t), False, omc, want_grad=True, nu=0.5)
    if precision == 64:
        assert abs(dm - drm) < 1e-10 * drm
        assert rel_err(gm, grm) < 1e-9


def axis_angles(N, seed):
    rng = np.random.default_rng(seed)
    alpha, beta = rng.uniform(-0.4, 0.4, N), rng.uniform(-0.4, 0.4, N)
    return alpha, beta, alpha + rng.uniform(-0.1, 0.1, N), beta + rng.uniform(-0.1, 0.1, N)


@pytest.mark.parametrize("precision", PRECISIONS)
def test_hs_distance_free_quantum_axis(precision):
    """free_quantum_axis on psi: distance and gradient, alpha / beta entries
    included, against the oracle on the exact sum, on a magnetisation sector
    and, in FP64, on the oracle's own chains."""
    print(f"[hs free axis] precision={precision}", flush=True)
    N, M = 7, 9
    b, W, _ = random_rbm(N, M, seed=5, scale=0.2)
    b2, W2, _ = random_rbm(N, M, seed=6, scale=0.2)
    alpha, beta, alpha_p, beta_p = axis = axis_angles(N, 71)
    tab = HAM.heisenberg([(i, (i + 1) % N) for i in range(N)])
    op_t = (1e-2j * tab[0], tab[1], tab[2])
    psi = R.Psi(alpha, beta, b, W, 1.0, True, True, precision=precision)
    psi2 = R.Psi(alpha_p, beta_p, b2, W2, 1.0, False, True, precision=precision)
    hs = R.HilbertSpaceDistance(N, psi2.num_params, True)
    ex = R.ExactSummation(N, True)
    g, d = hs.gradient(psi, psi2, gpu_op(op_t), False, ex, 0.5)
    gr, dr = O.hs_distance(O.RBM(b, W), O.RBM(b2, W2), O.Operator(*op_t), False, O.Exact(N), want_grad=True, nu=0.5,
                           axis=axis)
    assert abs(d - dr) < TOL[precision] * dr
    assert rel_err(g, gr) < TOL[precision] * 10
    assert rel_err(g[:2 * N], gr[:2 * N]) < TOL[precision] * 10
    assert abs(hs(psi, psi2, gpu_op(op_t), False, ex) - d) < 1e-12 * d
    ex = R.ExactSummation(N, True)
    ex.set_total_z_symmetry(1)
    g, d = hs.gradient(psi, psi2, gpu_op(op_t), False, ex, 0.5)
    gr, dr = O.hs_distance(O.RBM(b, W), O.RBM(b2, W2), O.Operator(*op_t), False, O.Exact(N, 1),
                           want_grad=True, nu=0.5, axis=axis)
    assert abs(d - dr) < TOL[precision] * dr
    assert rel_err(g, gr) < TOL[precision] * 10
    if precision == 64:
        mc = R.MonteCarloLoop(32 * 40, 1, 5, 32, True, seed=11)
        gm, dm = hs.gradient(psi, psi2, gpu_op(op_t), False, mc, 0.5)
        omc = O.MonteCarlo(N, 32 * 40, 1, 5, 32, seed=11, call_index=0)
        grm, drm = O.hs_distance(O.RBM(b, W), O.RBM(b2, W2), O.Operator(*op_t), False, omc, want_grad=True, nu=0.5,
                                 axis=axis)
        assert abs(dm - drm) < 1e-10 * drm
        assert rel_err(gm, grm) < 1e-9


# ------------------------------------------------------------------ Monte Carlo
@pytest.mark.parametrize("visible", [False, True])
def test_mc_chains_bit_identical_fp64(visible):
    """FP64 kernels replay the oracle's chains: identical spins at every
    sample, identical acceptance counts, E_loc to 1e-10."""
    print(f"[mc identity] visible={visible}", flush=True)
    N, M = 9, 14
    b, W, a = random_rbm(N, M, seed=31, scale=0.3, visible=visible)
    tab = HAM.tfim_1d(N, 1.0, 0.8)
    psi = gpu_psi(b, W, a, precision=64)
    chains, per = 64, 30
    mc = R.MonteCarloLoop(chains * per, 2, 5, chains, True, seed=99)
    ev = R.ExpectationValue(True)
    g, E = ev.gradient(psi, gpu_op(tab), mc)
    spins, lp, el = mc.last_samples()
    omc = O.MonteCarlo(N, chains * per, 2, 5, chains, seed=99, call_index=0)
    rs, rlp, rel, _ = O.records(O.RBM(b, W, a), O.Operator(*tab), omc)
    assert np.array_equal(spins, rs)
    assert rel_err(lp, rlp) < 1e-12 and rel_err(el, rel) < 1e-10
    acc = mc.acceptance_rate
    assert abs(acc - omc.acceptances / (omc.acceptances + omc.rejections)) < 1e-7
    gr, Er = O.gradient(O.RBM(b, W, a), O.Operator(*tab), O.MonteCarlo(N, chains * per, 2, 5, chains, seed=99), mode=1)
    assert abs(E - Er) < 1e-10 * abs(Er) and rel_err(g, gr) < 1e-10
    # the next call continues the stream: different samples
    ev.gradient(psi, gpu_op(tab), mc)
    spins2, _, _ = mc.last_samples()
    assert not np.array_equal(spins, spins2)
    rs2, _, _, _ = O.records(O.RBM(b, W, a), O.Operator(*tab), O.MonteCarlo(N, chains * per, 2, 5, chains, seed=99, call_index=1))
    assert np.array_equal(spins2, rs2)


def test_mc_fp32_follows_fp64_chains():
    N, M = 10, 20
    b, W, _ = random_rbm(N, M, seed=41, scale=0.05, fp32_exact=True)
    tab = HAM.tfim_1d(N)
    chains, per = 256, 20
    out = {}
    for precision in (32, 64):
        mc = R.MonteCarloLoop(chains * per, 1, 10, chains, True, seed=5)
        R.ExpectationValue(True)(gpu_psi(b, W, precision=precision), gpu_op(tab), mc)
        out[precision] = mc.last_samples()
    same = np.mean(np.all(out[32][0] == out[64][0], axis=1))
    assert same > 0.98
    ok = np.all(out[32][0] == out[64][0], axis=1)
    assert rel_err(out[32][2][ok], out[64][2][ok]) < 1e-5


def _chain_mean_se(x, chains):
    per_chain = x.reshape(-1, chains).mean(axis=0)
    return per_chain.mean(), per_chain.std(ddof=1) / np.sqrt(chains)


@pytest.mark.parametrize("precision", PRECISIONS)
def test_mc_energy_within_three_standard_errors(precision):
    """1-D TFIM at reduced size (N = 12, M = 24): sampled <H> vs ExactSummation."""
    N, M = 12, 24
    b, W, _ = random_rbm(N, M, seed=51, scale=0.1)
    tab = HAM.tfim_1d(N, 1.0, 1.0)
    psi = gpu_psi(b, W, precision=precision)
    ex = R.ExactSummation(N, True)
    psi.normalize(ex)
    ev = R.ExpectationValue(True)
    E_exact = ev(psi, gpu_op(tab), ex).real
    chains, per = 1024, 100
    mc = R.MonteCarloLoop(chains * per, 1, 50, chains, True, seed=3)
    E_mc = ev(psi, gpu_op(tab), mc)
    _, _, el = mc.last_samples()
    m, se = _chain_mean_se(el.real, chains)
    assert abs(m - E_mc.real) < 1e-9 * abs(m)
    assert abs(E_mc.real - E_exact) < 3 * se, (E_mc, E_exact, se)
    assert 0.05 < mc.acceptance_rate < 1.0


def test_mc_gradient_matches_exact_statistically():
    N, M = 10, 12
    b, W, _ = random_rbm(N, M, seed=61, scale=0.15)
    tab = HAM.tfim_1d(N, 1.0, 1.0)
    psi = gpu_psi(b, W, precision=32)
    ex = R.ExactSummation(N, True)
    psi.normalize(ex)
    ev = R.ExpectationValue(True)
    g_ex, _ = ev.gradient(psi, gpu_op(tab), ex)
    reps = [ev.gradient(psi, gpu_op(tab), R.MonteCarloLoop(512 * 50, 1, 30, 512, True, seed=100 + r))[0] for r in range(6)]
    reps = np.array(reps)
    mean, se = reps.mean(axis=0), reps.std(axis=0, ddof=1) / np.sqrt(len(reps))
    big = np.abs(g_ex) > 0.05 * np.abs(g_ex).max()
    z = np.abs(mean - g_ex)[big] / np.maximum(np.abs(se[big]), 1e-12)
    assert np.median(z) < 1.5 and np.mean(z < 4) > 0.95


# ------------------------------------------------------------------ full size: 1-D TFIM N = 20, M = 80
@pytest.fixture(scope="module")
def c2_golden():
    return np.load(os.path.join(GOLDEN, "c2_tfim_rbm.npz"))


@pytest.mark.parametrize("precision", PRECISIONS)
def test_c2_exact_summation_matches_golden(c2_golden, precision):
    """N = 20, M = 80 TFIM with N(0, 0.01^2) parameters, ExactSummation over
    all 2^20 states: the normalising prefactor, energy, 1720-entry gradient,
    and the Hilbert-space distance (and its gradient) to the W + N(0, 1e-3^2)
    copy, against the oracle's values (tests/golden/c2_tfim_rbm.npz): 1e-10
    in complex128, 1e-5 in complex64.  The complex64 distance is formed from
    the per-unit log-cosh differences of the two angle sets
    (k_rbm_log_ratio), which keeps the small log psi' - log psi accurate.
    In complex128 the distance d = sqrt(1 - F) of the nearby states
    (F = |<psi|psi'>|^2 / (<psi|psi><psi'|psi'>) = 1 - 1.2e-5) turns relative
    errors of F's 2^20-term sums into (1 - d^2) / (2 d^2) = 4e4 times larger
    ones of d, so there the 1e-10 bar holds F and the gradient of F,
    -2 d grad d, which carry no such factor."""
    print(f"[c2 golden] precision={precision}", flush=True)
    d = c2_golden
    tol = TOL[precision]
    N = 20
    op = gpu_op(HAM.tfim_1d(N))
    ident = R.Operator(np.ones(1, complex), np.zeros((1, 1), np.int32), -np.ones((1, 1), np.int32), True)
    ex = R.ExactSummation(N, True)
    psi = gpu_psi(d["b"], d["W"], precision=precision)
    psi.normalize(ex)
    errs = {"prefactor": abs(psi.prefactor - d["prefactor"]) / d["prefactor"]}
    psi.prefactor = float(d["prefactor"])
    g, E = R.ExpectationValue(True).gradient(psi, op, ex)
    errs["energy"] = abs(E - d["energy"]) / abs(d["energy"])
    errs["gradient"] = rel_err(g, d["gradient"])
    hs = R.HilbertSpaceDistance(N, psi.num_params, True)
    psi2 = gpu_psi(d["b"], d["W2"], precision=precision)
    hg, hd = hs.gradient(psi, psi2, ident, True, ex, 1.0)
    d_ref, g_ref = float(d["hs_distance"]), d["hs_gradient"]
    if precision == 64:
        errs["hs_fidelity"] = abs((1 - hd ** 2) - (1 - d_ref ** 2)) / (1 - d_ref ** 2)
        errs["hs_fidelity_gradient"] = rel_err(-2 * hd * hg, -2 * d_ref * g_ref)
    else:
        errs["hs_distance"] = abs(hd - d_ref) / d_ref
        errs["hs_gradient"] = rel_err(hg, g_ref)
    print("  " + "  ".join(f"{k} {v:.2e}" for k, v in errs.items()), flush=True)
    for k, v in errs.items():
        assert v < tol, (k, v)


def test_c2_sampled_energy_and_gradient_against_exact(c2_golden):
    """The plain GPU chain at the full configuration (4096 chains, 200
    thermalisation and 1000 measurement sweeps of 20 flips) against the exact
    2^20 values: the energy within 3 standard errors (per-chain means), and a
    z-test of the 1720-entry gradient over six independent seeds (median
    |z| < 1.5, 95 % of |z| < 4 on the entries above 5 % of the largest)."""
    print("[c2 sampled vs exact]", flush=True)
    d = c2_golden
    N, chains, per = 20, 4096, 1000
    psi = gpu_psi(d["b"], d["W"], prefactor=float(d["prefactor"]), precision=32)
    op = gpu_op(HAM.tfim_1d(N))
    ev = R.ExpectationValue(True)
    grads = []
    for rep in range(6):
        mc = R.MonteCarloLoop(chains * per, 1, 200, chains, True, seed=2 + rep)
        g, E_mc = ev.gradient(psi, op, mc)
        grads.append(g)
        if rep == 0:
            _, _, el = mc.last_samples()
            m, se = _chain_mean_se(el.real, chains)
            assert abs(E_mc.real - d["energy"].real) < 3 * se, (E_mc, d["energy"], se)
    grads = np.array(grads)
    mean, se = grads.mean(axis=0), grads.std(axis=0, ddof=1) / np.sqrt(len(grads))
    g_ex = d["gradient"]
    big = np.abs(g_ex) > 0.05 * np.abs(g_ex).max()
    z = np.abs(mean - g_ex)[big] / np.maximum(np.abs(se[big]), 1e-300)
    print(f"  gradient z over {int(big.sum())} entries: median {np.median(z):.2f}, "
          f"fraction below 4: {np.mean(z < 4):.3f}", flush=True)
    assert np.median(z) < 1.5 and np.mean(z < 4) > 0.95


@pytest.mark.parametrize("N,M", [(100, 200), (36, 144), (128, 16), (50, 7), (20, 80)])
def test_mc_gradient_contraction_shapes(N, M):
    """The O_k contraction at the benchmark shapes (up to 9 row tiles and 26
    column tiles): the complex64 gradient of recorded Metropolis samples against
    the FP64 centred estimator 2/n sum_s (E_s - <E>) conj(O_k(s)) on the same
    samples and local energies, O_k from the oracle's my_tanh.  The complex64
    local energies and log psi of these larger models (sums of up to 100 ratios
    over 200 units) are held to 1e-5 against the oracle, apart from the
    configurations whose FP32 and FP64 angles lie on different sides of the
    my_logcosh branch cut (counted and bounded by the jump,
    helpers.check_local_energies)."""
    print(f"[gradient shapes] N={N} M={M}", flush=True)
    b, W, _ = random_rbm(N, M, seed=3 * N + M, scale=0.05)
    psi = gpu_psi(b, W, precision=32)
    tab = HAM.tfim_1d(N, 1.0, 1.0)
    chains, per = 512, 6
    mc = R.MonteCarloLoop(chains * per, 1, 5, chains, True, seed=N)
    g, E = R.ExpectationValue(True).gradient(psi, gpu_op(tab), mc)
    spins, lp, el = mc.last_samples()
    s = spins_pm(spins, N)
    T = O.tanh(b[None, :] + s @ W)
    Ok = np.concatenate([T, (s[:, :, None] * T[:, None, :]).reshape(len(s), N * M)], axis=1)
    El = O.local_energy(O.RBM(b, W), O.Operator(*tab), spins)
    check_local_energies(f"TFIM N={N} M={M}", el, El, lp, O.log_psi(O.RBM(b, W), spins), b, W, s, tab)
    ref = (2.0 / len(s)) * ((el - el.mean())[:, None] * Ok.conj()).sum(axis=0)
    err = rel_err(g[2 * N:], ref)
    print(f"  gradient rel err {err:.2e}", flush=True)
    assert err < 1e-5


def test_profiling_switches_bracket_gpu_work():
    """start_profiling / stop_profiling (source/RBMonGPU.cu:11-17) as roctx
    collection switches: they bracket device work and leave results intact."""
    b, W, _ = random_rbm(8, 8, seed=91, scale=0.1)
    psi = gpu_psi(b, W, precision=64)
    ex = R.ExactSummation(8, True)
    op = gpu_op(HAM.tfim_1d(8))
    E0 = R.ExpectationValue(True)(psi, op, ex)
    R.start_profiling()
    E1 = R.ExpectationValue(True)(psi, op, ex)
    R.stop_profiling()
    assert E1 == E0
