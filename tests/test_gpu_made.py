"""GPU parity of the learned proposal (MADE) and of the independent
Metropolis-Hastings ensemble against oracle/made.py.

Tolerances: training-path losses and gradients (FP32 MFMA against FP64) to
1e-5 / 2e-4 relative; log q of the bfloat16 sampler against the oracle's
mirror of the same arithmetic to 1e-4 absolute (log q ~ -N ln 2); the seeded
initialisation and the accept/reject decisions of the chains bit-exactly;
sampled distributions and energies statistically (z-scores, standard errors)."""
import numpy as np
import pytest

from helpers import O, random_rbm, rel_err
from oracle import made as MD

import rbmongpu_amd as R
from rbmongpu_amd import hamiltonians as HAM

pytestmark = pytest.mark.gpu


def rand_params(N, H, seed, scale=0.5):
    rng = np.random.default_rng(seed)
    M1, M2 = MD.masks(N, H)
    p = MD.pack(scale * rng.normal(size=(H, N)) * M1, scale * rng.normal(size=H),
                scale * rng.normal(size=(N, H)) * M2, scale * rng.normal(size=N))
    return p.astype(np.float32)


def rand_words(N, n, seed):
    rng = np.random.default_rng(seed)
    X = (rng.random((n, N)) < 0.5).astype(np.uint64)
    return MD.words(X), X.astype(np.float64)


def gpu_psi(b, W, precision=64):
    N = W.shape[0]
    return R.Psi(np.zeros(N), np.zeros(N), b, W, 1.0, False, True, precision=precision)


@pytest.mark.parametrize("N,H,seed", [(10, 64, 1), (100, 256, 1234), (65, 100, 7)])
def test_seeded_initialisation_is_bit_exact(N, H, seed):
    print(f"[made init] N={N} H={H}", flush=True)
    md = R.MADE(N, H, seed=seed)
    ref = MD.init_params(N, H, seed).astype(np.float32)
    assert np.array_equal(md.params, ref)


@pytest.mark.parametrize("N,H,n", [(12, 64, 200), (100, 256, 1000), (70, 96, 333)])
def test_loss_and_gradient(N, H, n):
    print(f"[made grad] N={N} H={H} n={n}", flush=True)
    p = rand_params(N, H, seed=N + H)
    words, X = rand_words(N, n, seed=n)
    md = R.MADE(N, H, params=p)
    assert np.array_equal(md.params, p)
    loss, g = md.loss_and_gradient(words)
    pr = p.astype(np.float64)
    assert abs(loss - MD.loss(pr, X, N, H)) < 1e-5 * abs(MD.loss(pr, X, N, H))
    gr = MD.grad(pr, X, N, H)
    assert rel_err(g, gr) < 2e-4
    M1, M2 = MD.masks(N, H)
    masked = MD.pack(M1, np.ones(H), M2, np.ones(N)) == 0
    assert np.all(g[masked] == 0)


def test_adam_steps_follow_the_oracle():
    N, H, n, lr = 16, 64, 256, 1e-2
    p = rand_params(N, H, seed=3)
    words, X = rand_words(N, n, seed=4)
    md = R.MADE(N, H, params=p)
    losses = md.train(words, 3, batch=n, lr=lr)
    q = p.astype(np.float64)
    m = np.zeros_like(q)
    v = np.zeros_like(q)
    gmin = np.full_like(q, np.inf)
    ref_losses = []
    for t in range(1, 4):
        ref_losses.append(MD.loss(q, X, N, H))
        g = MD.grad(q, X, N, H)
        gmin = np.minimum(gmin, np.abs(g))
        q, m, v = MD.adam(q, g, m, v, t, lr=lr)
    assert np.allclose(losses, ref_losses, rtol=1e-5)
    got = md.params.astype(np.float64)
    # Adam's first steps move by about lr * sign(g): compare where the gradient is
    # resolved well above FP32 rounding
    firm = gmin > 1e-4 * np.abs(MD.grad(p.astype(np.float64), X, N, H)).max()
    assert firm.sum() > 0.5 * (gmin > 0).sum()
    assert np.abs(got - q)[firm].max() < 1e-3 * lr
    M1, M2 = MD.masks(N, H)
    masked = MD.pack(M1, np.ones(H), M2, np.ones(N)) == 0
    assert np.all(got[masked] == 0)    # masked weights get no gradient and stay zero


@pytest.mark.parametrize("N,H", [(10, 64), (100, 256), (65, 100), (7, 512)])
def test_log_q_matches_oracle_inference(N, H):
    print(f"[made log q] N={N} H={H}", flush=True)
    p = rand_params(N, H, seed=11 * N + H, scale=0.3)
    words, X = rand_words(N, 300, seed=N)
    md = R.MADE(N, H, params=p)
    got = md.log_q(words)
    ref = MD.log_q_inference(p.astype(np.float64), X, N, H)
    assert np.abs(got - ref).max() < 1e-4
    # and the inference model is close to the FP32 model it was rounded from
    assert np.abs(got - MD.log_q(p.astype(np.float64), X, N, H)).max() < 0.05 * N ** 0.5


# (128, 256): W1 does not fit in LDS as FP32 (196 KB) and the sampler keeps it
# in bfloat16; the other shapes take the FP32 copy.  N = 128 fills all four
# 32-bit words of a configuration.
# N = 33 and 65 end in a partial block of 4 positions just past a 32-bit word.
@pytest.mark.parametrize("N,H", [(20, 64), (33, 64), (65, 128), (100, 256), (128, 256)])
def test_sampler_log_q_is_its_own_score(N, H):
    print(f"[made sampler score] N={N} H={H}", flush=True)
    md = R.MADE(N, H, params=rand_params(N, H, seed=5, scale=0.3))
    spins, lq = md.sample(4096, seed=9)
    assert np.array_equal(lq, md.log_q(spins))
    X = MD.bits(spins, N)
    assert np.abs(lq - MD.log_q_inference(md.params.astype(np.float64), X, N, H)).max() < 1e-4
    # no bit at or above N is set
    if N < 64:
        assert np.all(spins[:, 0] >> np.uint64(N) == 0) and np.all(spins[:, 1] == 0)
    elif N < 128:
        assert np.all(spins[:, 1] >> np.uint64(N - 64) == 0)
    spins2, _ = md.sample(4096, seed=10)
    assert not np.array_equal(spins, spins2)


@pytest.mark.parametrize("N,H", [(20, 64), (100, 256), (30, 320), (128, 256)])
def test_sampler_draws_depend_on_the_proposal_index_only(N, H):
    """Proposal g's uniforms come from its own Philox counter, so a sample of
    n + 1 (odd) starts with the sample of n, whatever rows share a wavefront."""
    md = R.MADE(N, H, params=rand_params(N, H, seed=7, scale=0.3))
    a, la = md.sample(2048, seed=3)
    b, lb = md.sample(2049, seed=3)
    assert np.array_equal(a, b[:2048]) and np.array_equal(la, lb[:2048])
    c, lc = md.sample(1, seed=3)
    assert np.array_equal(c, a[:1]) and np.array_equal(lc, la[:1])
    assert np.array_equal(lb, md.log_q(b))


def test_sampler_draws_q():
    N, H = 6, 16
    p = rand_params(N, H, seed=21, scale=0.8)
    md = R.MADE(N, H, params=p)
    n = 400_000
    spins, _ = md.sample(n, seed=1)
    idx = spins[:, 0].astype(np.int64)
    counts = np.bincount(idx, minlength=2 ** N)
    q = np.exp(MD.log_q_inference(p.astype(np.float64), MD.all_configs(N), N, H))
    z = (counts - n * q) / np.sqrt(n * q * (1 - q))
    assert np.abs(z).max() < 4.5


def test_training_learns_a_distribution():
    """Fit a random teacher MADE from its own samples: the trained loss
    approaches the teacher's entropy."""
    N, H = 12, 64
    teacher = R.MADE(N, H, params=rand_params(N, H, seed=8, scale=0.8))
    data, lq = teacher.sample(100_000, seed=2)
    entropy = -lq.mean()
    student = R.MADE(N, H, seed=3)
    l0, _ = student.loss_and_gradient(data[:4096])
    losses = student.train(data, 1500, batch=1024, lr=3e-3, seed=4)
    lf, _ = student.loss_and_gradient(data[:20000])
    assert l0 > entropy + 0.2
    assert lf < entropy + 0.05, (l0, lf, entropy)
    assert losses[-50:].mean() < losses[:50].mean()


def test_independent_mh_replays_the_oracle_chain():
    """FP64 wavefunction: every accept/reject decision and every recorded
    configuration follows the oracle's independent-MH scan over the same
    proposals and Philox uniforms."""
    N, M, H = 8, 12, 32
    b, W, _ = random_rbm(N, M, seed=71, scale=0.4)
    psi = gpu_psi(b, W, precision=64)
    md = R.MADE(N, H, seed=5)
    chains, per, therm, seed = 64, 40, 7, 17
    ens = R.LearnedMonteCarloLoop(md, chains * per, therm, chains, True, seed=seed)
    tab = HAM.tfim_1d(N, 1.0, 0.7)
    R.ExpectationValue(True)(psi, R.Operator(*tab, True), ens)
    for call in range(2):
        if call:
            R.ExpectationValue(True)(psi, R.Operator(*tab, True), ens)
        prop, plq, plp = ens.last_proposals()
        spins, lp, el = ens.last_samples()
        assert prop.shape[0] == chains * (therm + per + 1)
        # proposal scores
        Xp = MD.bits(prop, N)
        assert np.abs(plq - MD.log_q_inference(md.params.astype(np.float64), Xp, N, H)).max() < 1e-4
        assert rel_err(plp, O.log_psi(O.RBM(b, W), prop)) < 1e-12
        acc = 0
        for c in range(chains):
            t = np.arange(therm + per + 1)
            x = MD.philox_np(c, t, call, 2, seed & 0xFFFFFFFF, seed >> 32)[0]
            U = (x.astype(np.float64) + 0.5) * 2.0 ** -32
            g = t * chains + c
            held, a = MD.imh_chain(plp.real[g], plq[g], U, therm, per)
            acc += a
            assert np.array_equal(spins[np.arange(per) * chains + c], prop[g[held]])
        assert abs(ens.acceptance_rate - acc / (chains * (therm + per))) < 1e-6
        assert rel_err(el, O.local_energy(O.RBM(b, W), O.Operator(*tab), spins)) < 1e-10


@pytest.mark.parametrize("precision", [64, 32])
def test_independent_mh_samples_psi_squared(precision):
    """Any proposal with full support leaves |psi|^2 stationary: an untrained
    network over N = 6 spins, histogram of all samples against |psi|^2 with
    chain-level standard errors."""
    N, M, H = 6, 10, 16
    b, W, _ = random_rbm(N, M, seed=81, scale=0.3)
    psi = gpu_psi(b, W, precision=precision)
    md = R.MADE(N, H, seed=6)
    chains, per = 2048, 200
    # an untrained proposal accepts ~10 %: thermalise well past the start bias
    ens = R.LearnedMonteCarloLoop(md, chains * per, 300, chains, True, seed=2)
    R.ExpectationValue(True)(psi, R.Operator(*HAM.tfim_1d(N), True), ens)
    spins, _, _ = ens.last_samples()
    idx = spins[:, 0].astype(np.int64).reshape(per, chains)
    lpsi = O.log_psi(O.RBM(b, W), MD.words(MD.all_configs(N).astype(np.uint64)))
    prob = np.exp(2 * lpsi.real)
    prob /= prob.sum()
    freq = np.stack([(idx == s).mean(axis=0) for s in range(2 ** N)])   # per state, per chain
    mean = freq.mean(axis=1)
    se = freq.std(axis=1, ddof=1) / np.sqrt(chains)
    # states no chain visited have no chain-level spread: floor with the i.i.d. error
    se = np.maximum(se, np.sqrt(prob * (1 - prob) / (chains * per)))
    z = (mean - prob) / se
    assert np.abs(z).max() < 4.5
    assert 0.05 < ens.acceptance_rate < 1.0


def test_learned_energy_and_gradient_match_exact_summation():
    N, M, H = 10, 20, 64
    b, W, _ = random_rbm(N, M, seed=91, scale=0.25)
    tab = HAM.tfim_1d(N, 1.0, 1.0)
    psi = gpu_psi(b, W, precision=32)
    ex = R.ExactSummation(N, True)
    psi.normalize(ex)
    ev = R.ExpectationValue(True)
    op = R.Operator(*tab, True)
    g_ex, E_ex = ev.gradient(psi, op, ex)
    # train the proposal on plain-chain samples, then sample with it
    mc = R.MonteCarloLoop(1024 * 100, 1, 50, 1024, True, seed=4)
    ev(psi, op, mc)
    data, _, _ = mc.last_samples()
    md = R.MADE(N, H, seed=1)
    md.train(data, 400, batch=2048, lr=3e-3, seed=2)
    chains, per = 1024, 100
    ens = R.LearnedMonteCarloLoop(md, chains * per, 5, chains, True, seed=9)
    g, E = ev.gradient(psi, op, ens)
    _, _, el = ens.last_samples()
    per_chain = el.real.reshape(per, chains).mean(axis=0)
    se = per_chain.std(ddof=1) / np.sqrt(chains)
    assert abs(E.real - E_ex.real) < 3 * se, (E, E_ex, se)
    assert ens.acceptance_rate > 0.3
    assert rel_err(g, g_ex) < 0.1


@pytest.mark.parametrize("N,M", [(100, 200), (37, 50), (128, 16), (5, 3)])
def test_proposal_log_psi_on_mfma_matches_oracle(N, M):
    """complex64 RBMs score the proposals with the theta GEMM on bfloat16 MFMA
    with the weights split into three exact bfloat16 parts: FP32-level error
    against the FP64 oracle, like the FP32 evaluation kernel's."""
    print(f"[proposal log psi] N={N} M={M}", flush=True)
    b, W, _ = random_rbm(N, M, seed=N + M, scale=0.05)
    psi = gpu_psi(b, W, precision=32)
    md = R.MADE(N, 64, seed=1)
    chains = 256
    ens = R.LearnedMonteCarloLoop(md, chains * 3, 2, chains, True, seed=1)
    R.ExpectationValue(True)(psi, R.Operator(*HAM.tfim_1d(N), True), ens)
    prop, _, plp = ens.last_proposals()
    ref = O.log_psi(O.RBM(b, W), prop)
    ev = R.log_psi_batch(psi, prop)
    e_gemm, e_eval = rel_err(plp, ref), rel_err(ev, ref)
    print(f"  rel err vs FP64: split-bf16 GEMM {e_gemm:.2e}, FP32 evaluation kernel {e_eval:.2e}", flush=True)
    assert e_gemm < 1e-5
    assert e_gemm < 4 * e_eval + 1e-7
    # the same configurations through the evaluation kernel
    assert rel_err(ev, plp) < 1e-5


def _tau_int(series, c=5.0):
    """Sokal's automatic windowing on a (T, chains) series (global mean)."""
    x = series - series.mean()
    T = x.shape[0]
    f = np.fft.rfft(x, n=2 * T, axis=0)
    acov = np.fft.irfft(f * np.conj(f), axis=0)[:T].mean(axis=1) / T
    rho = acov / acov[0]
    tau = 0.5
    for w in range(1, T):
        tau += rho[w]
        if w >= c * tau:
            break
    return float(tau)


def test_learned_proposal_gain_on_a_bimodal_state():
    """What the trained network gains: on a two-mode |psi|^2 (one strong hidden
    unit favours all-up and all-down) single flips almost never cross between
    the modes, so the magnetisation of a plain chain decorrelates over hundreds
    of sweeps; the MADE proposal, trained on plain-chain samples of both modes
    (different chains sit in different modes), jumps between them.  Both
    energies agree with ExactSummation."""
    N, M, H = 16, 4, 64
    rng = np.random.default_rng(3)
    b = np.zeros(M, complex)
    W = 0.01 * (rng.normal(size=(N, M)) + 1j * rng.normal(size=(N, M)))
    W[:, 0] = 0.5            # P(m) peaks at |m| = 12 and is ~400x lower at m = 0
    tab = HAM.tfim_1d(N, 1.0, 0.5)
    psi = gpu_psi(b, W, precision=32)
    ex = R.ExactSummation(N, True)
    psi.normalize(ex)
    ev = R.ExpectationValue(True)
    op = R.Operator(*tab, True)
    E_ex = ev(psi, op, ex).real
    chains, per = 512, 100
    plain = R.MonteCarloLoop(chains * per, 1, 50, chains, True, seed=21)
    E_plain = ev(psi, op, plain).real
    sp, _, el_p = plain.last_samples()
    md = R.MADE(N, H, seed=4)
    md.train(sp, 600, batch=2048, lr=3e-3, seed=5)
    learned = R.LearnedMonteCarloLoop(md, chains * per, 10, chains, True, seed=22)
    E_learned = ev(psi, op, learned).real
    sl, _, el_l = learned.last_samples()

    def magnetisation(spins):
        up = np.array([bin(int(s)).count("1") for s in spins[:, 0]])
        return (2.0 * up - N).reshape(per, chains)

    tau_p, tau_l = _tau_int(magnetisation(sp)), _tau_int(magnetisation(sl))
    print(f"  tau_int(m): plain {tau_p:.1f}, learned {tau_l:.2f}; acceptance learned {learned.acceptance_rate:.3f}",
          flush=True)
    assert tau_l * 5 < tau_p
    assert tau_l < 5
    for name, E, el in (("learned", E_learned, el_l), ("plain", E_plain, el_p)):
        per_chain = el.real.reshape(per, chains).mean(axis=0)
        se = per_chain.std(ddof=1) / np.sqrt(chains)
        print(f"  {name}: E {E:.6f} vs exact {E_ex:.6f}, {abs(E - E_ex) / se:.2f} standard errors", flush=True)
        assert abs(E - E_ex) < 3 * se + 1e-6 * abs(E_ex), (E, E_ex, se)
