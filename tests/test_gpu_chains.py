"""Monte Carlo chain state on the GPU: the complex64 sampler's incremental
angle updates (include/quantum_state/Psi.hpp:113-119) stay within a bound of
the angles recomputed from the chain's configuration over a chain as long as
the learned-proposal tier's plain chain (200 + 500 sweeps of 100 flips, 70 000
proposals), and the opt-in persistent chains continue exactly where the
previous call left them (the reference restarts every call,
include/spin_ensembles/MonteCarloLoop.hpp:79-105)."""
import numpy as np
import pytest

from helpers import O, dense_hamiltonian, random_rbm, spins_pm

import rbmongpu_amd as R
from rbmongpu_amd import hamiltonians as HAM

pytestmark = pytest.mark.gpu


def final_angles(b, W, mc, chains, per):
    """Angles from scratch (FP64) of each chain's last configuration, which is
    its last recorded sample (the last proposal ends the last sweep)."""
    spins, _, _ = mc.last_samples()
    last = spins[(per - 1) * chains: per * chains]
    return b[None, :] + spins_pm(last, W.shape[0]) @ W


@pytest.mark.parametrize("resync", [16, 0])
def test_fp32_angle_drift_over_a_long_chain(resync):
    """10x10 TFIM shape (N = 100, M = 200), 128 chains x (200 + 500) sweeps:
    |theta_kernel - theta(s)| at the chain's end.  With the default re-sync
    every 16 sweeps (at most 16 N + 32 updates since the last one) the drift
    stays below 2e-5 max|theta|; without re-sync it is printed for
    comparison."""
    N, M, chains, per = 100, 200, 128, 500
    b, W, _ = random_rbm(N, M, seed=71, scale=0.03)
    psi = R.Psi(np.zeros(N), np.zeros(N), b, W, 1.0, False, True, precision=32)
    mc = R.MonteCarloLoop(chains * per, 1, 200, chains, True, seed=7)
    mc.set_angle_resync(resync)
    mc.record_angles(True)
    R.ExpectationValue(True)(psi, R.Operator(*HAM.tfim_2d(10, 10, 1.0, 3.04), True), mc)
    th_k = mc.last_angles()
    assert th_k.shape == (chains, M)
    th = final_angles(b, W, mc, chains, per)
    drift = np.abs(th_k - th).max() / np.abs(th).max()
    print(f"resync every {resync} sweeps: max |theta_kernel - theta(s)| / max |theta| = {drift:.3e}, "
          f"acceptance {mc.acceptance_rate:.3f}", flush=True)
    if resync:
        assert drift < 2e-5


def test_fp64_chain_angles_match_scratch():
    """complex128 chains keep incremental angles (no re-sync, as the
    reference) that agree with the recomputed ones to FP64 rounding."""
    N, M, chains, per = 20, 80, 64, 100
    b, W, _ = random_rbm(N, M, seed=73, scale=0.1)
    psi = R.Psi(np.zeros(N), np.zeros(N), b, W, 1.0, False, True, precision=64)
    mc = R.MonteCarloLoop(chains * per, 1, 50, chains, True, seed=9)
    mc.record_angles(True)
    R.ExpectationValue(True)(psi, R.Operator(*HAM.tfim_1d(N), True), mc)
    th = final_angles(b, W, mc, chains, per)
    assert np.abs(mc.last_angles() - th).max() < 1e-11 * np.abs(th).max()


@pytest.mark.parametrize("move", ["single_flip", "exchange"])
@pytest.mark.parametrize("precision", [32, 64])
def test_persistent_chains_continue_from_the_last_configuration(precision, move):
    """With num_sweeps = 0 every sample of a call is the chain's configuration
    after thermalisation.  A persistent ensemble's second call skips
    thermalisation and starts where the first ended, so it returns the same
    samples; the default ensemble restarts from new random configurations."""
    N, M, chains, per = 12, 24, 256, 3
    b, W, _ = random_rbm(N, M, seed=75, scale=0.2)
    psi = R.Psi(np.zeros(N), np.zeros(N), b, W, 1.0, False, True, precision=precision)
    op = R.Operator(*HAM.tfim_1d(N), True)
    ev = R.ExpectationValue(True)
    pers = R.MonteCarloLoop(chains * per, 0, 30, chains, True, seed=3)
    if move == "exchange":
        pers.set_total_z_symmetry(0)
    pers.persistent_chains = True
    assert pers.persistent_chains
    plain = R.MonteCarloLoop(pers)
    plain.persistent_chains = False
    ev(psi, op, pers)
    a, _, _ = pers.last_samples()
    ev(psi, op, pers)
    b2, _, _ = pers.last_samples()
    assert np.array_equal(a, b2)
    if move == "exchange":
        s = np.where((a[:, :1] >> np.arange(N, dtype=np.uint64)) & np.uint64(1), 1, -1)
        assert np.all(s.sum(axis=1) == 0)
    ev(psi, op, plain)
    c1, _, _ = plain.last_samples()
    assert np.array_equal(c1, a)          # the same first call
    ev(psi, op, plain)
    c2, _, _ = plain.last_samples()
    assert not np.array_equal(c2, c1)     # a restart
    # turning the flag off and on starts afresh
    pers.persistent_chains = False
    pers.persistent_chains = True
    ev(psi, op, pers)
    d, _, _ = pers.last_samples()
    assert not np.array_equal(d, a)


def test_persistent_chains_sample_psi_squared():
    """Persistent chains without re-thermalisation keep sampling |psi|^2:
    five successive calls' energies agree with ExactSummation within 3
    standard errors (per-chain means)."""
    N, M, chains, per = 10, 20, 512, 20
    b, W, _ = random_rbm(N, M, seed=77, scale=0.2)
    tab = HAM.tfim_1d(N, 1.0, 0.9)
    psi = R.Psi(np.zeros(N), np.zeros(N), b, W, 1.0, False, True, precision=32)
    op = R.Operator(*tab, True)
    ex = R.ExactSummation(N, True)
    psi.normalize(ex)
    ev = R.ExpectationValue(True)
    E_exact = ev(psi, op, ex).real
    mc = R.MonteCarloLoop(chains * per, 1, 50, chains, True, seed=5)
    mc.persistent_chains = True
    means = []
    for call in range(5):
        ev(psi, op, mc)
        _, _, el = mc.last_samples()
        means.append(el.real.reshape(per, chains).mean(axis=0))
    per_chain = np.concatenate(means)
    m, se = per_chain.mean(), per_chain.std(ddof=1) / np.sqrt(len(per_chain))
    assert abs(m - E_exact) < 3 * se, (m, E_exact, se)


def test_persistent_chains_need_an_rbm():
    from helpers import deep_layers
    N = 6
    sizes = [6]
    biases, conns, weights, final = deep_layers(N, sizes, [N], seed=1)
    deep = R.PsiDeep(np.zeros(N), np.zeros(N), biases, conns, weights, final, 1.0, False, True)
    mc = R.MonteCarloLoop(64, 1, 5, 64, True)
    mc.persistent_chains = True
    with pytest.raises(NotImplementedError):
        R.ExpectationValue(True)(deep, R.Operator(*HAM.tfim_1d(N), True), mc)



def test_two_launch_sampling_matches_one_launch(monkeypatch):
    """Low-occupancy complex64 chains (at most one wavefront per SIMD) run in
    several launches split where one launch re-synchronises the angles, each
    launch's samples but the last's evaluated on a second stream while the
    next launch samples: samples, log psi, local energies, acceptance and the
    gradient equal one launch's bit for bit (RBMHIP_MC_SPLIT=0 forces one
    launch)."""
    N, M, chains, per = 20, 80, 1024, 64
    b, W, _ = random_rbm(N, M, seed=75, scale=0.1)
    psi = R.Psi(np.zeros(N), np.zeros(N), b, W, 1.0, False, True, precision=32)
    op = R.Operator(*HAM.tfim_1d(N), True)
    out = {}
    for split in ("0", "1"):
        monkeypatch.setenv("RBMHIP_MC_SPLIT", split)
        mc = R.MonteCarloLoop(chains * per, 1, 10, chains, True, seed=11)
        R.timing_enable(True)
        R.timing_reset()
        g, E = R.ExpectationValue(True).gradient(psi, op, mc)
        R.timing_enable(False)
        launches = R.timing_query("mc")[1]
        spins, lp, el = mc.last_samples()
        out[split] = (launches, g, E, spins, lp, el, mc.acceptance_rate)
        print(f"RBMHIP_MC_SPLIT={split}: {launches} sampler launches, E = {E}", flush=True)
    assert out["0"][0] == 1 and out["1"][0] > 1
    for x, y in zip(out["0"][1:], out["1"][1:]):
        assert np.array_equal(np.asarray(x), np.asarray(y))
