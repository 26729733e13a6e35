"""GPU checks of MonteCarloLoop.set_total_z_symmetry (exchange move) and
set_fast_sweep (multi-flip move): complex128 chains reproduce the oracle's
chains bit for bit (configurations, log psi 1e-12, E_loc 1e-10); complex64 and
BF16 chains stay in their sector and agree with ExactSummation over the same
sector within 3 standard errors."""
import numpy as np
import pytest

from helpers import O, deep_layers, random_rbm, rel_err

import rbmongpu_amd as R
from rbmongpu_amd import hamiltonians as HAM

pytestmark = pytest.mark.gpu


def gpu_psi(b, W, a=None, precision=64):
    N = W.shape[0]
    return R.Psi(np.zeros(N), np.zeros(N), b, W, 1.0, False, True, a=a, precision=precision)


def gpu_deep(N, biases, conns, weights, final, precision):
    return R.PsiDeep(np.zeros(N), np.zeros(N), list(biases), [np.asarray(c, np.uint32) for c in conns], list(weights),
                     np.asarray(final, complex), 1.0, False, True, precision=precision)


def make_mc(move, N, chains, per, sweeps, therm, seed):
    mc = R.MonteCarloLoop(chains * per, sweeps, therm, chains, True, seed=seed)
    if move == "exchange":
        mc.set_total_z_symmetry(0 if N % 2 == 0 else 1)
        ref = O.MonteCarlo(N, chains * per, sweeps, therm, chains, seed=seed, total_z=0 if N % 2 == 0 else 1)
    else:
        mc.set_fast_sweep(2)
        ref = O.MonteCarlo(N, chains * per, sweeps, therm, chains, seed=seed, fast_sweep=2)
    return mc, ref


@pytest.mark.parametrize("move", ["exchange", "fast_sweep"])
@pytest.mark.parametrize("model", ["rbm", "rbm_visible", "deep"])
def test_moves_bit_identical_to_oracle_fp64(move, model):
    print(f"[moves fp64] {move} {model}", flush=True)
    N = 8
    tab = HAM.heisenberg([(i, (i + 1) % N) for i in range(N)])
    if model == "deep":
        biases, conns, weights, final = deep_layers(N, [16, 8], [4, 16], seed=9, scale=0.25)
        psi = gpu_deep(N, biases, conns, weights, final, 64)
        ref = O.Deep(N, biases, conns, weights, final)
    else:
        b, W, a = random_rbm(N, 10, seed=19, scale=0.3, visible=model == "rbm_visible")
        psi = gpu_psi(b, W, a, 64)
        ref = O.RBM(b, W, a)
    chains, per = 32, 12
    mc, omc = make_mc(move, N, chains, per, 2, 3, seed=41)
    E = R.ExpectationValue(True)(psi, R.Operator(*tab, True), mc)
    spins, lp, el = mc.last_samples()
    rs, rlp, rel, _ = O.records(ref, O.Operator(*tab), omc)
    assert np.array_equal(spins, rs)
    assert rel_err(lp, rlp) < 1e-12 and rel_err(el, rel) < 1e-10
    assert abs(E - rel.mean()) < 1e-10 * abs(E)
    assert abs(mc.acceptance_rate - omc.acceptances / (omc.acceptances + omc.rejections)) < 1e-6
    if move == "exchange":
        ups = np.array([bin(int(s)).count("1") for s in spins[:, 0]])
        assert np.all(ups == N // 2)


@pytest.mark.parametrize("precision", [32, 16])
def test_exchange_energy_matches_sector_exact_summation(precision):
    print(f"[exchange energy] precision={precision}", flush=True)
    N = 12
    tab = HAM.j1j2_chain(N)
    if precision == 16:
        biases, conns, weights, final = deep_layers(N, [24, 12], [12, 24], seed=3, scale=0.08)
        psi = gpu_deep(N, biases, conns, weights, final, 16)
    else:
        b, W, _ = random_rbm(N, 24, seed=21, scale=0.15)
        psi = gpu_psi(b, W, precision=32)
    ex = R.ExactSummation(N, True)
    ex.set_total_z_symmetry(0)
    psi.normalize(ex)
    ev = R.ExpectationValue(True)
    E_exact = ev(psi, R.Operator(*tab, True), ex).real
    chains, per = 512, 60
    mc = R.MonteCarloLoop(chains * per, 1, 30, chains, True, seed=13)
    mc.set_total_z_symmetry(0)
    E_mc = ev(psi, R.Operator(*tab, True), mc).real
    spins, _, el = mc.last_samples()
    ups = np.array([bin(int(s)).count("1") for s in spins[:, 0]])
    assert np.all(ups == N // 2)
    per_chain = el.real.reshape(per, chains).mean(axis=0)
    se = per_chain.std(ddof=1) / np.sqrt(chains)
    assert abs(E_mc - E_exact) < 3 * se, (E_mc, E_exact, se)


def test_fast_sweep_energy_matches_exact_summation():
    N = 10
    b, W, _ = random_rbm(N, 20, seed=23, scale=0.15)
    tab = HAM.tfim_1d(N)
    psi = gpu_psi(b, W, precision=32)
    ex = R.ExactSummation(N, True)
    psi.normalize(ex)
    ev = R.ExpectationValue(True)
    E_exact = ev(psi, R.Operator(*tab, True), ex).real
    chains, per = 512, 80
    mc = R.MonteCarloLoop(chains * per, 4, 20, chains, True, seed=17)
    mc.set_fast_sweep(2)
    E_mc = ev(psi, R.Operator(*tab, True), mc).real
    _, _, el = mc.last_samples()
    per_chain = el.real.reshape(per, chains).mean(axis=0)
    se = per_chain.std(ddof=1) / np.sqrt(chains)
    assert abs(E_mc - E_exact) < 3 * se, (E_mc, E_exact, se)


def test_move_errors():
    N = 6
    b, W, _ = random_rbm(N, 8, seed=1, scale=0.1)
    psi = gpu_psi(b, W, precision=32)
    op = R.Operator(*HAM.tfim_1d(N), True)
    mc = R.MonteCarloLoop(64, 1, 1, 32, True)
    mc.set_total_z_symmetry(1)           # parity of N = 6 needs an even sector
    with pytest.raises(ValueError):
        R.ExpectationValue(True)(psi, op, mc)
    mc = R.MonteCarloLoop(64, 1, 1, 32, True)
    mc.set_total_z_symmetry(0)
    mc.set_fast_sweep(2)
    with pytest.raises(ValueError):
        R.ExpectationValue(True)(psi, op, mc)
