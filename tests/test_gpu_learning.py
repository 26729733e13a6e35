"""The training driver (rbmongpu_amd.learning.LearningByGradientDescent,
pyRBMonGPU/LearningByGradientDescent.py) on the GPU: unitary-evolution
learning (optimize_for, padam with and without the L2 regularisation) closes
most of the distance to exp(i t H) psi_0, and the ground-state search
(find_low_laying_state, imaginary-time steps with 1 - 0.1 H) lowers the
energy toward the exact ground state; both on ExactSummation of small
systems, against dense numpy references."""
import numpy as np
import pytest

from helpers import dense_hamiltonian, random_rbm

import rbmongpu_amd as R
from rbmongpu_amd import hamiltonians as HAM

pytestmark = pytest.mark.gpu


def gpu_psi(b, W, precision=64):
    N = W.shape[0]
    return R.Psi(np.zeros(N), np.zeros(N), b, W, 1.0, False, True, precision=precision)


def test_optimize_for_learns_a_small_time_step():
    print("[learning] optimize_for", flush=True)
    N, M = 6, 12
    b, W, _ = random_rbm(N, M, seed=81, scale=0.3)
    ex = R.ExactSummation(N, True)
    psi_0 = gpu_psi(b, W)
    psi_0.normalize(ex)
    tab = HAM.tfim_1d(N, 1.0, 1.0)
    t = 0.05
    op = (1j * t * tab[0], tab[1], tab[2])
    hs = R.HilbertSpaceDistance(N, psi_0.num_params, True)
    d0 = hs(psi_0, psi_0, R.Operator(*op, True), False, ex)
    L = R.LearningByGradientDescent(psi_0, ex)
    psi = L.optimize_for(psi_0, +psi_0, op, is_unitary=False, eta=1e-2)
    d1 = hs(psi_0, psi, R.Operator(*op, True), False, ex)
    # the learned state against the dense exp(i t H) psi_0
    H = dense_hamiltonian(N, *tab)
    e, U = np.linalg.eigh(H)
    v0 = psi_0.vector
    target = U @ (np.exp(1j * e * t) * (U.conj().T @ v0))
    v1 = psi.vector
    fid = abs(np.vdot(v1, target)) ** 2 / (np.vdot(v1, v1).real * np.vdot(target, target).real)
    fid0 = abs(np.vdot(v0, target)) ** 2 / (np.vdot(v0, v0).real * np.vdot(target, target).real)
    print(f"distance {d0:.4g} -> {d1:.4g}; infidelity {1 - fid0:.3g} -> {1 - fid:.3g}; best {L.best_solution['name']}",
          flush=True)
    assert d1 < 0.3 * d0
    assert 1 - fid < 0.3 * (1 - fid0)
    assert L.report["best_solution_name"] in ("padam", "padam_0")


def test_find_low_laying_state_lowers_the_energy():
    print("[learning] find_low_laying_state", flush=True)
    N, M = 6, 12
    b, W, _ = random_rbm(N, M, seed=83, scale=0.1)
    ex = R.ExactSummation(N, True)
    psi_init = gpu_psi(b, W)
    psi_init.normalize(ex)
    tab = HAM.tfim_1d(N, 1.0, 1.0)
    E_exact = np.linalg.eigvalsh(dense_hamiltonian(N, *tab))[0]
    L = R.LearningByGradientDescent(psi_init, ex)
    psi = L.find_low_laying_state(psi_init, tab, min_steps=200, max_steps=600, eta=1e-2)
    E0, E1 = L.energy_history[0], L.energy_history[-1]
    E_psi = R.ExpectationValue(True)(psi, R.Operator(*tab, True), ex).real
    print(f"energy {E0:.5f} -> {E1:.5f} (psi {E_psi:.5f}), exact {E_exact:.5f}, {len(L.energy_history)} steps", flush=True)
    assert E1 < E0
    assert E1 - E_exact < 1e-2 * (E0 - E_exact)
    assert abs(E_psi - E1) < 1e-6 * abs(E1) + 1e-3 * (E0 - E_exact)
