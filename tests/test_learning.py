"""Host side of the training driver (rbmongpu_amd.learning): the L2
regularisation rule (pyRBMonGPU/L2Regularization.py:1-13), the operator
tables the driver builds itself (identity, 1 - 0.1 H) against dense matrices,
and the convergence tests on synthetic histories.  The GPU steps are covered
by tests/test_gpu_learning.py."""
import numpy as np
import pytest

from helpers import dense_hamiltonian

import rbmongpu_amd as R
from rbmongpu_amd import hamiltonians as HAM
from rbmongpu_amd.learning import affine_operator, identity_operator


def test_l2_regularization_rule():
    reg = R.L2Regularization(1e-2, 0.96)
    p = np.array([0.5 + 1j, -2e-4 + 3j, -0.3, 1e-3])
    g = reg.gradient(7, p)
    lam = 1e-2 * 0.96 ** 7
    assert np.allclose(g, [2 * lam * 0.5, 0.0, 2 * lam * -0.3, 2 * lam * 1e-3], rtol=1e-15, atol=0)
    assert reg.lambda_(0) == 1e-2


def _dense(op, N):
    return dense_hamiltonian(N, op.coefficients, op.pauli_types, op.pauli_indices)


def test_identity_and_affine_operators_are_the_dense_expressions():
    N = 5
    H = R.Operator(*HAM.heisenberg([(i, (i + 1) % N) for i in range(N)]), True)
    Hd = _dense(H, N)
    assert np.allclose(_dense(identity_operator(), N), np.eye(2 ** N), atol=0)
    A = affine_operator(H, -0.1, 1.0)
    assert np.allclose(_dense(A, N), np.eye(2 ** N) - 0.1 * Hd, atol=1e-14)
    assert A.is_hermitian
    B = affine_operator(HAM.tfim_1d(N), 0.5j, -2.0)
    assert np.allclose(_dense(B, N), -2.0 * np.eye(2 ** N) + 0.5j * dense_hamiltonian(N, *HAM.tfim_1d(N)), atol=1e-14)


def test_convergence_tests_on_synthetic_histories():
    N, M = 4, 3
    psi = R.Psi(np.zeros(N), np.zeros(N), np.zeros(M, complex), np.zeros((N, M), complex), 1.0, False, True)
    L = R.LearningByGradientDescent(psi, R.ExactSummation(N, True))
    assert not L.mc
    L.distance_history = list(np.exp(-np.linspace(0, 3, 120)))
    assert L.is_descending and not L.is_learning_poorly and not L.is_highly_fluctuating
    L.distance_history = [0.2] * 120
    assert not L.is_descending
    assert L.has_converged([1.0] * 100) and not L.has_converged(list(np.linspace(1.0, 0.9, 100)))
    assert psi.first_layer_params_slice == slice(2 * N, psi.num_params)


def _driver_with_fake_objective(cost_of_step):
    """A driver whose objective is host arithmetic: gradient = params (so the
    optimiser moves), cost = cost_of_step(step)."""
    N, M = 3, 2
    psi = R.Psi(np.zeros(N), np.zeros(N), np.zeros(M, complex), np.zeros((N, M), complex), 1.0, False, True)
    L = R.LearningByGradientDescent(psi, R.ExactSummation(N, True))
    L.psi_init = psi
    L.algorithm_kwargs = dict(eta=1e-3)
    L._objective = lambda step, params: (np.ones_like(params), cost_of_step(step))
    return L


def test_descent_stages_extend_only_while_descending():
    """LearningByGradientDescent.py:238-252: 400 steps, then 200 more while
    the smoothed log cost still falls by more than 0.05 over the last 50
    steps, checked while at most 1000 steps were taken (so at most 1200)."""
    steady = _driver_with_fake_objective(lambda step: 0.3)
    steady.do_the_gradient_descent("padam")
    assert len(steady.distance_history) == 400
    falling = _driver_with_fake_objective(lambda step: np.exp(-step / 100.0))
    falling.do_the_gradient_descent("padam")
    assert len(falling.distance_history) == 1200
    # falls until step 500, flat afterwards: extended once (to 600), then the
    # last 50 steps are flat and the run stops
    knee = _driver_with_fake_objective(lambda step: np.exp(-min(step, 500) / 50.0))
    knee.do_the_gradient_descent("padam")
    assert len(knee.distance_history) == 600


def test_optimize_for_retries_learning_rates_and_raises():
    """optimize_for accepts a run whose smoothed final cost is below 0.4,
    tries the caller's rate, then 0.1 and 0.01, and raises DidNotConverge when
    no run is accepted (LearningByGradientDescent.py:301-329)."""
    L = _driver_with_fake_objective(lambda step: 0.5)
    seen = []
    real = L._descent_run

    def spy(psi_init, method):
        seen.append((L.algorithm_kwargs["eta"], method))
        return real(psi_init, method)

    L._descent_run = spy
    from rbmongpu_amd.learning import _TrackState
    import rbmongpu_amd.learning as LM
    orig = LM._TrackState
    LM._TrackState = lambda driver: (lambda step, params: (np.ones_like(params), 0.5))
    try:
        with pytest.raises(R.DidNotConverge):
            L.optimize_for(L.psi_init, L.psi_init, HAM.tfim_1d(3), eta=0.3)
    finally:
        LM._TrackState = orig
    assert seen == [(0.3, "padam"), (0.3, "padam_0"), (0.1, "padam"), (0.1, "padam_0"),
                    (0.01, "padam"), (0.01, "padam_0")]
    assert _TrackState is orig
