"""Host side of the training driver (rbmongpu_amd.learning): the L2
regularisation rule (pyRBMonGPU/L2Regularization.py:1-13), the operator
tables the driver builds itself (identity, 1 - 0.1 H) against dense matrices,
and the convergence tests on synthetic histories.  The GPU steps are covered
by tests/test_gpu_learning.py."""
import numpy as np

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
