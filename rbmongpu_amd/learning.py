"""The reference's training driver, LearningByGradientDescent
(pyRBMonGPU/LearningByGradientDescent.py:34-365) and L2Regularization
(pyRBMonGPU/L2Regularization.py:1-13), over this package's GPU calls.

Each step of the driver is one ``HilbertSpaceDistance.gradient`` (unitary
evolution: learn psi ~ U psi_0) or ``ExpectationValue.gradient`` /
``fluctuation`` (ground state) on the GPU; the optimiser (padam by default,
``gradient_descent``) and the bookkeeping are host numpy, as in the reference.

Operators are this package's ``Operator`` objects or ``(coefficients,
pauli_types, pauli_indices)`` tables.  The reference builds them from
QuantumExpression's PauliExpression, which is not vendored; the two
expressions its driver forms itself, the identity and ``1 - 0.1 H``, are
built here on the tables (``identity_operator``, ``affine_operator``).

Deliberate differences from the reference, each where the reference cannot
run as written:
* ``find_low_laying_state`` asks ``get_gradient_descent_algorithm`` for an
  algorithm without a name and indexes the generator as a dict
  (LearningByGradientDescent.py:229-230); here it runs padam on the
  ground-state gradient, ``min_steps`` first, then 100 at a time until the
  energy history converges or ``max_steps`` is reached;
* ``avoid_correlations`` needs HilbertSpaceCorrelations, whose import is
  commented out in the reference (line 10): a non-zero value raises
  NotImplementedError, zero is ignored;
* ``intelligent_optimize_for`` needs PauliExpression term algebra (exp of
  terms, max_norm): it raises NotImplementedError;
* ``save_report`` writes under ``folder`` (default: the current directory)
  instead of a fixed site path; ``min_report`` stores the operator tables;
* the reference's ``HilbertSpaceDistance.gradient`` calls without ``nu``
  (lines 131-133, 142-144, 270-272), which its binding requires, pass nu = 1
  here, and ``do_preconditioning``'s gradient function returns the
  (gradient, distance) pair padam expects.
"""
from __future__ import annotations

import json
import math
import os
from collections import namedtuple
from itertools import islice
from pathlib import Path

import numpy as np
from scipy.ndimage import gaussian_filter1d

from . import gradient_descent as _gd
from ._pyRBMonGPU import ExpectationValue, HilbertSpaceDistance, MonteCarloLoop, Operator
from .json_numpy import NumpyEncoder

modes = namedtuple("modes", "unitary_evolution groundstate")._make(range(2))


class DidNotConverge(RuntimeError):
    pass


class UnstableNetwork(RuntimeError):
    pass


class PoorLearning(RuntimeError):
    pass


class L2Regularization:
    """lambda(step) = lambda0 gamma^step; gradient 2 lambda Re p with entries
    |Re p| < 1e-3 set to zero (L2Regularization.py:1-13)."""

    def __init__(self, lambda0, gamma):
        self.lambda0 = lambda0
        self.gamma = gamma

    def lambda_(self, step):
        return self.lambda0 * self.gamma ** step

    def gradient(self, step, params):
        params_real = np.array(params.real, dtype=float)
        params_real[abs(params_real) < 1e-3] = 0
        return self.lambda_(step) * 2 * params_real


# ---------------------------------------------------------------- operator tables
def _tables(operator):
    if isinstance(operator, Operator):
        return operator.coefficients, operator.pauli_types, operator.pauli_indices
    c, t, i = operator
    return np.asarray(c, dtype=complex), np.asarray(t, dtype=np.int32), np.asarray(i, dtype=np.int32)


def identity_operator(gpu=True):
    """The identity as one Pauli string with no factors."""
    return Operator(np.ones(1, complex), np.zeros((1, 1), np.int32), -np.ones((1, 1), np.int32), gpu)


def affine_operator(operator, scale, shift, gpu=True):
    """shift + scale * operator (the reference forms 1 - 0.1 H this way)."""
    c, t, i = _tables(operator)
    L = t.shape[1]
    return Operator(np.concatenate([[complex(shift)], complex(scale) * c]),
                    np.vstack([np.zeros((1, L), np.int32), t]),
                    np.vstack([-np.ones((1, L), np.int32), i]), gpu)


def _as_operator(operator, gpu):
    return operator if isinstance(operator, Operator) else Operator(*_tables(operator), gpu)


class LearningByGradientDescent:
    """LearningByGradientDescent.py:34-365 (see the module docstring for the
    differences)."""

    def __init__(self, psi, spin_ensemble):
        self.gpu = psi.gpu
        self.mc = type(spin_ensemble) is MonteCarloLoop
        self.spin_ensemble = spin_ensemble
        self.num_params = psi.num_params
        self.hilbert_space_distance = HilbertSpaceDistance(psi.N, self.num_params, psi.gpu)
        self.expectation_value = ExpectationValue(self.gpu)
        self.regularization = None
        self.algorithm_kwargs = {}

    # ------------------------------------------------------------ convergence tests (lines 49-101)
    @property
    def smoothed_distance_history(self):
        distances = np.array(self.distance_history)
        sigma = 5 if self.mc else 1
        return gaussian_filter1d(distances, sigma, mode="nearest")

    def smoothed(self, x):
        return gaussian_filter1d(np.array(x[-100:]), 5, mode="nearest")

    @property
    def slopes(self):
        distances = self.smoothed_distance_history
        return abs(distances[1:] - distances[:-1])

    @property
    def graph(self):
        return np.log(self.smoothed_distance_history)

    @property
    def is_highly_fluctuating(self):
        graph = np.log(self.distance_history[-50:])
        diff_graph = graph[1:] - graph[:-1]
        return np.std(diff_graph) > 0.125

    @property
    def is_descending(self):
        graph = self.graph[-50:]
        return graph[-1] - graph[0] < -0.05

    @property
    def is_learning_poorly(self):
        graph = self.graph
        return np.count_nonzero(graph > graph[0] + 0.2) > 40

    def has_converged(self, x, threshold=None):
        threshold = threshold or (3e-3 if self.mc else 1e-3)
        smoothed_x = self.smoothed(x)
        return abs(smoothed_x[0] - smoothed_x[-1]) < threshold

    def push_params(self, params):
        stack_size = 3
        if not hasattr(self, "last_params_stack"):
            self.last_params_stack = np.empty((stack_size, self.psi.num_params), dtype=complex)
            self.last_params_idx = 0
        self.last_params_stack[self.last_params_idx] = params
        self.last_params_idx = (self.last_params_idx + 1) % stack_size

    @property
    def last_params_avg(self):
        return np.mean(self.last_params_stack, axis=0)

    # ------------------------------------------------------------ one step (lines 103-157)
    def set_params_and_return_gradient(self, step, params):
        self.psi.params = params
        if not self.mc:
            self.psi.normalize(self.spin_ensemble)

        if self.mode == modes.unitary_evolution:
            gradient, distance = self.hilbert_space_distance.gradient(
                self.psi_0, self.psi, self.operator, self.is_unitary, self.spin_ensemble, 0.0)
            distance -= self.distance_0
            if distance < 0:
                gradient *= -1
            gradient *= self.gradient_prefactor
            if self.regularization is not None:
                sl = self.psi.first_layer_params_slice
                gradient[sl] += self.regularization.gradient(step, self.psi.params[sl])
        else:
            if self.imaginary_time_evolution:
                gradient, distance = self.hilbert_space_distance.gradient(
                    self.psi, self.psi, self.opt_operator, True, self.spin_ensemble, 1.0)
            else:
                gradient, energy = self.expectation_value.gradient(self.psi, self.operator, self.spin_ensemble)
                distance = 0.0
            delta_energy, energy = self.expectation_value.fluctuation(self.psi, self.operator, self.spin_ensemble)
            self.delta_energy_history.append(delta_energy)
            energy = energy.real
            self.energy_history.append(energy)

            distances = []
            for excluded_state in self.excluded_states:
                es_gradient, es_distance = self.hilbert_space_distance.gradient(
                    excluded_state, self.psi, self.identity, True, self.spin_ensemble, 1.0)
                distances.append(es_distance)
                gradient -= self.level_repulsion * max(1, min(10, math.log(es_distance.real))) * es_gradient
            self.excluded_states_distances.append(distances)

        if getattr(self, "avoid_correlations", 0):
            raise NotImplementedError("avoid_correlations needs HilbertSpaceCorrelations, which the reference does not ship")

        return gradient, distance

    def get_gradient_descent_algorithm(self, name):
        return getattr(_gd, f"{name}_generator")(
            self.psi_init.params, lambda step, params: self.set_params_and_return_gradient(step, params),
            **self.algorithm_kwargs)

    # ------------------------------------------------------------ reports (lines 170-206)
    @property
    def min_report(self):
        c, t, i = _tables(self.operator)
        result = {"operator": json.loads(json.dumps(
            dict(coefficients=c, pauli_types=t, pauli_indices=i), cls=NumpyEncoder))}
        for key in ("is_unitary", "distance_history", "energy_history", "delta_energy_history",
                    "excluded_states_distances", "algorithm_kwargs"):
            if hasattr(self, key):
                result[key] = json.loads(json.dumps(getattr(self, key), cls=NumpyEncoder))
        return result

    @property
    def report(self):
        result = self.min_report
        if hasattr(self, "psi_0"):
            result["psi_0"] = self.psi_0.to_json()
        if hasattr(self, "best_solution"):
            result["best_solution_name"] = self.best_solution["name"]
        return result

    def save_report(self, folder=None):
        job_name = os.environ.get("JOB_NAME", "default")
        run_id = int(os.environ.get("RUN_ID", 0))
        folder = Path(folder if folder is not None else ".") / job_name
        folder.mkdir(parents=True, exist_ok=True)
        with open(folder / f"data_{run_id}.json", "w+") as f:
            json.dump(self.report, f, indent=2)

    # ------------------------------------------------------------ drivers (lines 208-365)
    def find_low_laying_state(self, psi_init, operator, excluded_states=(), avoid_correlations=0, level_repulsion=1,
                              imaginary_time_evolution=True, min_steps=200, max_steps=1500, **algorithm_kwargs):
        self.psi_0 = psi_init
        self.psi_init = psi_init
        self.psi = +psi_init
        self.identity = identity_operator(self.gpu)
        self.operator = _as_operator(operator, self.gpu)
        self.opt_operator = affine_operator(self.operator, -0.1, 1.0, self.gpu)
        self.excluded_states = list(excluded_states)
        self.energy_history = []
        self.delta_energy_history = []
        self.mode = modes.groundstate
        self.imaginary_time_evolution = imaginary_time_evolution
        self.excluded_states_distances = []
        self.avoid_correlations = avoid_correlations
        self.level_repulsion = level_repulsion
        self.algorithm_kwargs = algorithm_kwargs

        algorithm_iter = self.get_gradient_descent_algorithm("padam")
        steps = min_steps
        list(islice(algorithm_iter, min_steps))
        while steps < max_steps and not self.has_converged(self.energy_history):
            list(islice(algorithm_iter, 100))
            steps += 100
        return self.psi

    def do_the_gradient_descent(self, algorithm_name="padam"):
        self.psi = +self.psi_init
        algorithm = self.get_gradient_descent_algorithm(algorithm_name)
        self.gradient_prefactor = 1
        num_steps = 400
        self.distance_history = list(islice(algorithm, num_steps))
        while num_steps <= 1000 and self.is_descending:
            self.distance_history += list(islice(algorithm, 200))
            num_steps += 200
        print(algorithm_name, num_steps, self.smoothed_distance_history[-1])

    def do_preconditioning(self, psi, psi_ref, **algorithm_kwargs):
        print("preconditioning")
        id_op = identity_operator(self.gpu)
        self.distance_history = []

        def gradient(step, params):
            psi.params = params
            if not self.mc:
                psi.normalize(self.spin_ensemble)
            g, distance = self.hilbert_space_distance.gradient(psi_ref, psi, id_op, True, self.spin_ensemble, 1.0)
            self.distance_history.append(distance)
            return g, distance

        psi = +psi
        _gd.padam(psi.params, 75, gradient, **algorithm_kwargs)
        return psi

    def optimize_for(self, psi_0, psi_init, operator, is_unitary=False, regularization=None, distance_0=0,
                     methods=("padam", "padam_0"), **algorithm_kwargs):
        self.psi_0 = psi_0
        self.psi_init = psi_init
        if "eta" not in algorithm_kwargs:
            algorithm_kwargs["eta"] = 1e-1
        self.algorithm_kwargs = algorithm_kwargs
        self.distance_0 = distance_0
        self.regularization = regularization
        self.operator = _as_operator(operator, self.gpu)
        self.is_unitary = is_unitary
        self.mode = modes.unitary_evolution

        for eta in [algorithm_kwargs["eta"], 0.1, 0.01]:
            algorithm_kwargs["eta"] = eta
            print(f"eta = {eta}")
            self.solutions = []
            for method_name in methods:
                if method_name.endswith("0"):
                    method = method_name[:-2]
                    self.regularization = L2Regularization(1e-2, 0.96)
                else:
                    method = method_name
                    self.regularization = None
                self.psi_init = psi_init
                self.do_the_gradient_descent(method)
                if not math.isnan(self.distance_history[-1]) and self.smoothed_distance_history[-1] < 0.4:
                    self.solutions.append(dict(name=method_name, distance_history=self.distance_history, psi=+self.psi))
            if self.solutions:
                break
            print("Try again using different learning rate")

        if not self.solutions:
            raise DidNotConverge("did not converge")

        self.best_solution = min(self.solutions, key=lambda solution: solution["distance_history"][-1])
        self.distance_history = self.best_solution["distance_history"]
        self.psi = self.best_solution["psi"]
        print(self.smoothed_distance_history[-1])
        if not self.mc:
            self.psi.normalize(self.spin_ensemble)
        return self.psi

    def intelligent_optimize_for(self, psi, operator, exp=True, regularization=None, reversed_order=False):
        raise NotImplementedError("intelligent_optimize_for needs QuantumExpression's PauliExpression algebra, "
                                  "which is not vendored with the reference")
