"""Training driver over this package's GPU calls.

Public interface: that of the reference's ``LearningByGradientDescent``
(pyRBMonGPU/LearningByGradientDescent.py:36-365) and ``L2Regularization``
(pyRBMonGPU/L2Regularization.py:1-13) -- the class and method names, their
arguments, the attributes callers read afterwards (``distance_history``,
``energy_history``, ``delta_energy_history``, ``excluded_states_distances``,
``solutions``, ``best_solution``, ``psi``) and the report keys.

Organisation:

* ``Criteria`` holds every number that decides when a run of the optimiser is
  still descending, is too noisy, learns poorly or has levelled off, with the
  values the reference uses;
* ``CostHistory`` answers those questions for one cost history;
* an *objective* turns a parameter vector into ``(gradient, cost)`` on the GPU:
  ``_TrackState`` (learn psi ~ U psi_0, or any operator applied to psi_0) and
  ``_LowerEnergy`` (imaginary-time steps with ``1 - 0.1 H``, or plain energy
  gradients, with repulsion from excluded states);
* ``LearningByGradientDescent`` installs an objective, runs the optimiser in
  stages and keeps the records.

Operators are this package's ``Operator`` objects or ``(coefficients,
pauli_types, pauli_indices)`` tables.  The reference builds them from
QuantumExpression's PauliExpression, which is not vendored; the two operators
the driver forms itself, the identity and ``shift + scale * H``, are built on
the tables (``identity_operator``, ``affine_operator``).

Where the reference cannot run as written, this module does the following
instead:

* ``find_low_laying_state`` runs padam on the ground-state objective (the
  reference asks ``get_gradient_descent_algorithm`` for an unnamed algorithm
  and indexes the generator, lines 226-227);
* ``avoid_correlations`` needs HilbertSpaceCorrelations, whose import the
  reference leaves commented out (line 10): a non-zero value raises
  NotImplementedError;
* ``intelligent_optimize_for`` needs PauliExpression term algebra: it raises
  NotImplementedError;
* ``save_report`` writes under ``folder`` (default: the working directory),
  and the reports store the operator's tables;
* the ground-state and preconditioning distance gradients, which the reference
  calls without the binding's required ``nu`` (lines 130-132, 144-146,
  273-275), use nu = 1, and the preconditioning objective returns the
  ``(gradient, cost)`` pair that padam consumes;
* progress goes to the ``rbmongpu_amd.learning`` logger, not to stdout.
"""
from __future__ import annotations

import enum
import json
import logging
import math
import os
from dataclasses import dataclass
from itertools import islice
from pathlib import Path

import numpy as np
from scipy.ndimage import gaussian_filter1d

from . import gradient_descent as _gd
from ._pyRBMonGPU import ExactSummation, ExpectationValue, HilbertSpaceDistance, Operator
from .json_numpy import NumpyEncoder

_log = logging.getLogger(__name__)


class modes(enum.IntEnum):
    """What the driver optimises (the reference's ``modes`` record)."""
    unitary_evolution = 0
    groundstate = 1


class DidNotConverge(RuntimeError):
    """No learning rate of the schedule produced an accepted solution."""


class UnstableNetwork(RuntimeError):
    pass


class PoorLearning(RuntimeError):
    pass


class L2Regularization:
    """Penalty lambda(step) |Re p|^2 on the first layer with lambda(step) =
    lambda0 gamma^step.  Its gradient 2 lambda(step) Re p ignores entries with
    |Re p| < 1e-3 (pyRBMonGPU/L2Regularization.py:1-13)."""

    def __init__(self, lambda0, gamma):
        self.lambda0 = lambda0
        self.gamma = gamma

    def lambda_(self, step):
        return self.lambda0 * self.gamma ** step

    def gradient(self, step, params):
        re = np.real(np.asarray(params)).astype(float)
        return 2.0 * self.lambda_(step) * np.where(np.abs(re) < 1e-3, 0.0, re)


# ------------------------------------------------------------------ operators
def _tables(operator):
    if isinstance(operator, Operator):
        return operator.coefficients, operator.pauli_types, operator.pauli_indices
    c, t, i = operator
    return np.asarray(c, dtype=complex), np.asarray(t, dtype=np.int32), np.asarray(i, dtype=np.int32)


def identity_operator(gpu=True):
    """The identity: one Pauli string without factors."""
    return Operator(np.ones(1, complex), np.zeros((1, 1), np.int32), -np.ones((1, 1), np.int32), gpu)


def affine_operator(operator, scale, shift, gpu=True):
    """shift * 1 + scale * operator, as one table (the identity string first)."""
    c, t, i = _tables(operator)
    width = t.shape[1]
    coefficients = np.concatenate([[complex(shift)], complex(scale) * np.asarray(c, complex)])
    types = np.concatenate([np.zeros((1, width), np.int32), t])
    indices = np.concatenate([np.full((1, width), -1, np.int32), i])
    return Operator(coefficients, types, indices, gpu)


def _as_operator(operator, gpu):
    return operator if isinstance(operator, Operator) else Operator(*_tables(operator), gpu)


# ------------------------------------------------------------------ cost histories
@dataclass(frozen=True)
class Criteria:
    """Thresholds on a cost history (the values of
    LearningByGradientDescent.py:49-101).  Costs are compared on a log scale
    after Gaussian smoothing ("nearest" edges)."""

    sigma_sampled: float = 5.0   # smoothing width for Monte Carlo costs
    sigma_exact: float = 1.0     # ... and for exact-summation costs
    recent: int = 50             # steps seen by the descent and jitter tests
    min_drop: float = 0.05       # descending: smoothed log cost fell by more than this over `recent`
    max_jitter: float = 0.125    # noisy: std of the raw log cost's step changes above this
    poor_rise: float = 0.2       # a step is poor when its smoothed log cost exceeds the first by this
    max_poor: int = 40           # learning poorly: more poor steps than this
    level_span: int = 100        # levelled off: judged on the last `level_span` values
    level_sigma: float = 5.0
    level_sampled: float = 3e-3  # ... first and last smoothed values closer than this (Monte Carlo)
    level_exact: float = 1e-3    # ... (exact summation)


CRITERIA = Criteria()


class CostHistory:
    """One run's cost per step, with the tests the driver's schedule uses."""

    def __init__(self, values, sampled, criteria=CRITERIA):
        self.values = np.asarray(values, dtype=float)
        self.sampled = sampled
        self.criteria = criteria

    def smoothed(self):
        sigma = self.criteria.sigma_sampled if self.sampled else self.criteria.sigma_exact
        return gaussian_filter1d(self.values, sigma, mode="nearest")

    def log_smoothed(self):
        return np.log(self.smoothed())

    def step_changes(self):
        return np.abs(np.diff(self.smoothed()))

    def descending(self):
        tail = self.log_smoothed()[-self.criteria.recent:]
        return bool(tail[0] - tail[-1] > self.criteria.min_drop)

    def noisy(self):
        jumps = np.diff(np.log(self.values[-self.criteria.recent:]))
        return bool(np.std(jumps) > self.criteria.max_jitter)

    def poor(self):
        curve = self.log_smoothed()
        return bool(np.sum(curve > curve[0] + self.criteria.poor_rise) > self.criteria.max_poor)

    @staticmethod
    def levelled_off(values, sampled, threshold=None, criteria=CRITERIA):
        if not threshold:
            threshold = criteria.level_sampled if sampled else criteria.level_exact
        tail = gaussian_filter1d(np.asarray(values[-criteria.level_span:], dtype=float), criteria.level_sigma,
                                 mode="nearest")
        return bool(abs(tail[-1] - tail[0]) < threshold)


# ------------------------------------------------------------------ objectives
class _TrackState:
    """Cost: the Hilbert-space distance between psi and operator * psi_0 minus
    distance_0, with the gradient's sign following the cost's; gradient_prefactor
    scales it and the L2 penalty acts on the first layer."""

    def __init__(self, driver):
        self.d = driver

    def __call__(self, step, params):
        d = self.d
        psi = d._load(params)
        gradient, cost = d.hilbert_space_distance.gradient(d.psi_0, psi, d.operator, d.is_unitary, d.spin_ensemble,
                                                           0.0)
        cost -= d.distance_0
        gradient = (-1.0 if cost < 0 else 1.0) * d.gradient_prefactor * gradient
        if d.regularization is not None:
            layer = psi.first_layer_params_slice
            gradient[layer] += d.regularization.gradient(step, psi.params[layer])
        return gradient, cost


class _LowerEnergy:
    """Cost: the distance between psi and (1 - 0.1 H) psi (imaginary-time
    step), or 0 with the plain energy gradient.  Every call records the energy
    and its fluctuation, and pushes psi away from each excluded state with
    weight level_repulsion * clip(log distance, 1, 10)."""

    def __init__(self, driver):
        self.d = driver

    def __call__(self, step, params):
        d = self.d
        psi = d._load(params)
        if d.imaginary_time_evolution:
            gradient, cost = d.hilbert_space_distance.gradient(psi, psi, d.opt_operator, True, d.spin_ensemble, 1.0)
        else:
            gradient, _ = d.expectation_value.gradient(psi, d.operator, d.spin_ensemble)
            cost = 0.0
        sigma, energy = d.expectation_value.fluctuation(psi, d.operator, d.spin_ensemble)
        d.delta_energy_history.append(sigma)
        d.energy_history.append(energy.real)
        overlaps = []
        for other in d.excluded_states:
            push, dist = d.hilbert_space_distance.gradient(other, psi, d.identity, True, d.spin_ensemble, 1.0)
            overlaps.append(dist)
            weight = d.level_repulsion * min(10.0, max(1.0, math.log(dist.real)))
            gradient = gradient - weight * push
        d.excluded_states_distances.append(overlaps)
        return gradient, cost


# ------------------------------------------------------------------ driver
class LearningByGradientDescent:
    """Optimises a wavefunction on a spin ensemble by gradient descent on the
    GPU (interface of LearningByGradientDescent.py:36-365; differences in the
    module docstring)."""

    #: steps of one descent: a first stage, then further stages while the
    #: smoothed cost is still descending
    STAGES = (400, 200, 200, 200, 200)
    #: learning rates tried by optimize_for after the caller's own
    ETA_FALLBACKS = (0.1, 0.01)
    #: a run is a solution when its smoothed final cost is below this
    ACCEPT = 0.4

    def __init__(self, psi, spin_ensemble):
        self.gpu = psi.gpu
        # sampled costs (any Monte Carlo ensemble, the learned one included)
        # are smoothed more and judged with a looser plateau threshold
        self.mc = not isinstance(spin_ensemble, ExactSummation)
        self.spin_ensemble = spin_ensemble
        self.num_params = psi.num_params
        self.hilbert_space_distance = HilbertSpaceDistance(psi.N, self.num_params, psi.gpu)
        self.expectation_value = ExpectationValue(self.gpu)
        self.regularization = None
        self.algorithm_kwargs = {}
        self.gradient_prefactor = 1
        self._objective = None

    # -------------------------------------------------------------- history views
    def _history(self):
        return CostHistory(self.distance_history, self.mc)

    @property
    def smoothed_distance_history(self):
        return self._history().smoothed()

    def smoothed(self, x):
        return gaussian_filter1d(np.asarray(x[-CRITERIA.level_span:], dtype=float), CRITERIA.level_sigma,
                                 mode="nearest")

    @property
    def slopes(self):
        return self._history().step_changes()

    @property
    def graph(self):
        return self._history().log_smoothed()

    @property
    def is_highly_fluctuating(self):
        return self._history().noisy()

    @property
    def is_descending(self):
        return self._history().descending()

    @property
    def is_learning_poorly(self):
        return self._history().poor()

    def has_converged(self, x, threshold=None):
        return CostHistory.levelled_off(x, self.mc, threshold)

    def push_params(self, params):
        """Keep the last three parameter vectors (a ring)."""
        ring = getattr(self, "last_params_stack", None)
        if ring is None:
            ring = self.last_params_stack = np.zeros((3, self.psi.num_params), dtype=complex)
            self.last_params_idx = 0
        ring[self.last_params_idx] = params
        self.last_params_idx = (self.last_params_idx + 1) % len(ring)

    @property
    def last_params_avg(self):
        return self.last_params_stack.mean(axis=0)

    # -------------------------------------------------------------- one step
    def _load(self, params):
        """Install params in self.psi (normalised on an exact summation)."""
        self.psi.params = params
        if not self.mc:
            self.psi.normalize(self.spin_ensemble)
        return self.psi

    def set_params_and_return_gradient(self, step, params):
        if getattr(self, "avoid_correlations", 0):
            raise NotImplementedError("avoid_correlations needs HilbertSpaceCorrelations, which the reference does "
                                      "not ship")
        return self._objective(step, params)

    def get_gradient_descent_algorithm(self, name):
        factory = getattr(_gd, name + "_generator")
        return factory(self.psi_init.params, self.set_params_and_return_gradient, **self.algorithm_kwargs)

    # -------------------------------------------------------------- reports
    @property
    def min_report(self):
        c, t, i = _tables(self.operator)
        report = {"operator": dict(coefficients=c, pauli_types=t, pauli_indices=i)}
        for key in ("is_unitary", "distance_history", "energy_history", "delta_energy_history",
                    "excluded_states_distances", "algorithm_kwargs"):
            if hasattr(self, key):
                report[key] = getattr(self, key)
        return json.loads(json.dumps(report, cls=NumpyEncoder))

    @property
    def report(self):
        report = self.min_report
        if hasattr(self, "psi_0"):
            report["psi_0"] = self.psi_0.to_json()
        if hasattr(self, "best_solution"):
            report["best_solution_name"] = self.best_solution["name"]
        return report

    def save_report(self, folder=None):
        target = Path(folder if folder is not None else ".") / os.environ.get("JOB_NAME", "default")
        target.mkdir(parents=True, exist_ok=True)
        path = target / "data_{}.json".format(int(os.environ.get("RUN_ID", 0)))
        path.write_text(json.dumps(self.report, indent=2))
        return path

    # -------------------------------------------------------------- drivers
    def find_low_laying_state(self, psi_init, operator, excluded_states=(), avoid_correlations=0, level_repulsion=1,
                              imaginary_time_evolution=True, min_steps=200, max_steps=1500, **algorithm_kwargs):
        """padam on the energy objective: min_steps, then 100 at a time until
        the energy history levels off or max_steps is reached."""
        self.psi_0 = self.psi_init = psi_init
        self.psi = +psi_init
        self.operator = _as_operator(operator, self.gpu)
        self.opt_operator = affine_operator(self.operator, -0.1, 1.0, self.gpu)
        self.identity = identity_operator(self.gpu)
        self.excluded_states = list(excluded_states)
        self.level_repulsion = level_repulsion
        self.avoid_correlations = avoid_correlations
        self.imaginary_time_evolution = imaginary_time_evolution
        self.energy_history, self.delta_energy_history, self.excluded_states_distances = [], [], []
        self.algorithm_kwargs = algorithm_kwargs
        self.mode = modes.groundstate
        self._objective = _LowerEnergy(self)

        steps = self.get_gradient_descent_algorithm("padam")
        taken = len(list(islice(steps, min_steps)))
        while taken < max_steps and not self.has_converged(self.energy_history):
            taken += len(list(islice(steps, 100)))
        return self.psi

    def do_the_gradient_descent(self, algorithm_name="padam"):
        """One descent from psi_init: the first stage of STAGES, then each
        further stage while the smoothed cost keeps descending."""
        self.psi = +self.psi_init
        self.gradient_prefactor = 1
        steps = self.get_gradient_descent_algorithm(algorithm_name)
        self.distance_history = list(islice(steps, self.STAGES[0]))
        for length in self.STAGES[1:]:
            if not self.is_descending:
                break
            self.distance_history.extend(islice(steps, length))
        _log.info("%s: %d steps, smoothed cost %.4g", algorithm_name, len(self.distance_history),
                  self.smoothed_distance_history[-1])

    def do_preconditioning(self, psi, psi_ref, **algorithm_kwargs):
        """75 padam steps that pull a copy of psi towards psi_ref (identity
        operator, unitary form); returns the copy."""
        _log.info("preconditioning towards the reference state")
        unit = identity_operator(self.gpu)
        trained = +psi
        self.distance_history = []

        def objective(step, params):
            trained.params = params
            if not self.mc:
                trained.normalize(self.spin_ensemble)
            gradient, cost = self.hilbert_space_distance.gradient(psi_ref, trained, unit, True, self.spin_ensemble, 1.0)
            self.distance_history.append(cost)
            return gradient, cost

        _gd.padam(trained.params, 75, objective, **algorithm_kwargs)
        return trained

    def _descent_run(self, psi_init, method_name):
        """One descent of `method_name` ("<optimiser>" or "<optimiser>_0", the
        latter with the L2 penalty); a solution record or None."""
        base, _, suffix = method_name.rpartition("_")
        regularised = base != "" and suffix == "0"
        self.regularization = L2Regularization(1e-2, 0.96) if regularised else None
        self.psi_init = psi_init
        self.do_the_gradient_descent(base if regularised else method_name)
        final = self.distance_history[-1]
        if math.isnan(final) or self.smoothed_distance_history[-1] >= self.ACCEPT:
            return None
        return dict(name=method_name, distance_history=self.distance_history, psi=+self.psi)

    def optimize_for(self, psi_0, psi_init, operator, is_unitary=False, regularization=None, distance_0=0,
                     methods=("padam", "padam_0"), **algorithm_kwargs):
        """Learn psi ~ operator psi_0 starting from psi_init.  Every method is
        run at the caller's learning rate (default 0.1); if none is accepted,
        at each rate of ETA_FALLBACKS in turn.  The accepted run with the
        lowest final cost wins; DidNotConverge if there is none."""
        self.psi_0 = psi_0
        self.psi_init = psi_init
        algorithm_kwargs.setdefault("eta", 1e-1)
        self.algorithm_kwargs = algorithm_kwargs
        self.distance_0 = distance_0
        self.regularization = regularization
        self.operator = _as_operator(operator, self.gpu)
        self.is_unitary = is_unitary
        self.mode = modes.unitary_evolution
        self._objective = _TrackState(self)

        for eta in (algorithm_kwargs["eta"],) + self.ETA_FALLBACKS:
            algorithm_kwargs["eta"] = eta
            _log.info("learning rate %g", eta)
            self.solutions = [s for s in (self._descent_run(psi_init, m) for m in methods) if s is not None]
            if self.solutions:
                break
            _log.info("no accepted run at learning rate %g", eta)
        else:
            raise DidNotConverge("no method converged at any learning rate")

        self.best_solution = min(self.solutions, key=lambda s: s["distance_history"][-1])
        self.distance_history = self.best_solution["distance_history"]
        self.psi = self.best_solution["psi"]
        _log.info("best: %s, smoothed cost %.4g", self.best_solution["name"], self.smoothed_distance_history[-1])
        if not self.mc:
            self.psi.normalize(self.spin_ensemble)
        return self.psi

    def intelligent_optimize_for(self, psi, operator, exp=True, regularization=None, reversed_order=False):
        raise NotImplementedError("intelligent_optimize_for needs QuantumExpression's PauliExpression algebra, "
                                  "which is not vendored with the reference")
