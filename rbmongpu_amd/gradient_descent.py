"""First-order optimisers over complex parameter vectors with the interface of
the reference's driver module (pyRBMonGPU/gradient_descent.py): a generator
per method that calls ``gradient_function(step, params)`` and yields the new
parameters, and a list-returning wrapper that runs ``num_steps`` of it.

Conventions kept from the reference because callers depend on them:

* a gradient is ``d/dRe p + i d/dIm p`` as returned by
  ``ExpectationValue.gradient``; the adaptive methods work on the real vector
  ``[Re p, Im p]`` (``unpack`` / ``pack``, gradient_descent.py:5-10);
* ``adam`` divides the moments by ``1 - beta`` rather than ``1 - beta^t``
  (gradient_descent.py:147-148), ``padam`` uses the running maximum of the
  second moment to the power ``p`` and yields the cost its gradient function
  returns (167-187), ``adamax`` accepts a ``params_filter`` (190-220);
* ``eta`` may be a function of the step where the reference allows it.

``sr_generator`` (no reference counterpart) takes stochastic-reconfiguration
steps with the GPU solve of ``rbmongpu_amd.sr_update``.

The optimisers are plain numpy on the host: the parameters are a few
thousand numbers and every step is dominated by the gradient evaluation on
the GPU.
"""
from itertools import islice

import numpy as np

__all__ = ["pack", "unpack", "vanilla_generator", "vanilla", "momentum", "nag_generator", "nag", "adagrad", "adadelta",
           "rmsprop_generator", "rmsprop", "adam_generator", "adam", "padam_generator", "padam", "adamax_generator",
           "adamax", "sr_generator"]


def pack(array):
    """[Re p, Im p] -> p."""
    half = len(array) // 2
    return array[:half] + 1j * array[half:]


def unpack(array):
    """p -> [Re p, Im p]."""
    return np.concatenate([np.real(array), np.imag(array)])


def _rate(eta, step):
    return eta(step) if callable(eta) else eta


def _run(gen, num_steps):
    return list(islice(gen, num_steps))


# ---------------------------------------------------------------- plain steps
def vanilla_generator(params, gradient_function, eta=1e-3):
    step = 0
    while True:
        params -= eta * gradient_function(step, params)
        yield +params
        step += 1


def vanilla(params, num_steps, gradient_function, eta):
    return _run(vanilla_generator(params, gradient_function, eta), num_steps)


def momentum(params, num_steps, gradient_function, eta, gamma=0.9):
    out, velocity = [], 0
    for step in range(num_steps):
        velocity = gamma * velocity + eta * gradient_function(step, params)
        params -= velocity
        out.append(+params)
    return out


def nag_generator(params, gradient_function, eta=1e-1, gamma=0.9, **kwargs):
    """Nesterov: the gradient is taken at the look-ahead point."""
    velocity, step = 0, 0
    while True:
        velocity = gamma * velocity + eta * gradient_function(step, params - gamma * velocity)
        params -= velocity
        yield +params
        step += 1


def nag(params, num_steps, gradient_function, eta=1e-1, gamma=0.9):
    return _run(nag_generator(params, gradient_function, eta, gamma), num_steps)


# ---------------------------------------------------------------- adaptive steps on [Re p, Im p]
def adagrad(params, num_steps, gradient_function, eta, epsilon):
    x, G, out = unpack(params), 0, []
    for step in range(num_steps):
        g = unpack(gradient_function(step, pack(x)))
        G = G + g * g
        x = x - eta / (G + epsilon) * g
        out.append(pack(x))
    return out


def adadelta(params, num_steps, gradient_function, gamma, epsilon):
    x, out = unpack(params), []
    delta, acc_delta, acc_grad = 0, 0, 0
    for step in range(num_steps):
        g = unpack(gradient_function(step, pack(x)))
        acc_delta = gamma * acc_delta + (1 - gamma) * delta ** 2
        acc_grad = gamma * acc_grad + (1 - gamma) * g * g
        delta = -np.sqrt(acc_delta + epsilon) / np.sqrt(acc_grad + epsilon) * g
        x = x + delta
        out.append(pack(x))
    return out


def rmsprop_generator(params, gradient_function, eta, gamma=0.9, epsilon=1e-3, **kwargs):
    x, G, step = unpack(params), 0, 0
    while True:
        g = unpack(gradient_function(step, pack(x)))
        G = gamma * G + (1 - gamma) * g * g
        x = x - _rate(eta, step) / (G + epsilon) * g
        yield pack(x)
        step += 1


def rmsprop(params, num_steps, gradient_function, eta, gamma, epsilon):
    return _run(rmsprop_generator(params, gradient_function, eta, gamma, epsilon), num_steps)


def adam_generator(params, gradient_function, beta1=0.9, beta2=0.999, eta=4e-3, epsilon=1e-3, **kwargs):
    x, m, v, step = unpack(params), 0, 0, 0
    while True:
        g = unpack(gradient_function(step, pack(x)))
        m = beta1 * m + (1 - beta1) * g
        v = beta2 * v + (1 - beta2) * g * g
        # moments divided by 1 - beta, not 1 - beta^t, as in the reference
        x = x - eta / (np.sqrt(v / (1 - beta2)) + epsilon) * (m / (1 - beta1))
        yield pack(x)
        step += 1


def adam(params, num_steps, gradient_function, beta1=0.9, beta2=0.999, eta=4e-3, epsilon=1e-3):
    return _run(adam_generator(params, gradient_function, beta1, beta2, eta, epsilon), num_steps)


def padam_generator(params, gradient_function, beta1=0.9, beta2=0.999, eta=1e-1, epsilon=1e-4, p=1 / 4, **kwargs):
    """gradient_function returns (gradient, cost); the generator yields the
    cost of each step."""
    x, m, v, vmax, step = unpack(params), 0, 0, 0, 0
    while True:
        grad, cost = gradient_function(step, pack(x))
        g = unpack(grad)
        m = beta1 * m + (1 - beta1) * g
        v = beta2 * v + (1 - beta2) * g * g
        vmax = np.maximum(vmax, v)
        x = x - _rate(eta, step) / (vmax ** p + epsilon) * m
        yield cost
        step += 1


def padam(params, num_steps, gradient_function, beta1=0.9, beta2=0.999, eta=1e-1, epsilon=1e-4, p=1 / 4):
    return _run(padam_generator(params, gradient_function, beta1, beta2, eta, epsilon, p), num_steps)


def adamax_generator(params, gradient_function, beta1=0.9, beta2=0.999, eta=1e-3, epsilon=1e-4, params_filter=None,
                     **kwargs):
    x, m, u, step = unpack(params), 0, 0, 0
    while True:
        g = unpack(gradient_function(step, pack(x)))
        m = beta1 * m + (1 - beta1) * g
        u = np.maximum(beta2 * u, np.abs(g))
        x = x - _rate(eta, step) / (u + epsilon) * (m / (1 - beta1))
        if params_filter is not None:
            x = unpack(params_filter(step, pack(x)))
        yield pack(x)
        step += 1


def adamax(params, num_steps, gradient_function, beta1=0.9, beta2=0.999, eta=1e-3, epsilon=1e-4, params_filter=None):
    return _run(adamax_generator(params, gradient_function, beta1, beta2, eta, epsilon, params_filter), num_steps)


# ---------------------------------------------------------------- stochastic reconfiguration
def sr_generator(psi, operator, spin_ensemble, eta=5e-2, epsilon=1e-4):
    """Stochastic-reconfiguration steps p -= eta (S + epsilon I)^-1 g on the GPU
    (centred S on the matrix cores, FP64 Cholesky solve); yields the energy
    measured on the samples of each step.  psi is updated in place."""
    from . import sr_update
    step = 0
    while True:
        update, energy = sr_update(psi, operator, spin_ensemble, epsilon)
        psi.params = psi.params - _rate(eta, step) * update
        yield energy
        step += 1
