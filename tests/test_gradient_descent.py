"""The optimisers of rbmongpu_amd.gradient_descent (interface of the
reference's pyRBMonGPU/gradient_descent.py) on f(p) = sum |p - p0|^2, whose
gradient in the d/dRe + i d/dIm convention is 2 (p - p0): every method reaches
p0; adam's first two steps under a constant gradient pin its moment division
by 1 - beta (gradient_descent.py:147-148) against the textbook 1 - beta^t;
padam yields the cost; adamax applies its params_filter."""
import numpy as np
import pytest

from rbmongpu_amd import gradient_descent as GD

P0 = np.array([0.3 - 0.2j, -1.0 + 0.5j, 0.05 + 0.0j])


def grad(step, p):
    return 2.0 * (p - P0)


@pytest.mark.parametrize("method,kwargs,steps,tol", [
    ("vanilla", dict(eta=0.1), 200, 1e-8),
    ("momentum", dict(eta=0.05, gamma=0.5), 300, 1e-8),
    ("nag", dict(eta=0.05, gamma=0.5), 300, 1e-8),
    ("adagrad", dict(eta=0.5, epsilon=1.0), 3000, 1e-3),
    ("rmsprop", dict(eta=0.05, gamma=0.9, epsilon=1.0), 3000, 1e-6),   # divides by G + eps, not its root
    ("adam", dict(eta=0.01), 3000, 2e-2),
    ("adamax", dict(eta=0.01), 3000, 2e-2),
])
def test_methods_reach_the_minimum(method, kwargs, steps, tol):
    p = np.zeros_like(P0)
    hist = getattr(GD, method)(p.copy(), steps, grad, **kwargs)
    assert len(hist) == steps
    assert np.abs(hist[-1] - P0).max() < tol


def test_adadelta_decreases_the_cost():
    hist = GD.adadelta(np.zeros_like(P0), 2000, grad, gamma=0.9, epsilon=1e-4)
    costs = [np.sum(np.abs(h - P0) ** 2) for h in hist]
    assert costs[-1] < 0.1 * np.sum(np.abs(P0) ** 2)


def test_adam_moment_correction_is_the_references():
    eta, b1, b2 = 0.1, 0.9, 0.999
    g = np.array([0.5 + 0.25j])
    gen = GD.adam_generator(np.zeros(1, complex), lambda s, p: g, beta1=b1, beta2=b2, eta=eta, epsilon=0.0)
    p1, p2 = next(gen), next(gen)
    x = GD.unpack(g)
    assert np.allclose(GD.unpack(p1), -eta * np.sign(x))
    step2 = eta * (1 + b1) * x / (np.sqrt(1 + b2) * np.abs(x))
    assert np.allclose(GD.unpack(p2 - p1), -step2)
    textbook = eta * x / np.abs(x)       # m/(1-b1^2) over sqrt(v/(1-b2^2)) for a constant gradient
    assert not np.allclose(step2, textbook)


def test_padam_yields_cost_and_adamax_filters():
    costs = GD.padam(np.zeros_like(P0), 50, lambda s, p: (grad(s, p), float(np.sum(np.abs(p - P0) ** 2))), eta=0.05)
    assert len(costs) == 50 and costs[-1] < costs[0]
    hist = GD.adamax(np.zeros_like(P0), 5, grad, eta=0.1, params_filter=lambda s, p: p.real + 0j)
    assert all(np.all(h.imag == 0) for h in hist)


def test_pack_unpack_roundtrip():
    assert np.array_equal(GD.pack(GD.unpack(P0)), P0)
