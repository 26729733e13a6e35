"""Writes the golden fixtures under tests/golden/ from the FP64 oracle.

The reference ships no golden vectors (its tests draw unseeded random
networks), so the fixtures are the oracle's outputs on seeded inputs; the
oracle itself is pinned by tests/test_oracle.py.  Re-run with
`python tests/golden/make_golden.py` after an intentional oracle change.

c1_tfim_rbm.npz: 1-D TFIM N=10 (J=h=1, periodic), RBM M=20 with parameters
  drawn i.i.d. N(0, 0.01^2) (seed 1235), with and without the visible-bias
  extension: psi over all 1024 states, E_loc per state, the normalised
  energy, gradient and fluctuation under ExactSummation, and the
  Hilbert-space distance (+ gradient) to a copy with W += N(0, 1e-3^2).
deep_small.npz: PsiDeep N=6, layers [12, 6] (connectivity [3, 12]) on a
  random operator: log psi, E_loc, O_k at three configurations, gradient.
c2_tfim_rbm.npz: 1-D TFIM N=20 (J=h=1, periodic), RBM M=80 with parameters
  drawn i.i.d. N(0, 0.01^2) (seed 1236): the normalising prefactor, energy
  and the 1720-entry gradient under ExactSummation over all 2^20 states
  (the centred estimator, equal to the reference's E_k form for this
  Hermitian operator), and the Hilbert-space distance (identity operator,
  unitary form, nu = 1) and its gradient to a copy with W += N(0, 1e-3^2).
  `python tests/golden/make_golden.py c2` rewrites only this file (about two
  minutes of oracle time).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from oracle import oracle as O  # noqa: E402
from rbmongpu_amd import hamiltonians as HAM  # noqa: E402
from helpers import basis_words, deep_layers, random_operator  # noqa: E402


def c1(visible):
    rng = np.random.default_rng(1235)
    N, M = 10, 20
    cn = lambda *s: 0.01 * (rng.normal(size=s) + 1j * rng.normal(size=s))  # noqa: E731
    a = cn(N) if visible else None
    b, W = cn(M), cn(N, M)
    W2 = W + 1e-3 * (rng.normal(size=(N, M)) + 1j * rng.normal(size=(N, M)))
    tab = HAM.tfim_1d(N, 1.0, 1.0)
    op = O.Operator(*tab)
    psi = O.RBM(b, W, a)
    vec = O.psi_vector(psi)
    el = O.local_energy(psi, op, basis_words(N))
    psi.prefactor = 1.0 / O.psi_norm(psi, O.Exact(N))
    grad, energy = O.gradient(psi, op, O.Exact(N), mode=0)
    fluct, _ = O.fluctuation(psi, op, O.Exact(N))
    psi2 = O.RBM(b, W2, a)
    hs_grad, hs_d = O.hs_distance(psi, psi2, op, True, O.Exact(N), want_grad=True, nu=1.0)
    pre = "vis_" if visible else ""
    out = {pre + "b": b, pre + "W": W, pre + "W2": W2, pre + "psi": vec, pre + "e_loc": el,
           pre + "prefactor": np.array(psi.prefactor), pre + "energy": np.array(energy), pre + "gradient": grad,
           pre + "fluctuation": np.array(fluct), pre + "hs_distance": np.array(hs_d), pre + "hs_gradient": hs_grad}
    if visible:
        out["vis_a"] = a
    return out


def deep_small():
    N, M, C = 6, [12, 6], [3, 12]
    biases, conns, weights, final = deep_layers(N, M, C, seed=77)
    tab = random_operator(N, 8, 3, seed=78, hermitian_coeffs=True)
    psi = O.Deep(N, biases, conns, weights, final)
    words = basis_words(N)
    lp = O.log_psi(psi, words)
    el = O.local_energy(psi, O.Operator(*tab), words)
    configs = np.array([0, 21, 63])
    ok = np.stack([O.o_k(psi, w) for w in O.spins_words(configs)])
    psi.prefactor = 1.0 / O.psi_norm(psi, O.Exact(N))
    grad, energy = O.gradient(psi, O.Operator(*tab), O.Exact(N), mode=0)
    out = {"coeffs": tab[0], "types": tab[1], "idx": tab[2], "final": final, "log_psi": lp, "e_loc": el,
           "ok_configs": configs, "o_k": ok, "prefactor": np.array(psi.prefactor), "gradient": grad,
           "energy": np.array(energy)}
    for l, (b, c, w) in enumerate(zip(biases, conns, weights)):
        out[f"b{l}"], out[f"c{l}"], out[f"w{l}"] = b, c, w
    return out


def c2():
    rng = np.random.default_rng(1236)
    N, M = 20, 80
    cn = lambda *s: 0.01 * (rng.normal(size=s) + 1j * rng.normal(size=s))  # noqa: E731
    b, W = cn(M), cn(N, M)
    W2 = W + 1e-3 * (rng.normal(size=(N, M)) + 1j * rng.normal(size=(N, M)))
    op = O.Operator(*HAM.tfim_1d(N, 1.0, 1.0))
    psi = O.RBM(b, W)
    psi.prefactor = 1.0 / O.psi_norm(psi, O.Exact(N))
    grad, energy = O.gradient(psi, op, O.Exact(N), mode=1)
    ident = O.Operator(np.ones(1, complex), np.zeros((1, 1), np.int32), -np.ones((1, 1), np.int32))
    hs_grad, hs_d = O.hs_distance(psi, O.RBM(b, W2), ident, True, O.Exact(N), want_grad=True, nu=1.0)
    return {"b": b, "W": W, "W2": W2, "prefactor": np.array(psi.prefactor), "energy": np.array(energy),
            "gradient": grad, "hs_distance": np.array(hs_d), "hs_gradient": hs_grad}


if __name__ == "__main__":
    if sys.argv[1:] != ["c2"]:
        np.savez_compressed(os.path.join(HERE, "c1_tfim_rbm.npz"), **c1(False), **c1(True))
        np.savez_compressed(os.path.join(HERE, "deep_small.npz"), **deep_small())
    np.savez_compressed(os.path.join(HERE, "c2_tfim_rbm.npz"), **c2())
    print("wrote", os.listdir(HERE))
