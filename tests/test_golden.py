"""The oracle reproduces the committed golden fixtures (regression guard of
the checker itself); the GPU comparisons against the same fixtures live in
test_gpu_rbm.py / test_gpu_deep.py."""
import os

import numpy as np
import pytest

from helpers import O, basis_words, rel_err
from rbmongpu_amd import hamiltonians as HAM

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("pre", ["", "vis_"])
def test_oracle_reproduces_c1_fixture(pre):
    d = np.load(os.path.join(GOLDEN, "c1_tfim_rbm.npz"))
    a = d["vis_a"] if pre else None
    psi = O.RBM(d[pre + "b"], d[pre + "W"], a)
    op = O.Operator(*HAM.tfim_1d(10, 1.0, 1.0))
    assert rel_err(O.psi_vector(psi), d[pre + "psi"]) < 1e-14
    assert rel_err(O.local_energy(psi, op, basis_words(10)), d[pre + "e_loc"]) < 1e-14
    psi.prefactor = float(d[pre + "prefactor"])
    g, E = O.gradient(psi, op, O.Exact(10), mode=0)
    assert abs(E - d[pre + "energy"]) < 1e-12
    assert rel_err(g, d[pre + "gradient"]) < 1e-12
    psi2 = O.RBM(d[pre + "b"], d[pre + "W2"], a)
    hg, hd = O.hs_distance(psi, psi2, op, True, O.Exact(10), want_grad=True, nu=1.0)
    assert abs(hd - d[pre + "hs_distance"]) < 1e-13
    assert rel_err(hg, d[pre + "hs_gradient"]) < 1e-11


def test_oracle_reproduces_deep_fixture():
    d = np.load(os.path.join(GOLDEN, "deep_small.npz"))
    L = 2
    psi = O.Deep(6, [d[f"b{l}"] for l in range(L)], [d[f"c{l}"] for l in range(L)], [d[f"w{l}"] for l in range(L)],
                 d["final"])
    assert rel_err(O.log_psi(psi, basis_words(6)), d["log_psi"]) < 1e-14
    op = O.Operator(d["coeffs"], d["types"], d["idx"])
    assert rel_err(O.local_energy(psi, op, basis_words(6)), d["e_loc"]) < 1e-13
    ok = np.stack([O.o_k(psi, w) for w in O.spins_words(d["ok_configs"])])
    assert rel_err(ok, d["o_k"]) < 1e-14


def test_oracle_reproduces_c2_fixture_norm_and_distance():
    """The c2 fixture (N = 20, 2^20 states): the normalisation and the
    Hilbert-space distance recomputed by the oracle (about 10 s; the energy
    and gradients take minutes and are regenerated with make_golden.py)."""
    d = np.load(os.path.join(GOLDEN, "c2_tfim_rbm.npz"))
    psi = O.RBM(d["b"], d["W"])
    assert abs(1.0 / O.psi_norm(psi, O.Exact(20)) - d["prefactor"]) < 1e-13 * d["prefactor"]
    psi.prefactor = float(d["prefactor"])
    ident = O.Operator(np.ones(1, complex), np.zeros((1, 1), np.int32), -np.ones((1, 1), np.int32))
    hd = O.hs_distance(psi, O.RBM(d["b"], d["W2"]), ident, True, O.Exact(20))
    assert abs(hd - d["hs_distance"]) < 1e-12 * d["hs_distance"]
