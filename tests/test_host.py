"""Host-side tests that need no GPU: the C ABI library loads and exports
every symbol declared in include/rbmhip.h, and the drop-in objects keep the
reference's parameter layouts, table formats and error behaviour."""
import ctypes
import json
import os
import re

import numpy as np
import pytest

from helpers import ROOT, dense_hamiltonian, random_operator

import rbmongpu_amd as R
from rbmongpu_amd import hamiltonians as HAM

LIB = os.path.join(ROOT, "rbmongpu_amd", "librbmhip.so")
HEADER = os.path.join(ROOT, "include", "rbmhip.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rbmhip_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    syms = declared_symbols()
    assert len(syms) >= 40
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing


def test_abi_host_objects_without_gpu():
    lib = ctypes.CDLL(LIB)
    lib.rbmhip_last_error.restype = ctypes.c_char_p
    P = ctypes.c_void_p
    lib.rbmhip_rbm_create.argtypes = [ctypes.c_int, ctypes.c_int, P, P, P, ctypes.c_double, ctypes.c_int, P]
    lib.rbmhip_psi_num_params.argtypes = [P, P]
    lib.rbmhip_psi_destroy.argtypes = [P]
    N, M = 3, 4
    b = np.zeros(2 * M)
    W = np.zeros(2 * N * M)
    h = ctypes.c_void_p()
    assert lib.rbmhip_rbm_create(N, M, b.ctypes.data, W.ctypes.data, None, 1.0, 32, ctypes.byref(h)) == 0
    n = ctypes.c_int64()
    assert lib.rbmhip_psi_num_params(h, ctypes.byref(n)) == 0
    assert n.value == 2 * N + M + N * M
    # a bad precision is rejected with RBMHIP_ERR_ARG and a message
    h2 = ctypes.c_void_p()
    assert lib.rbmhip_rbm_create(N, M, b.ctypes.data, W.ctypes.data, None, 1.0, 16, ctypes.byref(h2)) == -1
    assert b"precision" in lib.rbmhip_last_error()
    assert lib.rbmhip_psi_destroy(h) == 0


def test_compute_without_gpu_fails_loudly():
    if R.device_count() > 0:
        pytest.skip("a GPU is visible")
    psi = R.new_neural_network(3, 4, gpu=True)
    with pytest.raises(RuntimeError, match="no HIP device"):
        psi.vector


def test_gpu_false_is_rejected():
    with pytest.raises(ValueError, match="GPU only"):
        R.Psi(np.zeros(2), np.zeros(2), np.zeros(2, complex), np.zeros((2, 2), complex), 1.0, False, False)
    with pytest.raises(ValueError):
        R.ExactSummation(4, False)
    with pytest.raises(ValueError):
        R.MonteCarloLoop(8, 1, 1, 4, False)


def test_psi_params_layout_roundtrip():
    """[alpha, beta, b, W(row-major)] as in source/quantum_state/Psi.cu:71-92."""
    N, M = 3, 5
    rng = np.random.default_rng(0)
    alpha, beta = rng.normal(size=N), rng.normal(size=N)
    b = rng.normal(size=M) + 1j * rng.normal(size=M)
    W = rng.normal(size=(N, M)) + 1j * rng.normal(size=(N, M))
    psi = R.Psi(alpha, beta, b, W, 2.0, False, True)
    assert psi.num_params == 2 * N + M + N * M
    assert psi.N == N and psi.M == M and psi.num_angles == M and psi.gpu and psi.prefactor == 2.0
    p = psi.params
    assert np.allclose(p[:N], alpha) and np.allclose(p[N:2 * N], beta)
    assert np.allclose(p[2 * N:2 * N + M], b) and np.allclose(p[2 * N + M:], W.ravel())
    p2 = p * (1 + 0.5j)
    psi.params = p2
    assert np.allclose(psi.b, p2[2 * N:2 * N + M]) and np.allclose(psi.W, p2[2 * N + M:].reshape(N, M))
    assert np.allclose(psi.alpha, p2[:N].real)
    with pytest.raises(ValueError):
        psi.params = p[:-1]
    c = +psi
    c.params = p
    assert np.allclose(psi.params[2 * N:], p2[2 * N:]) and np.allclose(psi.params[:2 * N], p2[:2 * N].real)
    assert np.allclose(c.params, p)


def test_psi_visible_bias_extension_layout():
    N, M = 3, 2
    a = np.arange(N) + 1j
    psi = R.Psi(np.zeros(N), np.zeros(N), np.zeros(M, complex), np.zeros((N, M), complex), 1.0, False, True, a=a)
    assert psi.num_params == 2 * N + M + N * M + N
    assert np.allclose(psi.params[-N:], a) and np.allclose(psi.a, a)


def test_psi_deep_params_layout_and_properties():
    N, M, C = 4, [8, 4], [2, 8]
    psi = R.new_deep_neural_network(N, M, C, gpu=True, rng=np.random.default_rng(1))
    assert psi.num_params == 2 * N + (8 + 2 * 8) + (4 + 8 * 4) + 4
    assert [w.shape for w in psi.W] == [(2, 8), (8, 4)]
    assert [c.shape for c in psi.connections] == [(2, 8), (8, 4)]
    p = psi.params
    assert np.allclose(p[2 * N:2 * N + 8], psi.b[0]) and np.allclose(p[2 * N + 8:2 * N + 24], psi.W[0].ravel())
    assert np.allclose(p[-4:], psi.final_weights)
    psi.params = 2 * p
    assert np.allclose(psi.params[2 * N:], 2 * p[2 * N:])
    psi.translational_invariance = True
    assert psi.translational_invariance
    assert psi.first_layer_params_slice == slice(2 * N, 2 * N + 16 + 8)


def test_operator_tables_roundtrip_and_canonical_form():
    coeffs, types, idx = HAM.tfim_1d(6)
    op = R.Operator(coeffs, types, idx, True)
    assert op.num_strings == 12 and op.max_string_length == 3
    assert np.array_equal(op.pauli_types, types) and np.array_equal(op.pauli_indices, idx)
    assert np.allclose(op.coefficients, coeffs)
    assert op.num_flip_masks == (6, 6)        # six ZZ (diagonal), six X masks
    heis = R.Operator(*HAM.heisenberg([(0, 1), (1, 2)]), True)
    assert heis.num_flip_masks == (2, 2)      # XX and YY share a flip mask
    with pytest.raises(ValueError, match="terminator"):
        R.Operator(np.ones(1, complex), np.ones((1, 2), np.int32), np.zeros((1, 2), np.int32), True)


@pytest.mark.parametrize("seed", range(6))
def test_operator_hermiticity_matches_dense(seed):
    N = 4
    coeffs, types, idx = random_operator(N, 6, 3, seed=seed, hermitian_coeffs=seed % 2 == 0)
    H = dense_hamiltonian(N, coeffs, types, idx)
    herm = np.allclose(H, H.conj().T, atol=1e-12)
    assert R.Operator(coeffs, types, idx, True).is_hermitian == herm
    Hs = H + H.conj().T   # strings on distinct sites are Hermitian: the adjoint conjugates the coefficient
    c2 = np.concatenate([coeffs, coeffs.conj()])
    t2, i2 = np.concatenate([types, types]), np.concatenate([idx, idx])
    assert np.allclose(dense_hamiltonian(N, c2, t2, i2), Hs)
    assert R.Operator(c2, t2, i2, True).is_hermitian


def test_ensembles_host_side():
    mc = R.MonteCarloLoop(4096 * 10, 1, 5, 4096, True)
    assert mc.num_steps == 4096 * 10
    mc2 = R.MonteCarloLoop(mc)
    assert mc2.num_steps == mc.num_steps
    with pytest.raises(ValueError, match="multiple"):
        R.MonteCarloLoop(100, 1, 1, 7, True)
    mc.set_fast_sweep(2)              # move settings are host state; copies keep them
    mc.set_total_z_symmetry(0)
    R.MonteCarloLoop(mc)
    ex = R.ExactSummation(10, True)
    assert ex.num_steps == 1024
    ex.set_total_z_symmetry(0)
    assert ex.num_steps == 252
    ex.set_total_z_symmetry(-2)
    assert ex.num_steps == 210


def test_spins_semantics():
    """include/Spins.h:125-148, 230-237, 280-285."""
    s = R.Spins(0b10110, 5)
    assert list(s.array(5)) == [-1, 1, 1, -1, 1]
    assert s.flip(0).configuration == 0b10111
    assert R.Spins(0xFF, 5).configuration == 0b11111
    assert s.rotate_left(2, 5).configuration == ((0b10110 << 2) | (0b10110 >> 3)) & 0b11111
    assert s.hamming_distance(R.Spins(0b00110, 5)) == 1
    # 2x3 lattice: shifting a single up spin moves it by one row and one column
    one = R.Spins(1, 6)
    assert one.shift_2d(1, 1, 2, 3).configuration == 1 << (1 * 3 + 1)


def test_json_roundtrip():
    psi = R.new_neural_network(3, 6, noise=1e-2, gpu=True, rng=np.random.default_rng(5))
    obj = json.loads(json.dumps(psi.to_json()))
    psi2 = R.Psi.from_json(obj, True)
    assert np.allclose(psi2.params, psi.params) and psi2.prefactor == psi.prefactor
    deep = R.new_deep_neural_network(4, [8, 4], [2, 8], gpu=True, rng=np.random.default_rng(6))
    deep2 = R.PsiDeep.from_json(json.loads(json.dumps(deep.to_json())), True)
    assert np.allclose(deep2.params, deep.params)
    assert all(np.array_equal(a, b) for a, b in zip(deep2.connections, deep.connections))


# Documents in the reference's on-disk format, written out by hand from
# pyRBMonGPU/json_numpy.py:5-38 (array objects), Psi.py:8-21 and
# PsiDeep.py:8-23 (the model records): a renamed key, another array encoding
# or another dtype string fails the equality below.
PSI_DOC = {
    "type": "Psi",
    "alpha": {"type": "ndarray", "dtype": "float64", "complex": False, "data": [0.5, 0.25]},
    "beta": {"type": "ndarray", "dtype": "float64", "complex": False, "data": [0.0, 1.0]},
    "b": {"type": "ndarray", "dtype": "complex128", "complex": True, "real": [1.0], "imag": [2.0]},
    "W": {"type": "ndarray", "dtype": "complex128", "complex": True, "real": [[0.5], [0.0]], "imag": [[-1.0], [2.0]]},
    "prefactor": 0.5,
}
DEEP_DOC = {
    "type": "PsiDeep",
    "alpha": {"type": "ndarray", "dtype": "float64", "complex": False, "data": [0.0, 0.0]},
    "beta": {"type": "ndarray", "dtype": "float64", "complex": False, "data": [0.0, 0.0]},
    "b": [{"type": "ndarray", "dtype": "complex128", "complex": True, "real": [0.25, -0.5], "imag": [0.0, 1.5]}],
    "connections": [{"type": "ndarray", "dtype": "uint32", "complex": False, "data": [[0, 1], [1, 0]]}],
    "W": [{"type": "ndarray", "dtype": "complex128", "complex": True, "real": [[1.0, 2.0], [3.0, 4.0]],
           "imag": [[0.5, 0.0], [0.0, -0.5]]}],
    "final_weights": {"type": "ndarray", "dtype": "complex128", "complex": True, "real": [10.0, 10.0],
                      "imag": [0.0, 0.0]},
    "prefactor": 1.0,
    "free_quantum_axis": False,
}


def test_json_documents_have_the_reference_layout():
    psi = R.Psi(np.array([0.5, 0.25]), np.array([0.0, 1.0]), np.array([1 + 2j]), np.array([[0.5 - 1j], [2j]]), 0.5,
                False, True)
    assert psi.to_json() == PSI_DOC
    # the text round trip keeps key order and value types
    assert json.loads(json.dumps(psi.to_json())) == PSI_DOC
    deep = R.PsiDeep(np.zeros(2), np.zeros(2), [np.array([0.25, -0.5 + 1.5j])],
                     [np.array([[0, 1], [1, 0]], dtype=np.uint32)], [np.array([[1 + 0.5j, 2], [3, 4 - 0.5j]])],
                     np.array([10.0, 10.0], complex), 1.0, False, True)
    assert deep.to_json() == DEEP_DOC


def test_json_documents_in_the_reference_layout_load():
    """A document in the reference's layout loads into the same model."""
    psi = R.Psi.from_json(json.loads(json.dumps(PSI_DOC)), True)
    assert np.array_equal(psi.alpha, [0.5, 0.25]) and np.array_equal(psi.beta, [0.0, 1.0])
    assert np.array_equal(psi.b, [1 + 2j]) and np.array_equal(psi.W, [[0.5 - 1j], [2j]])
    assert psi.prefactor == 0.5
    deep = R.PsiDeep.from_json(json.loads(json.dumps(DEEP_DOC)), True)
    assert np.array_equal(deep.W[0], [[1 + 0.5j, 2], [3, 4 - 0.5j]])
    assert np.array_equal(deep.b[0], [0.25, -0.5 + 1.5j])
    assert np.array_equal(deep.connections[0], [[0, 1], [1, 0]]) and deep.connections[0].dtype == np.uint32
    assert np.array_equal(deep.final_weights, [10.0, 10.0])
    # the encoder's array objects decode with the reference's rule
    arr = json.loads(json.dumps({"x": np.arange(3, dtype=np.int64)}, cls=R.NumpyEncoder), cls=R.NumpyDecoder)["x"]
    assert arr.dtype == np.int64 and list(arr) == [0, 1, 2]


def test_new_neural_network_init_matches_reference_rule():
    """W = noise * U(-1,1)(1+i) + initial_value on the M // N diagonal blocks."""
    psi = R.new_neural_network(4, 8, noise=0.0, gpu=True)
    W = psi.W
    iv = 0.01 + 1j * np.pi / 4
    assert np.allclose(W[:, :4], iv * np.eye(4)) and np.allclose(W[:, 4:], iv * np.eye(4))
    assert np.allclose(psi.b, 0)


def test_made_training_launch_extents_without_gpu():
    """Every kernel launch of a proposal-network training step stays within
    its buffer (host-side extent check of the launch arguments against the
    allocations), over network shapes up to the 100 -> 256 -> 100 learned
    proposal and batches around the 64-row padding and split-K boundaries."""
    lib = ctypes.CDLL(LIB)
    lib.rbmhip_made_check_launches.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int64, ctypes.c_int]
    lib.rbmhip_last_error.restype = ctypes.c_char_p
    for N in (1, 2, 7, 20, 63, 64, 65, 100, 128):
        for H in (1, 16, 64, 65, 256, 512):
            for batch in (1, 63, 64, 65, 128, 1000, 4096, 4097, 65536):
                rc = lib.rbmhip_made_check_launches(N, H, 1_007_616, batch)
                assert rc == 0, (N, H, batch, lib.rbmhip_last_error())
    assert lib.rbmhip_made_check_launches(0, 256, 10, 10) == -1


def test_deep_connections_2d_layouts():
    """2-D connection tables (pyRBMonGPU/new_neural_network.py:128-145):
    a lattice of width 1 reduces to the 1-D rule; every unit reads a c0 x c1
    window of distinct inputs; with stride 1 the window of unit (j0 + 1, j1)
    is that of (j0, j1) moved down one row (periodically)."""
    from rbmongpu_amd.new_neural_network import deep_connections
    one_d = deep_connections(8, [8, 4], [3, 8])
    assert all(np.array_equal(a, b) for a, b in zip(deep_connections((8, 1), [(8, 1), (4, 1)], [(3, 1), (8, 1)]), one_d))
    assert all(np.array_equal(a, b) for a, b in zip(deep_connections((1, 8), [(1, 8), (1, 4)], [(1, 3), (1, 8)]), one_d))
    (t,) = deep_connections((6, 6), [(6, 6)], [(3, 2)])
    assert t.shape == (6, 36)
    for unit in range(36):
        assert len(set(t[:, unit])) == 6
    for j0 in range(6):
        for j1 in range(6):
            here = {(int(x) // 6, int(x) % 6) for x in t[:, j0 * 6 + j1]}
            below = {(int(x) // 6, int(x) % 6) for x in t[:, ((j0 + 1) % 6) * 6 + j1]}
            assert below == {((r + 1) % 6, c) for r, c in here}
    # a layer with fewer units strides over the inputs: 3x3 units over 6x6 inputs
    (t2,) = deep_connections((6, 6), [(3, 3)], [(2, 2)])
    assert sorted({int(x) for x in t2.ravel()}) == list(range(36))
    deep = R.new_deep_neural_network((4, 4), [(4, 4), (2, 2)], [(2, 2), (4, 4)], gpu=True,
                                     rng=np.random.default_rng(3))
    assert deep.N == 16
    assert [w.shape for w in deep.W] == [(4, 16), (16, 4)]
