"""ctypes front end of the FP64 CPU oracle (oracle/rbm_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker.  The product (rbmongpu_amd) never
imports this module.

Models, operators and ensembles are passed as plain numpy arrays; the wrapper
keeps references alive for the duration of each call.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_SRC = os.path.join(_HERE, "rbm_oracle.c")


def build(force: bool = False) -> str:
    """Compile rbm_oracle.c with gcc into oracle/liboracle.so."""
    if force or not os.path.exists(_LIB_PATH) or os.path.getmtime(_LIB_PATH) < os.path.getmtime(_SRC):
        cmd = ["gcc", "-O2", "-std=c11", "-fPIC", "-shared", "-o", _LIB_PATH, _SRC, "-lm"]
        subprocess.run(cmd, check=True)
    return _LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        build()
        _lib = C.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


class Model(C.Structure):
    _fields_ = [
        ("kind", C.c_int32), ("N", C.c_int32), ("prefactor", C.c_double),
        ("M", C.c_int32), ("b", C.c_void_p), ("W", C.c_void_p), ("a", C.c_void_p),
        ("num_layers", C.c_int32), ("sizes", C.c_void_p), ("conn_counts", C.c_void_p),
        ("conns", C.c_void_p), ("weights", C.c_void_p), ("biases", C.c_void_p),
        ("final_weights", C.c_void_p), ("translational_invariance", C.c_int32),
        ("ref_compat_final", C.c_int32),
    ]


class Op(C.Structure):
    _fields_ = [("S", C.c_int32), ("L", C.c_int32), ("coeff", C.c_void_p),
                ("types", C.c_void_p), ("idx", C.c_void_p)]


class Ens(C.Structure):
    _fields_ = [
        ("kind", C.c_int32), ("N", C.c_int32), ("num_configs", C.c_int64), ("configs", C.c_void_p),
        ("num_samples", C.c_uint32), ("num_sweeps", C.c_uint32),
        ("num_thermalization_sweeps", C.c_uint32), ("num_markov_chains", C.c_uint32),
        ("seed", C.c_uint64), ("call_index", C.c_uint32),
        ("acceptances", C.c_uint64), ("rejections", C.c_uint64),
        ("move", C.c_int32), ("tries", C.c_int32), ("k_up", C.c_int32),
    ]


def _declare(L):
    P = C.c_void_p
    L.orc_logcosh.argtypes = [P, P, C.c_int64]
    L.orc_tanh.argtypes = [P, P, C.c_int64]
    L.orc_philox.argtypes = [P, P, P]
    L.orc_num_params.argtypes = [P]; L.orc_num_params.restype = C.c_int64
    L.orc_log_psi.argtypes = [P, P, C.c_int64, P]
    L.orc_o_k.argtypes = [P, P, P]
    L.orc_local_energy.argtypes = [P, P, P, C.c_int64, P]
    L.orc_ens_num_steps.argtypes = [P]; L.orc_ens_num_steps.restype = C.c_int64
    L.orc_records.argtypes = [P, P, P, P, P, P, P]
    L.orc_expectation.argtypes = [P, P, P, P]
    L.orc_fluctuation.argtypes = [P, P, P, P, P]
    L.orc_gradient.argtypes = [P, P, P, C.c_int, P, P]
    L.orc_fluctuation_gradient.argtypes = [P, P, P, P]
    L.orc_difference.argtypes = [P, P, P, C.c_int, P, P]
    L.orc_corrected.argtypes = [P, P, P, P, P]
    L.orc_psi_norm.argtypes = [P, P]; L.orc_psi_norm.restype = C.c_double
    L.orc_psi_vector.argtypes = [P, P]
    L.orc_s_matrix.argtypes = [P, P, P]
    L.orc_hs_distance.argtypes = [P, P, P, C.c_int, P, C.c_int, C.c_double, P]
    L.orc_hs_distance.restype = C.c_double
    L.orc_hs_distance_axis.argtypes = [P, P, P, C.c_int, P, C.c_int, C.c_double, P, P, P, P, P]
    L.orc_hs_distance_axis.restype = C.c_double


def _ptr(a):
    return None if a is None else a.ctypes.data


class RBM:
    """RBM wavefunction (reference Psi, include/quantum_state/Psi.hpp:31-53)."""

    def __init__(self, b, W, a=None, prefactor=1.0):
        self.b = np.ascontiguousarray(b, dtype=np.complex128)
        self.W = np.ascontiguousarray(W, dtype=np.complex128)
        self.a = None if a is None else np.ascontiguousarray(a, dtype=np.complex128)
        self.N, self.M = self.W.shape
        self.prefactor = float(prefactor)

    def struct(self):
        m = Model()
        m.kind = 0; m.N = self.N; m.prefactor = self.prefactor; m.M = self.M
        m.b = _ptr(self.b); m.W = _ptr(self.W); m.a = _ptr(self.a)
        return m

    @property
    def num_params(self):
        return 2 * self.N + self.M + self.N * self.M + (self.N if self.a is not None else 0)


class Deep:
    """PsiDeep (include/quantum_state/PsiDeep.hpp:39-84): per layer a C x size
    connection table and weight table, biases, and final weights."""

    def __init__(self, N, biases, connections, weights, final_weights, prefactor=1.0,
                 translational_invariance=False, ref_compat_final=True):
        self.N = int(N)
        self.sizes = np.array([len(b) for b in biases], dtype=np.int32)
        self.conn_counts = np.array([np.asarray(c).shape[0] for c in connections], dtype=np.int32)
        self.conns = np.ascontiguousarray(np.concatenate([np.asarray(c, dtype=np.uint32).ravel() for c in connections]))
        self.weights = np.ascontiguousarray(np.concatenate([np.asarray(w, dtype=np.complex128).ravel() for w in weights]))
        self.biases = np.ascontiguousarray(np.concatenate([np.asarray(b, dtype=np.complex128).ravel() for b in biases]))
        self.final_weights = np.ascontiguousarray(final_weights, dtype=np.complex128)
        self.prefactor = float(prefactor)
        self.translational_invariance = bool(translational_invariance)
        self.ref_compat_final = bool(ref_compat_final)

    def struct(self):
        m = Model()
        m.kind = 1; m.N = self.N; m.prefactor = self.prefactor
        m.num_layers = len(self.sizes)
        m.sizes = _ptr(self.sizes); m.conn_counts = _ptr(self.conn_counts); m.conns = _ptr(self.conns)
        m.weights = _ptr(self.weights); m.biases = _ptr(self.biases); m.final_weights = _ptr(self.final_weights)
        m.translational_invariance = int(self.translational_invariance)
        m.ref_compat_final = int(self.ref_compat_final)
        return m

    @property
    def num_params(self):
        return int(lib().orc_num_params(C.byref(self.struct())))


class Operator:
    """Pauli-string table (include/operator/Operator.hpp:30-39)."""

    def __init__(self, coefficients, pauli_types, pauli_indices):
        self.coeff = np.ascontiguousarray(coefficients, dtype=np.complex128)
        self.types = np.ascontiguousarray(pauli_types, dtype=np.int32)
        self.idx = np.ascontiguousarray(pauli_indices, dtype=np.int32)
        self.S, self.L = self.types.shape

    def struct(self):
        o = Op()
        o.S = self.S; o.L = self.L
        o.coeff = _ptr(self.coeff); o.types = _ptr(self.types); o.idx = _ptr(self.idx)
        return o


class Exact:
    """ExactSummation over 2^N states, or over a total-S_z sector."""

    def __init__(self, N, sector=None):
        self.N = N
        self.configs = None
        if sector is None:
            self.num_configs = 1 << N
        else:
            idx = np.arange(1 << N, dtype=np.uint64)
            pop = np.array([bin(int(i)).count("1") for i in idx])
            sel = idx[2 * pop - N == sector]
            self.configs = np.ascontiguousarray(np.stack([sel, np.zeros_like(sel)], axis=1))
            self.num_configs = len(sel)

    def struct(self):
        e = Ens()
        e.kind = 0; e.N = self.N; e.num_configs = self.num_configs; e.configs = _ptr(self.configs)
        return e


class MonteCarlo:
    """Metropolis ensemble with the shared Philox stream: single spin flips, or
    with total_z the exchange move in that S_z sector
    (MonteCarloLoop.set_total_z_symmetry), or with fast_sweep the multi-flip
    move (set_fast_sweep)."""

    def __init__(self, N, num_samples, num_sweeps, num_thermalization_sweeps, num_markov_chains, seed=0, call_index=0,
                 total_z=None, fast_sweep=None):
        self.N = N
        self.params = (num_samples, num_sweeps, num_thermalization_sweeps, num_markov_chains)
        self.seed = seed
        self.call_index = call_index
        self.total_z = total_z
        self.fast_sweep = fast_sweep
        self.acceptances = 0
        self.rejections = 0

    def struct(self):
        e = Ens()
        e.kind = 1; e.N = self.N
        e.num_samples, e.num_sweeps, e.num_thermalization_sweeps, e.num_markov_chains = self.params
        e.seed = self.seed; e.call_index = self.call_index
        if self.total_z is not None:
            e.move = 1; e.k_up = (self.total_z + self.N) // 2
        elif self.fast_sweep is not None:
            e.move = 2; e.tries = self.fast_sweep
        return e


class _Call:
    def __init__(self, psi, op, ens):
        self.m = psi.struct()
        self.o = op.struct() if op is not None else None
        self.e = ens.struct()
        self.ens = ens

    def args(self):
        return (C.byref(self.m), C.byref(self.o) if self.o is not None else None, C.byref(self.e))

    def done(self):
        if isinstance(self.ens, MonteCarlo):
            self.ens.acceptances = self.e.acceptances
            self.ens.rejections = self.e.rejections


def logcosh(z):
    z = np.ascontiguousarray(np.atleast_1d(z), dtype=np.complex128)
    out = np.empty_like(z)
    lib().orc_logcosh(z.ctypes.data, out.ctypes.data, z.size)
    return out


def tanh(z):
    z = np.ascontiguousarray(np.atleast_1d(z), dtype=np.complex128)
    out = np.empty_like(z)
    lib().orc_tanh(z.ctypes.data, out.ctypes.data, z.size)
    return out


def philox(ctr, key):
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    out = np.empty(4, dtype=np.uint32)
    lib().orc_philox(c.ctypes.data, k.ctypes.data, out.ctypes.data)
    return out


def spins_words(configs):
    """Python ints or an array of basis indices -> (n, 2) uint64 words."""
    configs = [int(c) for c in np.atleast_1d(configs)]
    return np.ascontiguousarray(
        np.array([[c & 0xFFFFFFFFFFFFFFFF, (c >> 64) & 0xFFFFFFFFFFFFFFFF] for c in configs], dtype=np.uint64))


def log_psi(psi, spins):
    s = np.ascontiguousarray(spins, dtype=np.uint64).reshape(-1, 2)
    out = np.empty(len(s), dtype=np.complex128)
    m = psi.struct()
    lib().orc_log_psi(C.byref(m), s.ctypes.data, len(s), out.ctypes.data)
    return out


def o_k(psi, config_words):
    s = np.ascontiguousarray(config_words, dtype=np.uint64).reshape(2)
    out = np.zeros(psi.num_params, dtype=np.complex128)
    m = psi.struct()
    lib().orc_o_k(C.byref(m), s.ctypes.data, out.ctypes.data)
    return out


def local_energy(psi, op, spins):
    s = np.ascontiguousarray(spins, dtype=np.uint64).reshape(-1, 2)
    out = np.empty(len(s), dtype=np.complex128)
    m, o = psi.struct(), op.struct()
    lib().orc_local_energy(C.byref(m), C.byref(o), s.ctypes.data, len(s), out.ctypes.data)
    return out


def records(psi, op, ens):
    c = _Call(psi, op, ens)
    n = int(lib().orc_ens_num_steps(C.byref(c.e)))
    spins = np.zeros((n, 2), dtype=np.uint64)
    lp = np.zeros(n, dtype=np.complex128)
    el = np.zeros(n, dtype=np.complex128)
    w = np.zeros(n, dtype=np.float64)
    m, o, e = c.args()
    lib().orc_records(m, o, e, spins.ctypes.data, lp.ctypes.data, el.ctypes.data if op is not None else None,
                      w.ctypes.data)
    c.done()
    return spins, lp, el, w


def expectation(psi, op, ens):
    c = _Call(psi, op, ens)
    out = np.zeros(1, dtype=np.complex128)
    lib().orc_expectation(*c.args(), out.ctypes.data)
    c.done()
    return complex(out[0])


def fluctuation(psi, op, ens):
    c = _Call(psi, op, ens)
    f = np.zeros(1); E = np.zeros(1, dtype=np.complex128)
    lib().orc_fluctuation(*c.args(), f.ctypes.data, E.ctypes.data)
    c.done()
    return float(f[0]), complex(E[0])


def gradient(psi, op, ens, mode=0):
    c = _Call(psi, op, ens)
    g = np.zeros(psi.num_params, dtype=np.complex128)
    E = np.zeros(1, dtype=np.complex128)
    lib().orc_gradient(*c.args(), mode, g.ctypes.data, E.ctypes.data)
    c.done()
    return g, complex(E[0])


def fluctuation_gradient(psi, op, ens):
    """ExpectationValue::fluctuation_gradient (ExpectationValue.cu:258-332)."""
    c = _Call(psi, op, ens)
    g = np.zeros(psi.num_params, dtype=np.complex128)
    lib().orc_fluctuation_gradient(*c.args(), g.ctypes.data)
    c.done()
    return g


def difference(psi, psi_prime, ops, ens):
    """ExpectationValue::difference (ExpectationValue.cu:334-446)."""
    m, mp, e = psi.struct(), psi_prime.struct(), ens.struct()
    structs = [o.struct() for o in ops]
    arr = (C.c_void_p * len(ops))(*[C.cast(C.pointer(x), C.c_void_p) for x in structs])
    out = np.zeros(len(ops), dtype=np.complex128)
    lib().orc_difference(C.byref(m), C.byref(mp), arr, len(ops), C.byref(e), out.ctypes.data)
    return list(out)


def corrected(psi, op, ens):
    """ExpectationValue::corrected (ExpectationValue.cu:74-127): (<E>/<X>, <X>)."""
    c = _Call(psi, op, ens)
    v = np.zeros(1, dtype=np.complex128); x = np.zeros(1)
    lib().orc_corrected(*c.args(), v.ctypes.data, x.ctypes.data)
    c.done()
    return complex(v[0]), float(x[0])


def psi_norm(psi, ens):
    c = _Call(psi, None, ens)
    m, _, e = c.args()
    return float(lib().orc_psi_norm(m, e))


def psi_vector(psi):
    out = np.zeros(1 << psi.N, dtype=np.complex128)
    m = psi.struct()
    lib().orc_psi_vector(C.byref(m), out.ctypes.data)
    return out


def s_matrix(psi, ens):
    c = _Call(psi, None, ens)
    P = psi.num_params
    S = np.zeros((P, P), dtype=np.complex128)
    m, _, e = c.args()
    lib().orc_s_matrix(m, e, S.ctypes.data)
    return S


def hs_distance(psi, psi_prime, op, is_unitary, ens, want_grad=False, nu=1.0, axis=None):
    """axis = (alpha, beta, alpha_prime, beta_prime) switches on the free
    quantum axis of the first wave function."""
    c = _Call(psi, op, ens)
    mp = psi_prime.struct()
    g = np.zeros(psi_prime.num_params, dtype=np.complex128)
    m, o, e = c.args()
    if axis is None:
        d = lib().orc_hs_distance(m, C.byref(mp), o, int(bool(is_unitary)), e, int(bool(want_grad)),
                                  float(np.float32(nu)), g.ctypes.data)
    else:
        ax = [np.ascontiguousarray(a, dtype=np.float64) for a in axis]
        d = lib().orc_hs_distance_axis(m, C.byref(mp), o, int(bool(is_unitary)), e, int(bool(want_grad)),
                                       float(np.float32(nu)), g.ctypes.data, *[a.ctypes.data for a in ax])
    c.done()
    return (g, d) if want_grad else d
