"""Pauli-string tables for the benchmark Hamiltonians, in the reference
Operator's array format (include/operator/Operator.hpp:212-227):
coefficients[S] complex, pauli_types[S, L] (0=I, 1=X, 2=Y, 3=Z) and
pauli_indices[S, L], each row terminated by -1.  The reference builds these
through the unvendored QuantumExpression package; here they are plain lists."""
import numpy as np


def table(strings):
    """strings: list of (coefficient, [(type, index), ...]) -> (coeffs, types, indices)."""
    L = max(len(s[1]) for s in strings) + 1
    S = len(strings)
    coeffs = np.zeros(S, dtype=np.complex128)
    types = np.zeros((S, L), dtype=np.int32)
    idx = -np.ones((S, L), dtype=np.int32)
    for n, (c, ops) in enumerate(strings):
        coeffs[n] = c
        for t, (pt, pi) in enumerate(ops):
            types[n, t] = pt
            idx[n, t] = pi
    return coeffs, types, idx


def tfim_1d(N, J=1.0, h=1.0, periodic=True):
    """H = -J sum_i Z_i Z_{i+1} - h sum_i X_i."""
    strings = []
    for i in range(N if periodic else N - 1):
        strings.append((-J, [(3, i), (3, (i + 1) % N)]))
    for i in range(N):
        strings.append((-h, [(1, i)]))
    return table(strings)


def square_bonds(Lx, Ly):
    bonds = []
    for y in range(Ly):
        for x in range(Lx):
            i = y * Lx + x
            bonds.append((i, y * Lx + (x + 1) % Lx))
            bonds.append((i, ((y + 1) % Ly) * Lx + x))
    return bonds


def tfim_2d(Lx, Ly, J=1.0, h=3.04):
    """2-D periodic TFIM: -J sum_<ij> Z_i Z_j - h sum_i X_i."""
    strings = [(-J, [(3, i), (3, j)]) for i, j in square_bonds(Lx, Ly)]
    strings += [(-h, [(1, i)]) for i in range(Lx * Ly)]
    return table(strings)


def heisenberg(bonds, J=1.0):
    """sum_<ij> J (X_i X_j + Y_i Y_j + Z_i Z_j): three strings per bond."""
    strings = []
    for i, j in bonds:
        for p in (1, 2, 3):
            strings.append((J, [(p, i), (p, j)]))
    return table(strings)


def heisenberg_2d(Lx, Ly, J=1.0):
    return heisenberg(square_bonds(Lx, Ly), J)


def j1j2_chain(N, J1=1.0, J2=0.5):
    strings = []
    for d, J in ((1, J1), (2, J2)):
        for i in range(N):
            for p in (1, 2, 3):
                strings.append((J, [(p, i), (p, (i + d) % N)]))
    return table(strings)


def operator(tab, gpu=True):
    from . import Operator
    return Operator(tab[0], tab[1], tab[2], gpu)
