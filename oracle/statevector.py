"""Independent FP64 state-vector simulator for qFlex-format circuits.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Built from textbook gate matrices (no tensor-network code, no reference gate
tables), so it pins the gate tables, their (input, output) storage
convention, the two-qubit splits and the whole contraction path.

Conventions (reference src/evaluate_circuit.cpp:22-36,
src/read_circuit.cpp:348-366): qubit number q = i*J + j on the I x J grid;
the bitstring position of an active qubit is its row-major rank among active
qubits; amplitude = <final| U |initial>.
"""
import math

import numpy as np

_I2 = np.eye(2, dtype=np.complex128)
_X = np.array([[0, 1], [1, 0]], dtype=np.complex128)
_Y = np.array([[0, -1j], [1j, 0]], dtype=np.complex128)
_W = (_X + _Y) / math.sqrt(2)


def _sqrt_involution(P):
    # For P^2 = I: sqrt(P) = (1+i)/2 I + (1-i)/2 P  (principal root).
    return 0.5 * (1 + 1j) * _I2 + 0.5 * (1 - 1j) * P


def one_qubit(name, params):
    if name == "h":
        return np.array([[1, 1], [1, -1]], dtype=np.complex128) / math.sqrt(2)
    if name == "t":
        return np.diag([1, np.exp(1j * math.pi / 4)]).astype(np.complex128)
    if name == "x_1_2":
        return _sqrt_involution(_X)
    if name == "y_1_2":
        return _sqrt_involution(_Y)
    if name == "hz_1_2":
        return _sqrt_involution(_W)
    if name == "rz":
        a = math.pi * params[0]
        return np.diag([np.exp(-0.5j * a), np.exp(0.5j * a)])
    raise ValueError(name)


def two_qubit(name, params):
    """4x4 matrix on (q1, q2), q1 the more significant bit."""
    if name == "cz":
        return np.diag([1, 1, 1, -1]).astype(np.complex128)
    if name == "cx":  # control q1, target q2
        m = np.eye(4, dtype=np.complex128)
        m[2:, 2:] = _X
        return m
    if name == "fsim":
        th, ph = math.pi * params[0], math.pi * params[1]
        c, s = math.cos(th), math.sin(th)
        return np.array([[1, 0, 0, 0], [0, c, -1j * s, 0], [0, -1j * s, c, 0],
                         [0, 0, 0, np.exp(-1j * ph)]], dtype=np.complex128)
    raise ValueError(name)


def final_state_vector(circuit, I, J, off, initial):
    """Returns psi as an array of shape (2,)*n over active qubits."""
    offs = set(map(tuple, off))
    active = [q for q in range(I * J) if divmod(q, J) not in offs]
    pos = {q: k for k, q in enumerate(active)}
    n = len(active)
    psi = np.zeros((2,) * n, dtype=np.complex128)
    psi[tuple(int(c) for c in initial)] = 1.0
    for g in circuit.gates:
        if len(g.qubits) == 1:
            a = pos[g.qubits[0]]
            U = one_qubit(g.name, g.params)
            psi = np.moveaxis(np.tensordot(U, psi, axes=([1], [a])), 0, a)
        else:
            a, b = pos[g.qubits[0]], pos[g.qubits[1]]
            U = two_qubit(g.name, g.params).reshape(2, 2, 2, 2)
            psi = np.tensordot(U, psi, axes=([2, 3], [a, b]))
            psi = np.moveaxis(psi, [0, 1], [a, b])
    return psi


def amplitudes(circuit, I, J, off, initial, finals):
    psi = final_state_vector(circuit, I, J, off, initial)
    return np.array([psi[tuple(int(c) for c in f)] for f in finals])
