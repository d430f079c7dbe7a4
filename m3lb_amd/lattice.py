"""Velocity sets in esoteric-pull order (src/lattice/Lattices.hpp:72-96, 125-159).

Index 0 is the rest velocity; (1,2), (3,4), ... are opposite pairs, so
opp(i) = i+1 for odd i and i-1 for even i (Lattices.hpp:40-45).  Weights are
exact rationals w_num / w_den.
"""
import numpy as np

from . import abi

D2Q9_C = np.array([[0, 1, -1, 0, 0, 1, -1, 1, -1],
                   [0, 0, 0, 1, -1, 1, -1, -1, 1],
                   [0, 0, 0, 0, 0, 0, 0, 0, 0]], dtype=np.int64).T
D2Q9_W = (np.array([16, 4, 4, 4, 4, 1, 1, 1, 1], dtype=np.int64), 36)

D3Q27_C = np.array([[0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0,
                     1, -1, 1, -1, 1, -1, -1, 1],
                    [0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1,
                     1, -1, 1, -1, -1, 1, 1, -1],
                    [0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1,
                     1, -1, -1, 1, 1, -1, 1, -1]], dtype=np.int64).T
D3Q27_W = (np.array([64] + [16] * 6 + [4] * 12 + [1] * 8, dtype=np.int64), 216)

CS2 = 1.0 / 3.0


def velocities(lattice):
    return D2Q9_C if lattice == abi.D2Q9 else D3Q27_C


def weights(lattice):
    num, den = D2Q9_W if lattice == abi.D2Q9 else D3Q27_W
    return num / den


def weights_rational(lattice):
    return D2Q9_W if lattice == abi.D2Q9 else D3Q27_W


def Q_of(lattice):
    return 9 if lattice == abi.D2Q9 else 27


def D_of(lattice):
    return 2 if lattice == abi.D2Q9 else 3


def opp(i):
    return 0 if i == 0 else (i + 1 if i & 1 else i - 1)
