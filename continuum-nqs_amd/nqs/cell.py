"""2D simulation cell helpers (same names as src/nqs/cell.py:14-46).

A cell is a 2x2 matrix A whose columns are the lattice vectors; the square box of side L is
A = L I.  frac/wrap accept numpy arrays or torch tensors.
"""
import numpy as np


def square_cell(L):
    return np.array([[float(L), 0.0], [0.0, float(L)]])


def rect_cell(Lx, Ly):
    return np.array([[float(Lx), 0.0], [0.0, float(Ly)]])


def hex_cell(a):
    a = float(a)
    return np.array([[a, a / 2.0], [0.0, a * np.sqrt(3) / 2.0]])


def as_cell(cell):
    """Scalar L -> L I, else the 2x2 matrix."""
    return square_cell(cell) if np.ndim(cell) == 0 else np.asarray(cell, dtype=float)


def reciprocal(A):
    return 2.0 * np.pi * np.linalg.inv(np.asarray(A, float)).T


def cell_area(A):
    return float(abs(np.linalg.det(np.asarray(A, float))))


def frac(r, A):
    """Cartesian -> fractional coordinates mod 1."""
    Ainv = np.linalg.inv(np.asarray(A, float))
    if hasattr(r, "new_tensor"):
        import torch
        return torch.remainder(r @ r.new_tensor(Ainv).T, 1.0)
    return np.mod(np.asarray(r) @ Ainv.T, 1.0)


def wrap(r, A):
    """Wrap a cartesian point into the cell."""
    A = np.asarray(A, float)
    s = frac(r, A)
    return s @ (s.new_tensor(A).T if hasattr(s, "new_tensor") else A.T)
