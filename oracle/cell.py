"""Oracle cell helpers (test infrastructure).  Follows src/nqs/cell.py:14-46 and
src/nqs/sampler.py:19-31 (scalar L -> L*I, wrap = mod(r A^-T, 1) A^T)."""
import numpy as np


def cell_matrix(cell):
    """sampler.py:19-22 -- scalar side L -> L*I, else the 2x2 matrix with lattice vectors as columns."""
    if np.ndim(cell) == 0:
        return np.array([[float(cell), 0.0], [0.0, float(cell)]])
    return np.asarray(cell, dtype=float)


def reciprocal(A):
    """cell.py:34-36 -- B = 2 pi (A^-1)^T."""
    return 2.0 * np.pi * np.linalg.inv(np.asarray(A, float)).T


def cell_area(A):
    """cell.py:39-40."""
    return float(abs(np.linalg.det(np.asarray(A, float))))
