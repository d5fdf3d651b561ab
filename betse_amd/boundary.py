"""Applied boundary voltage: the potential ``Phi_b`` of ``bound_V``.

The reference's voltage event (``TissueEventVoltage.fire``,
``betse/science/tissue/event/tisevevolt.py:75-89``) sets ``sim.bound_V`` on
one or two edges of the environment grid; ``get_current``
(``betse/science/physics/ion_current.py:84-90``) then solves
``Phi_b = lapENVinv . (-div_Jb)``, where ``div_Jb`` holds ``-bound_V/delta**2``
on the grid's edge rows and columns (columns first, then rows, so the top and
bottom values win at the corners) and zero inside, and ``lapENVinv`` inverts
the 5-point Laplacian whose edge rows are identity rows ('value' boundaries,
``finitediff.py:252-431``). That is the discrete harmonic extension of the edge
values: ``Phi_b`` equals ``bound_V`` on the edges, and inside it solves
Laplace's equation with them as Dirichlet data. ``delta`` cancels.

The interior solve here is a DST-I (type-1 discrete sine transform)
diagonalisation of the interior Dirichlet Laplacian, O(n^2 log n) instead of
the reference's dense (n^2)^2 product.
"""

from __future__ import annotations

from typing import Mapping, Tuple

import numpy as np
import scipy.fft

SIDES = ("T", "B", "L", "R")


def edge_values(bound_V: Mapping[str, float], shape: Tuple[int, int]) -> np.ndarray:
    """The grid with ``bound_V`` on its edges (``div_Jb``'s assignment order,
    ``ion_current.py:85-88``: right and left columns, then top and bottom rows)
    and zero inside."""
    ny, nx = shape
    g = np.zeros((ny, nx))
    g[:, -1] = bound_V["R"]
    g[:, 0] = bound_V["L"]
    g[-1, :] = bound_V["T"]
    g[0, :] = bound_V["B"]
    return g


def boundary_potential(bound_V: Mapping[str, float], shape: Tuple[int, int]) -> np.ndarray:
    """``Phi_b`` (flat, row-major ``ny*nx``) of the applied voltages ``bound_V``
    (keys ``T``, ``B``, ``L``, ``R``) on a grid of ``shape``."""
    for k in SIDES:
        if not np.isfinite(float(bound_V[k])):
            raise ValueError(f"bound_V[{k!r}] is not finite")
    ny, nx = shape
    phi = edge_values(bound_V, shape)
    n1, n2 = ny - 2, nx - 2
    if n1 > 0 and n2 > 0 and np.any(phi != 0):
        # interior rows: u[i-1,j] + u[i+1,j] + u[i,j-1] + u[i,j+1] - 4 u[i,j] = 0,
        # with the edge neighbours moved to the right-hand side
        rhs = np.zeros((n1, n2))
        rhs[0, :] -= phi[0, 1:-1]
        rhs[-1, :] -= phi[-1, 1:-1]
        rhs[:, 0] -= phi[1:-1, 0]
        rhs[:, -1] -= phi[1:-1, -1]
        # eigenvalues of the 1-D Dirichlet second difference tridiag(1, -2, 1)
        ly = 2.0 * np.cos(np.pi * np.arange(1, n1 + 1) / (n1 + 1)) - 2.0
        lx = 2.0 * np.cos(np.pi * np.arange(1, n2 + 1) / (n2 + 1)) - 2.0
        coef = scipy.fft.dstn(rhs, type=1) / (ly[:, None] + lx[None, :])
        phi[1:-1, 1:-1] = scipy.fft.idstn(coef, type=1)
    return phi.ravel()
