"""The host solve of the applied boundary potential ``Phi_b``
(``betse_amd.boundary.boundary_potential``) against the reference's own
``get_current`` output with a nonzero ``bound_V`` (golden vectors recorded by
``tests/golden/make_golden_vectors.py`` from ``physics/ion_current.py``, whose
solve is the dense ``lapENVinv`` product), against the oracle's sparse solve
of the same system on grids of other shapes, and against what the mathematics
fixes: a constant edge value extends to a constant, the solution is linear in
``bound_V``, equals the edge values on the edges (rows win at the corners)
and satisfies the discrete Laplace equation inside."""

import os

import numpy as np
import pytest
import scipy.sparse.linalg

from betse_amd.boundary import boundary_potential, edge_values
from betse_amd.cells import Cells
from betse_amd.parameters import Parameters
from oracle.betse_oracle import make_laplacian_value

HERE = os.path.dirname(os.path.abspath(__file__))


def oracle_phi(bound, shape, delta=1e-5):
    """``lapENVinv . (-div_Jb)`` of ``get_current`` (ion_current.py:84-90) by a sparse direct solve."""
    div = np.zeros(shape)
    div[:, -1] = -bound["R"] / delta ** 2
    div[:, 0] = -bound["L"] / delta ** 2
    div[-1, :] = -bound["T"] / delta ** 2
    div[0, :] = -bound["B"] / delta ** 2
    A = make_laplacian_value(shape[0], shape[1], delta, sparse=True).tocsc()
    return scipy.sparse.linalg.spsolve(A, -div.ravel())


def test_matches_reference_get_current():
    with np.load(os.path.join(HERE, "golden", "reference_get_current.npz"), allow_pickle=False) as z:
        L, R, T, B = (float(x) for x in z["c1_in_bound"])
        ref = np.asarray(z["c1_out_Phi_b"]).ravel()
        n_sq = z["A"].shape[0]
    # the golden vectors' cluster is the tiny preset's (its Laplacian is (ny*nx)^2)
    cells = Cells.make_world(Parameters.preset("tiny"))
    assert cells.X.size == n_sq
    got = boundary_potential({"L": L, "R": R, "T": T, "B": B}, cells.X.shape)
    assert np.abs(ref).max() > 0
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-13 * np.abs(ref).max())


@pytest.mark.parametrize("shape", [(3, 3), (3, 8), (7, 4), (16, 16), (31, 23), (64, 65)])
def test_matches_oracle_solve(shape):
    rng = np.random.RandomState(sum(shape))
    bound = {k: float(rng.uniform(-0.1, 0.1)) for k in "TBLR"}
    ref = oracle_phi(bound, shape)
    got = boundary_potential(bound, shape)
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-13 * np.abs(ref).max())


def test_properties():
    shape = (21, 34)
    ny, nx = shape
    # a constant edge value extends to that constant
    np.testing.assert_allclose(boundary_potential(dict(T=0.05, B=0.05, L=0.05, R=0.05), shape), 0.05,
                               rtol=1e-14, atol=0)
    # edges hold bound_V, rows win at the corners
    b = dict(T=0.3, B=-0.2, L=0.1, R=-0.4)
    phi = boundary_potential(b, shape).reshape(shape)
    np.testing.assert_allclose(phi[0, :], -0.2, rtol=1e-15)
    np.testing.assert_allclose(phi[-1, :], 0.3, rtol=1e-15)
    np.testing.assert_allclose(phi[1:-1, 0], 0.1, rtol=1e-15)
    np.testing.assert_allclose(phi[1:-1, -1], -0.4, rtol=1e-15)
    np.testing.assert_array_equal(edge_values(b, shape)[[0, -1, 0, -1], [0, 0, -1, -1]], [-0.2, 0.3, -0.2, 0.3])
    # discrete Laplace equation inside
    lap = phi[:-2, 1:-1] + phi[2:, 1:-1] + phi[1:-1, :-2] + phi[1:-1, 2:] - 4 * phi[1:-1, 1:-1]
    assert np.abs(lap).max() < 1e-14
    # maximum principle and linearity
    assert phi.max() <= 0.3 + 1e-15 and phi.min() >= -0.4 - 1e-15
    b2 = dict(T=-0.1, B=0.25, L=0.0, R=0.2)
    both = boundary_potential({k: b[k] + 2 * b2[k] for k in b}, shape)
    np.testing.assert_allclose(both, phi.ravel() + 2 * boundary_potential(b2, shape), rtol=0, atol=1e-15)
    # zero voltages: exactly zero
    assert not boundary_potential(dict(T=0, B=0, L=0, R=0), shape).any()
    with pytest.raises(ValueError):
        boundary_potential(dict(T=np.nan, B=0, L=0, R=0), shape)
