"""Geometry contract of the cluster generator (reference rules cited in betse_amd/cells.py)."""

import os

import numpy as np
import pytest

from betse_amd.cells import Cells, hex_lattice_points
from betse_amd.parameters import Parameters

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module", params=["tiny", "small", "random"])
def world(request):
    if request.param == "random":
        p = Parameters.preset("tiny")
        rng = np.random.RandomState(11)
        pts = rng.uniform(-60e-6, 60e-6, (400, 2))
        return p, Cells.make_world(p, points=pts)
    p = Parameters.preset(request.param)
    return p, Cells.make_world(p)


def test_csr_and_ordering(world):
    p, c = world
    assert np.all(np.diff(c.mem_to_cells) >= 0)
    assert c.mem_ptr[0] == 0 and c.mem_ptr[-1] == c.n_mems
    np.testing.assert_array_equal(np.diff(c.mem_ptr), c.num_mems)
    assert c.num_mems.min() >= 3


def test_partner_symmetry_and_distance(world):
    p, c = world
    nn = c.nn_i
    np.testing.assert_array_equal(nn[nn], np.arange(c.n_mems))
    inner = nn != np.arange(c.n_mems)
    assert np.all(c.mem_to_cells[nn[inner]] != c.mem_to_cells[inner])
    # partners sit across the shrunken gap: within the reference ball radius
    d = np.hypot(*(c.mem_mids_flat[inner] - c.mem_mids_flat[nn[inner]]).T)
    assert d.max() < 3.0 * (1 - p.scale_cell) * p.cell_radius * 2
    # partner normals are anti-parallel
    n1 = c.mem_vects_flat[inner, 2:4]
    n2 = c.mem_vects_flat[nn[inner], 2:4]
    np.testing.assert_allclose(np.sum(n1 * n2, axis=1), -1.0, atol=1e-2)
    np.testing.assert_array_equal(c.bflags_mems, np.nonzero(~inner)[0])


def test_outward_normals_and_volumes(world):
    p, c = world
    rad = c.mem_mids_flat - c.cell_centres[c.mem_to_cells]
    assert np.all(np.sum(rad * c.mem_vects_flat[:, 2:4], axis=1) > 0)
    np.testing.assert_allclose(c.mem_vol, 0.5 * c.R_rads * c.mem_sa)
    np.testing.assert_allclose(c.cell_vol, np.add.reduceat(c.mem_vol, c.mem_ptr[:-1]))
    np.testing.assert_allclose(c.diviterm, c.cell_vol / c.cell_sa)
    # wedges with the perpendicular distance (apothem) tile the polygon exactly;
    # the reference's R_rads (centre to midpoint) is >= the apothem
    area = np.array([0.5 * abs(np.sum(v[:, 0] * np.roll(v[:, 1], -1) - np.roll(v[:, 0], -1) * v[:, 1]))
                     for v in c.cell_verts])
    apothem = np.sum(rad * c.mem_vects_flat[:, 2:4], axis=1)
    wedge = np.add.reduceat(0.5 * apothem * c.mem_sa, c.mem_ptr[:-1])
    np.testing.assert_allclose(wedge, area * p.cell_height, rtol=1e-9)
    assert np.all(c.cell_vol >= wedge * (1 - 1e-12))


def test_grid_maps(world):
    p, c = world
    G = c.grid_n
    d = c.xypts[c.map_mem2ecm] - c.mem_mids_flat
    assert np.all(np.abs(d) <= c.delta / 2 * (1 + 1e-9))
    counts = np.bincount(c.map_mem2ecm, minlength=G * G)
    np.testing.assert_array_equal(np.diff(c.sq_ptr), counts)
    for k in c.envInds_inClust[:50]:
        lst = c.sq_mems[c.sq_ptr[k]:c.sq_ptr[k + 1]]
        assert np.all(np.diff(lst) > 0)
        np.testing.assert_allclose(c.memSa_per_envSquare[k], c.mem_sa[lst].sum())


def test_small_cluster_size():
    p = Parameters.preset("small")
    c = Cells.make_world(p)
    assert 300 < c.n_cells < 400 and 1800 < c.n_mems < 2400


def test_save_load_roundtrip(tmp_path):
    p = Parameters.preset("tiny")
    c = Cells.make_world(p)
    f = str(tmp_path / "w.npz")
    c.save(f)
    d = Cells.load(f)
    for k in Cells._SAVE_KEYS:
        np.testing.assert_array_equal(getattr(c, k), getattr(d, k))
    assert d.delta == c.delta


def test_committed_fixture_matches_generator():
    """The committed C1 world is what the generator builds from numpy seed 0."""
    f = os.path.join(ROOT, "tests", "golden", "world_small.npz")
    d = Cells.load(f)
    c = Cells.make_world(Parameters.preset("small"))
    for k in ("mem_to_cells", "nn_i", "map_mem2ecm", "sq_mems"):
        np.testing.assert_array_equal(getattr(c, k), getattr(d, k))
    for k in ("mem_sa", "cell_vol", "diviterm", "memSa_per_envSquare"):
        np.testing.assert_allclose(getattr(c, k), getattr(d, k), rtol=1e-14)


def test_lattice_jitter_bounds():
    pts0 = hex_lattice_points(50e-6, 10e-6, 0.0, 0)
    pts = hex_lattice_points(50e-6, 10e-6, 0.1, 0)
    assert np.abs(pts - pts0).max() <= 1e-6 * (1 + 1e-12)
