"""bt_validate: the geometry checks bt_create runs before any kernel gathers
through the index arrays, on host arrays (no GPU needed).

A caller in another language that passes a bad index array gets BT_EINVAL and
a message instead of an out-of-bounds device access."""

import os

import numpy as np
import pytest

from betse_amd import _native
from betse_amd.assemble import batch_geometry, make_clusters, make_params, validate
from betse_amd.cells import Cells
from betse_amd.parameters import Parameters
from betse_amd.sim import HostState, SimPhase, SimPhaseKind

pytestmark = pytest.mark.skipif(not os.path.exists(_native.LIB_PATH), reason="library not built")


@pytest.fixture(scope="module")
def batch():
    worlds = []
    for seed, radius in ((0, 30e-6), (3, 40e-6)):
        p = Parameters.preset("tiny")
        p.seed = seed
        p.cluster_radius = radius
        worlds.append((p, Cells.make_world(p)))
    states = []
    for p, c in worlds:
        st = HostState()
        ph = SimPhase(SimPhaseKind.INIT, p, c)
        st.init_core(ph)
        st.init_dynamics(ph)
        states.append(st)
    params = [p for p, _ in worlds]
    cells = [c for _, c in worlds]
    geo, (coff, moff, soff) = batch_geometry(cells)
    prm = make_params(states, params, cells, "fp64", True)
    cls = make_clusters(states, params, cells, soff)
    return prm, cls, geo, (coff, moff, soff)


def corrupt(geo, name, fn):
    g = {k: v.copy() for k, v in geo.items()}
    fn(g[name])
    return g


def test_valid_batch_passes(batch):
    prm, cls, geo, _ = batch
    rc, msg = validate(prm, cls, geo)
    assert rc == _native.BT_OK, msg


@pytest.mark.parametrize("name,fn,expect", [
    ("nn_mem", lambda a: a.__setitem__(5, 10 ** 7), "nn_mem out of range"),
    ("nn_mem", lambda a: a.__setitem__(0, -1), "nn_mem out of range"),
    ("nn_cell", lambda a: a.__setitem__(3, a[3] + 1), "nn_cell is not the cell of nn_mem"),
    ("mem_sq", lambda a: a.__setitem__(7, 10 ** 7), "mem_sq out of range"),
    ("mem_cell", lambda a: a.__setitem__(2, a[2] + 1), "mem_cell is not the CSR owner"),
    ("mem_pos", lambda a: a.__setitem__(slice(0, 2), a[[1, 0]]), "mem_pos is not the inverse"),
    ("sq_mems", lambda a: a.__setitem__(0, a[1]), "sq_mems"),
    ("sq_cluster", lambda a: a.__setitem__(0, 1), "sq_cluster does not match"),
    ("tile_cluster", lambda a: a.__setitem__(0, 5), "tile_cluster out of range"),
    ("tile_ptr", lambda a: a.__setitem__(1, a[1] + 40), "exceeds the tile limits"),
    ("mem_ptr", lambda a: a.__setitem__(1, 0), "a cell has no membranes"),
])
def test_bad_index_array_is_rejected(batch, name, fn, expect):
    prm, cls, geo, _ = batch
    rc, msg = validate(prm, cls, corrupt(geo, name, fn))
    assert rc == _native.BT_EINVAL
    assert expect in msg, msg


def test_cross_cluster_links_are_rejected(batch):
    """A gap-junction partner or a square in another cluster's range is in
    bounds for the arrays but wrong for the physics: rejected as well."""
    prm, cls, geo, (coff, moff, soff) = batch
    m = int(moff[1])                       # first membrane of cluster 1

    def swap_partner(g):
        g["nn_mem"][m] = 0
        g["nn_cell"][m] = g["mem_cell"][0]
    g = {k: v.copy() for k, v in geo.items()}
    swap_partner(g)
    rc, msg = validate(prm, cls, g)
    assert rc == _native.BT_EINVAL and "another cluster" in msg, msg
    g = corrupt(geo, "mem_sq", lambda a: a.__setitem__(m, 0))
    rc, msg = validate(prm, cls, g)
    assert rc == _native.BT_EINVAL and ("another cluster" in msg or "sq_mems" in msg), msg
