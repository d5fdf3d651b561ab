"""The batched workload's cluster draws (host only)."""

import numpy as np

from betse_amd.batch import batched_worlds


def test_batched_worlds_ranges():
    w = batched_worlds(n_clusters=6, cells_range=(100, 1200), grid_range=(32, 64), seed=1)
    cells = [c.n_cells for _, c in w]
    grids = [c.X.shape[0] for _, c in w]
    assert min(cells) >= 60 and max(cells) <= 1400
    assert min(grids) >= 32 and max(grids) <= 64
    # grids grow with the radius
    order = np.argsort([p.cluster_radius for p, _ in w])
    assert np.all(np.diff(np.asarray(grids)[order]) >= 0)
    # per-instance diffusivities and pump rates stay within one decade of the base
    for p, _ in w:
        assert 10 ** -0.5 <= p.alpha_NaK / 1e-7 <= 10 ** 0.5
        assert 10 ** -0.5 <= p.mem_perms["K"] / 15e-18 <= 10 ** 0.5
