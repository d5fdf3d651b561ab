"""Cell cluster geometry: the ``Cells`` contract the electrodiffusion step reads.

The reference builds its world in ``Cells.make_world`` (``betse/science/cells.py:414-526``)
from a DEC mesh. This module builds a cluster from a jittered hexagonal
lattice instead (the seed geometry is not on the step's path) and then derives
every array the step consumes with the reference's rules:

* polygons shrunk toward the centroid by ``scale_cell`` (``cells.py:1009-1016``),
  membranes are the edges ``(v[i-1], v[i])`` of each CCW polygon with normal
  ``(t_y, -t_x)`` (``cells.py:1044-1089``), flattened cell by cell so each
  cell's membranes are one contiguous range (``math/toolbox.py:21-64``);
* ``mem_sa = length * cell_height`` (``cells.py:1110``), ``R_rads`` centre to
  midpoint (``cells.py:1170-1173``), ``mem_vol = R_rads*mem_sa/2``,
  ``cell_vol = sum(mem_vol)``, ``diviterm = cell_vol/cell_sa`` (``cells.py:1378-1390``);
* gap-junction partners ``nn_i`` (``cells.py:1403-1446,1498-1533``): a membrane
  paired with the membrane of the neighbouring cell across the same Voronoi
  edge, self-paired (and flagged in ``bflags_mems``) on the cluster boundary;
* a square environment grid of ``grid_size``x``grid_size`` centres,
  ``map_mem2ecm``/``map_cell2ecm`` by nearest centre (``cells.py:1754-1758``),
  ``memSa_per_envSquare`` and ``envInds_inClust`` (``cells.py:1800-1818``),
  tight-junction target squares (``cells.py:1824-1833``).

Beyond the reference's arrays it provides the device-side layouts: the cell
CSR ``mem_ptr`` and the grid-side CSR ``sq_ptr``/``sq_mems`` (the sparse form of
``M_divmap_mem2ecm``, sorted by square then membrane index).
"""

from __future__ import annotations

import math
from typing import Optional

import numpy as np
from scipy.spatial import Voronoi

from .parameters import Parameters


def _polygon_area_centroid(v: np.ndarray):
    x, y = v[:, 0], v[:, 1]
    xs, ys = np.roll(x, -1), np.roll(y, -1)
    cross = x * ys - xs * y
    a = 0.5 * cross.sum()
    cx = ((x + xs) * cross).sum() / (6.0 * a)
    cy = ((y + ys) * cross).sum() / (6.0 * a)
    return a, np.array([cx, cy])


def hex_lattice_points(radius: float, spacing: float, jitter: float, seed: int) -> np.ndarray:
    """Jittered hexagonal lattice covering a disc of ``radius`` plus three rows.

    Rows are ``spacing*sqrt(3)/2`` apart and alternate rows shift by
    ``spacing/2``; each point moves by ``U(-jitter, +jitter)*spacing`` in x and
    y, drawn from NumPy's legacy ``RandomState(seed)`` stream (the generator the
    reference uses for its lattice disorder, ``cells.py:794-795``).
    """
    dy = spacing * math.sqrt(3.0) / 2.0
    extent = radius + 3.0 * spacing
    ny = int(math.ceil(extent / dy))
    nx = int(math.ceil(extent / spacing)) + 1
    js = np.arange(-ny, ny + 1)
    iis = np.arange(-nx, nx + 1)
    J, I = np.meshgrid(js, iis, indexing="ij")
    x = I * spacing + (np.abs(J) % 2) * (spacing / 2.0)
    y = J * dy
    pts = np.column_stack((x.ravel(), y.ravel())).astype(np.float64)
    rng = np.random.RandomState(seed)
    pts = pts + spacing * jitter * (2.0 * rng.rand(pts.shape[0], 2) - 1.0)
    return pts


def _ragged(rows) -> np.ndarray:
    """1-D object array of per-cell lists (what the reference stores as ragged arrays)."""
    out = np.empty(len(rows), dtype=object)
    for i, r in enumerate(rows):
        out[i] = r
    return out


class Cells:
    """Geometry arrays of one cluster and its environment grid.

    Attribute names follow the reference ``Cells`` object (App. A of the
    contract): ``cell_centres, mem_mids_flat, mem_vects_flat, mem_sa, mem_to_cells,
    num_mems, R_rads, mem_vol, cell_vol, cell_sa, diviterm, nn_i, mem_nn,
    bflags_mems, bflags_cells, cell_nn, gj_len, gj_default_weights, delta, X, Y,
    xypts, map_mem2ecm, map_cell2ecm, memSa_per_envSquare, envInds_inClust``.
    """

    # ------------------------------------------------------------------
    @classmethod
    def make_world(cls, p: Parameters, points: Optional[np.ndarray] = None) -> "Cells":
        self = cls()
        if points is None:
            points = hex_lattice_points(p.cluster_radius, p.lattice_spacing, p.lattice_jitter, p.seed)
        self._build_cells(p, points)
        self._build_grid(p)
        return self

    # ------------------------------------------------------------------
    def _build_cells(self, p: Parameters, points: np.ndarray) -> None:
        vor = Voronoi(points)
        inside = np.hypot(points[:, 0], points[:, 1]) <= p.cluster_radius
        keep = np.nonzero(inside)[0]
        polys = []
        regions_v = []
        for pi in keep:
            reg = vor.regions[vor.point_region[pi]]
            if len(reg) == 0 or -1 in reg:
                raise RuntimeError("unbounded Voronoi region inside the cluster")
            v = vor.vertices[reg]
            a, _ = _polygon_area_centroid(v)
            if a < 0:               # make every polygon counter-clockwise
                reg = reg[::-1]
                v = v[::-1]
            polys.append(v)
            regions_v.append(list(reg))
        ncell = len(polys)

        centres = np.zeros((ncell, 2))
        for c, v in enumerate(polys):
            _, centres[c] = _polygon_area_centroid(v)

        # shrink toward the centroid (cells.py:1009-1016)
        cell_verts = [p.scale_cell * (v - centres[c]) + centres[c] for c, v in enumerate(polys)]

        mids, nrm, tng, lengths, m2c, edge_key = [], [], [], [], [], []
        for c, v in enumerate(cell_verts):
            reg = regions_v[c]
            n = len(v)
            for i in range(n):
                pt1 = v[i - 1]
                pt2 = v[i]
                mid = (pt1 + pt2) / 2
                lgth = np.sqrt((pt2[0] - pt1[0]) ** 2 + (pt2[1] - pt1[1]) ** 2)
                tang_a = pt2 - pt1
                tang = tang_a / np.linalg.norm(tang_a)
                mids.append(mid)
                nrm.append([tang[1], -tang[0]])
                tng.append(tang)
                lengths.append(lgth)
                m2c.append(c)
                edge_key.append((min(reg[i - 1], reg[i]), max(reg[i - 1], reg[i])))

        self.cell_centres = centres
        self.cell_verts = cell_verts
        self.mem_mids_flat = np.asarray(mids)
        nrm = np.asarray(nrm)
        tng = np.asarray(tng)
        self.mem_vects_flat = np.column_stack((self.mem_mids_flat, nrm, tng))
        self.mem_to_cells = np.asarray(m2c, dtype=np.int64)
        nmem = len(m2c)
        self.mem_i = np.arange(nmem)
        self.cell_i = np.arange(ncell)
        self.mem_sa = np.asarray(lengths) * p.cell_height
        self.num_mems = np.bincount(self.mem_to_cells, minlength=ncell).astype(np.int64)
        self.mem_ptr = np.zeros(ncell + 1, dtype=np.int64)
        self.mem_ptr[1:] = np.cumsum(self.num_mems)
        self.cell_to_mems = _ragged([np.arange(self.mem_ptr[c], self.mem_ptr[c + 1]) for c in range(ncell)])

        # cell surface area: python sum over each cell's membranes (cells.py:1150-1155)
        self.cell_sa = np.asarray([sum(self.mem_sa[g]) for g in self.cell_to_mems])

        rads = self.mem_mids_flat - self.cell_centres[self.mem_to_cells]
        self.R_rads = np.sqrt(rads[:, 0] ** 2 + rads[:, 1] ** 2)
        self.mem_vol = (1 / 2) * self.R_rads * self.mem_sa
        # cell_vol = M_sum_mems . mem_vol (cells.py:1385): contiguous segment sums
        self.cell_vol = np.add.reduceat(self.mem_vol, self.mem_ptr[:-1])
        self.diviterm = self.cell_vol / self.cell_sa

        # gap-junction partners through the shared (unscaled) Voronoi edge
        partner = {}
        for m, key in enumerate(edge_key):
            partner.setdefault(key, []).append(m)
        nn = np.arange(nmem, dtype=np.int64)
        for key, ms in partner.items():
            if len(ms) == 2:
                a, b = ms
                nn[a] = b
                nn[b] = a
            elif len(ms) > 2:
                raise RuntimeError("edge shared by more than two membranes")
        self.nn_i = nn
        self.mem_nn = np.column_stack((self.mem_i, nn))
        self.bflags_mems = np.nonzero(nn == self.mem_i)[0]
        self.bflags_cells = self.mem_to_cells[self.bflags_mems]
        self.cell_nn_i = np.column_stack((self.mem_to_cells, self.mem_to_cells[nn]))
        cell_nn = []
        for c in range(ncell):
            s = []
            for m in self.cell_to_mems[c]:
                j = nn[m]
                if j != m and self.mem_to_cells[j] != c:
                    s.append(int(self.mem_to_cells[j]))
            cell_nn.append(s)
        self.cell_nn = _ragged(cell_nn)
        self.num_nn = np.asarray([len(s) for s in cell_nn])
        self.gj_len = p.cell_space                     # cells.py:1004
        self.gj_default_weights = np.ones(nmem)        # cells.py:526

    # ------------------------------------------------------------------
    def _build_grid(self, p: Parameters) -> None:
        half = p.world_scale * p.cluster_radius + 2.0 * p.lattice_spacing
        self.xmin, self.xmax = -half, half
        self.ymin, self.ymax = -half, half
        G = int(p.grid_size)
        self.grid_n = G
        self.delta = (self.xmax - self.xmin) / G                 # cells.py:1607
        xv = np.linspace(self.xmin, self.xmax, G + 1)
        yv = np.linspace(self.ymin, self.ymax, G + 1)
        vx, vy = np.meshgrid(xv, yv)
        # grid-square centres (finitediff.py:188-189)
        self.X = (vx[0:-1, 0:-1] + vx[0:-1, 1:]) / 2
        self.Y = (vy[0:-1, 0:-1] + vy[1:, 0:-1]) / 2
        self.xypts = np.column_stack((self.X.ravel(), self.Y.ravel()))
        x_c = self.X[0, :]
        y_c = self.Y[:, 0]

        def nearest(pts):
            ci = np.clip(np.searchsorted(xv, pts[:, 0]) - 1, 0, G - 1)
            ri = np.clip(np.searchsorted(yv, pts[:, 1]) - 1, 0, G - 1)
            # resolve to the nearest centre among the two candidates
            for d in (-1, 1):
                cj = np.clip(ci + d, 0, G - 1)
                better = np.abs(x_c[cj] - pts[:, 0]) < np.abs(x_c[ci] - pts[:, 0])
                ci = np.where(better, cj, ci)
                rj = np.clip(ri + d, 0, G - 1)
                better = np.abs(y_c[rj] - pts[:, 1]) < np.abs(y_c[ri] - pts[:, 1])
                ri = np.where(better, rj, ri)
            return (ri * G + ci).astype(np.int64)

        self.map_cell2ecm = nearest(self.cell_centres)
        self.map_mem2ecm = nearest(self.mem_mids_flat)
        if (self.map_mem2ecm // G == 0).any() or (self.map_mem2ecm // G == G - 1).any() \
                or (self.map_mem2ecm % G == 0).any() or (self.map_mem2ecm % G == G - 1).any():
            raise RuntimeError("cluster touches the fixed-concentration world boundary")

        n_sq = G * G
        self.memSa_per_envSquare = np.zeros(n_sq)
        self.mems_per_envSquare = np.zeros(n_sq)
        order = np.argsort(self.map_mem2ecm, kind="stable")
        counts = np.bincount(self.map_mem2ecm, minlength=n_sq)
        self.sq_ptr = np.zeros(n_sq + 1, dtype=np.int64)
        self.sq_ptr[1:] = np.cumsum(counts)
        self.sq_mems = order.astype(np.int64)
        for k in np.nonzero(counts)[0]:
            lst = self.sq_mems[self.sq_ptr[k]:self.sq_ptr[k + 1]]
            self.memSa_per_envSquare[k] = np.sum(self.mem_sa[lst])     # cells.py:1810-1813
            self.mems_per_envSquare[k] = len(lst)
        self.envInds_inClust = np.nonzero(counts)[0]

        # tight-junction targets (cells.py:1824-1833)
        bcells = np.unique(self.bflags_cells)
        all_bound = np.concatenate([self.cell_to_mems[c] for c in bcells]) if len(bcells) else np.zeros(0, np.int64)
        neigh = np.unique(np.concatenate([np.asarray(self.cell_nn[c], dtype=np.int64) for c in bcells])) \
            if len(bcells) else np.zeros(0, np.int64)
        interior = np.concatenate([self.cell_to_mems[c] for c in neigh]) if len(neigh) else np.zeros(0, np.int64)
        self.all_bound_mem_inds = self.map_mem2ecm[all_bound.astype(np.int64)]
        self.interior_bound_mem_inds = self.map_mem2ecm[interior.astype(np.int64)]
        self.ecm_inds_bound_cell = self.map_cell2ecm[self.bflags_cells]

    # ------------------------------------------------------------------
    @property
    def n_cells(self) -> int:
        return len(self.cell_i)

    @property
    def n_mems(self) -> int:
        return len(self.mem_i)

    @property
    def n_squares(self) -> int:
        return self.X.size

    # ------------------------------------------------------------------
    _SAVE_KEYS = ("cell_centres", "mem_mids_flat", "mem_vects_flat", "mem_to_cells", "mem_sa",
                  "num_mems", "mem_ptr", "cell_sa", "R_rads", "mem_vol", "cell_vol", "diviterm",
                  "nn_i", "bflags_mems", "bflags_cells", "gj_default_weights", "X", "Y", "xypts",
                  "map_cell2ecm", "map_mem2ecm", "memSa_per_envSquare", "mems_per_envSquare",
                  "sq_ptr", "sq_mems", "envInds_inClust", "all_bound_mem_inds",
                  "interior_bound_mem_inds", "ecm_inds_bound_cell")
    _SAVE_SCALARS = ("delta", "gj_len", "xmin", "xmax", "ymin", "ymax", "grid_n")

    def save(self, path: str) -> None:
        """Write the contract arrays to an ``.npz`` (no pickled objects)."""
        d = {k: np.asarray(getattr(self, k)) for k in self._SAVE_KEYS}
        for k in self._SAVE_SCALARS:
            d[k] = np.asarray(getattr(self, k))
        np.savez_compressed(path, **d)

    @classmethod
    def load(cls, path: str) -> "Cells":
        self = cls()
        with np.load(path, allow_pickle=False) as z:
            for k in cls._SAVE_KEYS:
                setattr(self, k, z[k])
            for k in cls._SAVE_SCALARS:
                setattr(self, k, z[k].item())
        ncell = len(self.cell_centres)
        self.cell_i = np.arange(ncell)
        self.mem_i = np.arange(len(self.mem_to_cells))
        self.cell_to_mems = _ragged([np.arange(self.mem_ptr[c], self.mem_ptr[c + 1]) for c in range(ncell)])
        self.mem_nn = np.column_stack((self.mem_i, self.nn_i))
        cell_nn = []
        for c in range(ncell):
            s = [int(self.mem_to_cells[j]) for m in self.cell_to_mems[c]
                 for j in (self.nn_i[m],) if j != m and self.mem_to_cells[j] != c]
            cell_nn.append(s)
        self.cell_nn = _ragged(cell_nn)
        self.num_nn = np.asarray([len(s) for s in cell_nn])
        return self
