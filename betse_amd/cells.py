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

import itertools
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
        regs = [vor.regions[vor.point_region[pi]] for pi in keep]
        for reg in regs:
            if len(reg) == 0 or -1 in reg:
                raise RuntimeError("unbounded Voronoi region inside the cluster")
        ncell = len(regs)
        counts = np.fromiter((len(r) for r in regs), dtype=np.int64, count=ncell)
        ptr = np.zeros(ncell + 1, dtype=np.int64)
        ptr[1:] = np.cumsum(counts)
        nv = int(ptr[-1])
        vid = np.fromiter(itertools.chain.from_iterable(regs), dtype=np.int64, count=nv)
        cell_of = np.repeat(np.arange(ncell), counts)
        pos = np.arange(nv) - ptr[cell_of]
        first = ptr[cell_of]
        last = first + counts[cell_of] - 1

        def area_centroid(V):
            nxt = np.where(pos == counts[cell_of] - 1, first, np.arange(nv) + 1)
            x, y = V[:, 0], V[:, 1]
            xs, ys = x[nxt], y[nxt]
            cross = x * ys - xs * y
            a = 0.5 * np.add.reduceat(cross, ptr[:-1])
            cx = np.add.reduceat((x + xs) * cross, ptr[:-1]) / (6.0 * a)
            cy = np.add.reduceat((y + ys) * cross, ptr[:-1]) / (6.0 * a)
            return a, np.column_stack((cx, cy))

        a, _ = area_centroid(vor.vertices[vid])
        # make every polygon counter-clockwise
        order = np.where(a[cell_of] < 0, first + (counts[cell_of] - 1 - pos), np.arange(nv))
        vid = vid[order]
        V = vor.vertices[vid]
        _, centres = area_centroid(V)

        # shrink toward the centroid (cells.py:1009-1016)
        cc = centres[cell_of]
        Vs = p.scale_cell * (V - cc) + cc
        self.cell_verts = [Vs[ptr[c]:ptr[c + 1]] for c in range(ncell)]

        # membrane k of a cell is the edge (v[k-1], v[k]) (cells.py:1044-1089)
        prv = np.where(pos == 0, last, np.arange(nv) - 1)
        pt1, pt2 = Vs[prv], Vs
        mids = (pt1 + pt2) / 2
        dx = pt2[:, 0] - pt1[:, 0]
        dy = pt2[:, 1] - pt1[:, 1]
        lengths = np.sqrt(dx ** 2 + dy ** 2)
        tang = np.column_stack((dx, dy)) / np.sqrt(dx * dx + dy * dy)[:, None]
        nrm = np.column_stack((tang[:, 1], -tang[:, 0]))

        self.cell_centres = centres
        self.mem_mids_flat = mids
        self.mem_vects_flat = np.column_stack((mids, nrm, tang))
        self.mem_to_cells = cell_of.astype(np.int64)
        nmem = nv
        self.mem_i = np.arange(nmem)
        self.cell_i = np.arange(ncell)
        self.mem_sa = lengths * p.cell_height
        self.num_mems = counts.copy()
        self.mem_ptr = ptr.copy()
        self.cell_to_mems = _ragged([np.arange(ptr[c], ptr[c + 1]) for c in range(ncell)])

        # cell surface area: sum over each cell's membranes (cells.py:1150-1155)
        self.cell_sa = np.add.reduceat(self.mem_sa, ptr[:-1])

        rads = self.mem_mids_flat - self.cell_centres[self.mem_to_cells]
        self.R_rads = np.sqrt(rads[:, 0] ** 2 + rads[:, 1] ** 2)
        self.mem_vol = (1 / 2) * self.R_rads * self.mem_sa
        # cell_vol = M_sum_mems . mem_vol (cells.py:1385): contiguous segment sums
        self.cell_vol = np.add.reduceat(self.mem_vol, ptr[:-1])
        self.diviterm = self.cell_vol / self.cell_sa

        # gap-junction partners through the shared (unscaled) Voronoi edge
        lo = np.minimum(vid[prv], vid)
        hi = np.maximum(vid[prv], vid)
        key = lo * (len(vor.vertices) + 1) + hi
        srt = np.argsort(key, kind="stable")
        ks = key[srt]
        same = ks[1:] == ks[:-1]
        if np.any(same[1:] & same[:-1]):
            raise RuntimeError("edge shared by more than two membranes")
        ia = srt[:-1][same]
        ib = srt[1:][same]
        nn = np.arange(nmem, dtype=np.int64)
        nn[ia] = ib
        nn[ib] = ia
        self.nn_i = nn
        self.mem_nn = np.column_stack((self.mem_i, nn))
        self.bflags_mems = np.nonzero(nn == self.mem_i)[0]
        self.bflags_cells = self.mem_to_cells[self.bflags_mems]
        self.cell_nn_i = np.column_stack((self.mem_to_cells, self.mem_to_cells[nn]))
        self._set_cell_nn()
        self.gj_len = p.cell_space                     # cells.py:1004
        self.gj_default_weights = np.ones(nmem)        # cells.py:526

    def _set_cell_nn(self) -> None:
        """Neighbour cells through gap-junction partners (cells.py:1536-1564), as
        CSR (``cell_nn_ptr``, ``cell_nn_idx``) and as ragged ``cell_nn``."""
        m = self.mem_i
        other = self.mem_to_cells[self.nn_i]
        ok = (self.nn_i != m) & (other != self.mem_to_cells)
        src = self.mem_to_cells[ok]
        dst = other[ok]
        ncell = len(self.cell_i)
        cnt = np.bincount(src, minlength=ncell)
        self.cell_nn_ptr = np.zeros(ncell + 1, dtype=np.int64)
        self.cell_nn_ptr[1:] = np.cumsum(cnt)
        self.cell_nn_idx = dst.astype(np.int64)       # src is sorted: membranes are cell-ordered
        self.num_nn = cnt
        self.cell_nn = _ragged([list(self.cell_nn_idx[self.cell_nn_ptr[c]:self.cell_nn_ptr[c + 1]])
                                for c in range(ncell)])

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
        nz = np.nonzero(counts)[0]
        if len(nz):
            # sum of mem_sa of the square's membranes in membrane order (cells.py:1810-1813)
            self.memSa_per_envSquare[nz] = np.add.reduceat(self.mem_sa[self.sq_mems], self.sq_ptr[nz])
            self.mems_per_envSquare[nz] = counts[nz]
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
        self._set_cell_nn()
        return self
