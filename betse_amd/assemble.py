"""Host-side assembly of the arrays and structures ``bt_create`` takes.

One simulation is a batch of one cluster. Cells, membranes and grid squares of
every cluster are concatenated with global indices; the per-cluster table
(``bt_cluster``) holds each grid's shape, offset and the constants that may
differ between independent clusters. Everything here is NumPy and ctypes:
no device work, so the same arrays can be checked by ``bt_validate`` on a
machine without a GPU.

The arrays follow the Cells contract of the reference (``cells.py``:
``mem_to_cells``, ``nn_i``, ``map_mem2ecm``, ``mem_sa``, ``cell_vol`` ...) and
the state of ``Simulator.init_core``/``init_dynamics`` (``sim.py:452-1012``).
"""

from __future__ import annotations

from typing import Dict, List, Sequence, Tuple

import numpy as np

from . import _native
from .channels import CHANNEL_KINDS
from .exceptions import BetseSimException


def make_tiles(mem_ptr: np.ndarray, tile_m: int = _native.BT_TILE_MEMBRANES,
               tile_c: int = _native.BT_TILE_CELLS) -> np.ndarray:
    """Greedy packing of consecutive whole cells into tiles of <= tile_m
    membranes and <= tile_c cells."""
    n = len(mem_ptr) - 1
    counts = np.diff(mem_ptr)
    if counts.max(initial=0) > tile_m:
        raise BetseSimException(f"a cell has more than {tile_m} membranes")
    if counts.min(initial=1) < 1:
        raise BetseSimException("a cell has no membranes")
    tiles = [0]
    acc = 0
    for c in range(n):
        if acc + counts[c] > tile_m or c - tiles[-1] == tile_c:
            tiles.append(c)
            acc = 0
        acc += counts[c]
    tiles.append(n)
    return np.asarray(tiles, dtype=np.int32)


# geometry arrays of bt_buffers (int32 unless noted)
GEOMETRY_INT = ("mem_ptr", "mem_cell", "nn_mem", "nn_cell", "mem_sq", "sq_ptr", "sq_mems", "mem_pos",
                "tile_ptr", "tile_cluster", "sq_cluster")
GEOMETRY_F64 = ("mem_sa", "cell_vol", "diviterm", "cell_sa", "mem_nx", "mem_ny")


def batch_geometry(cells_list: Sequence) -> Tuple[Dict[str, np.ndarray], Tuple[np.ndarray, np.ndarray, np.ndarray]]:
    """Concatenated geometry of the clusters and the (cell, membrane, square)
    offsets of each. ``sq_has_mem`` is uint8; ``cell_cluster`` maps cells to
    clusters (for the init residual)."""
    C = [c.n_cells for c in cells_list]
    M = [c.n_mems for c in cells_list]
    G = [c.X.size for c in cells_list]
    coff = np.concatenate([[0], np.cumsum(C)]).astype(np.int64)
    moff = np.concatenate([[0], np.cumsum(M)]).astype(np.int64)
    soff = np.concatenate([[0], np.cumsum(G)]).astype(np.int64)

    def cat(f):
        return np.concatenate([f(i, c) for i, c in enumerate(cells_list)])

    for c in cells_list:
        if len(c.sq_ptr) != c.X.size + 1:
            raise BetseSimException("grid-side CSR does not match the grid")
    mem_ptr = np.concatenate([cells_list[0].mem_ptr[:1]] + [c.mem_ptr[1:] + moff[i] for i, c in enumerate(cells_list)])
    sq_ptr = np.concatenate([cells_list[0].sq_ptr[:1]] + [c.sq_ptr[1:] + moff[i] for i, c in enumerate(cells_list)])
    sq_mems = cat(lambda i, c: c.sq_mems + moff[i])
    pos = np.empty(int(moff[-1]), dtype=np.int64)
    pos[sq_mems] = np.arange(int(moff[-1]))
    tiles, tcl = [np.zeros(1, dtype=np.int64)], []
    for i, c in enumerate(cells_list):
        t = make_tiles(c.mem_ptr)
        tiles.append(t[1:].astype(np.int64) + coff[i])
        tcl.append(np.full(len(t) - 1, i, dtype=np.int32))
    g = {
        "mem_ptr": mem_ptr, "mem_cell": cat(lambda i, c: c.mem_to_cells + coff[i]),
        "nn_mem": cat(lambda i, c: c.nn_i + moff[i]),
        "nn_cell": cat(lambda i, c: c.mem_to_cells[c.nn_i] + coff[i]),
        "mem_sq": cat(lambda i, c: c.map_mem2ecm + soff[i]),
        "sq_ptr": sq_ptr, "sq_mems": sq_mems, "mem_pos": pos,
        "tile_ptr": np.concatenate(tiles), "tile_cluster": np.concatenate(tcl),
        "sq_cluster": cat(lambda i, c: np.full(c.X.size, i, dtype=np.int32)),
        "mem_sa": cat(lambda i, c: c.mem_sa), "cell_vol": cat(lambda i, c: c.cell_vol),
        "diviterm": cat(lambda i, c: c.diviterm), "cell_sa": cat(lambda i, c: c.cell_sa),
        "mem_nx": cat(lambda i, c: c.mem_vects_flat[:, 2]), "mem_ny": cat(lambda i, c: c.mem_vects_flat[:, 3]),
    }
    for k in GEOMETRY_INT:
        g[k] = np.ascontiguousarray(g[k], dtype=np.int32)
    for k in GEOMETRY_F64:
        g[k] = np.ascontiguousarray(g[k], dtype=np.float64)
    g["sq_has_mem"] = (np.diff(sq_ptr) > 0).astype(np.uint8)
    g["cell_cluster"] = np.concatenate([np.full(n, i, dtype=np.int32) for i, n in enumerate(C)])
    return g, (coff, moff, soff)


def batch_state(states: Sequence, cells_list: Sequence, sim_phase: bool) -> Dict[str, np.ndarray]:
    """Concatenated FP64 parameters and state of the clusters' host states
    (``HostState`` after ``init_core``/``init_dynamics``). ``sim_phase``
    selects the SIM phase's membrane diffusivities ``Dm_scheduled + Dm_base``
    (``makeAllChanges``, ``tishandler.py:1321-1332``)."""
    nch = len(states[0].channel_specs)
    h = {
        "gjopen": np.concatenate([st.gj_block * np.ones(st.mdl) * c.gj_default_weights
                                  for st, c in zip(states, cells_list)]),
        "Dm": np.concatenate([(st.Dm_scheduled + st.Dm_base) if sim_phase else st.Dm_cells for st in states], axis=1),
        "D_env": np.concatenate([st.D_env * st.TJ_modulator for st in states], axis=1),
        "c_cell": np.concatenate([st.cc_cells for st in states], axis=1),
        "c_lag": np.concatenate([st.c_lag for st in states], axis=1),
        "vm_cell": np.concatenate([st.vm_cell for st in states]),
        "c_env": np.concatenate([st.cc_env for st in states], axis=1),
        "E_x": np.concatenate([st.E_env_x.ravel() for st in states]),
        "E_y": np.concatenate([st.E_env_y.ravel() for st in states]),
        "v_env": np.concatenate([st.v_env for st in states]),
        "gates": np.concatenate([st.chan_gates for st in states], axis=2) if nch else np.zeros((1, 2, 1)),
        "chan_mod": np.concatenate([st.chan_mod for st in states], axis=1) if nch else np.ones((1, 1)),
        "Phi_b": np.concatenate([st.Phi_b for st in states]),
    }
    if states[0].network is not None:
        h["c_net"] = np.concatenate([st.c_net for st in states], axis=1)
    return {k: np.ascontiguousarray(v, dtype=np.float64) for k, v in h.items()}


# Parameters fields that bt_params takes once for the whole batch
SHARED_FIELDS = ("dt", "T", "F", "R", "q", "kb", "cm", "tm", "gj_surface", "deltaGATP", "cATP", "cADP", "cPi",
                 "KmNK_Na", "KmNK_K", "KmNK_ATP", "alpha_Ca", "KmCa_Ca", "KmCa_ATP", "cell_height", "eo", "er",
                 "cell_radius", "true_cell_size", "smooth_cells", "mu", "NaKATP_block")


def check_compatible(states: Sequence, params: Sequence, cells_list: Sequence) -> None:
    """Raise unless the clusters share what ``bt_params`` holds once per batch:
    the ions, the channel models, the reaction network, the boundary
    concentrations, gap-junction diffusivities and length, and every constant
    of ``SHARED_FIELDS``. What may differ per cluster: the geometry and grid,
    Dm, D_env, ko_env and the NaK-ATPase rate (``bt_cluster``)."""
    s0, p0, c0 = states[0], params[0], cells_list[0]
    for st, p, c in zip(states, params, cells_list):
        if st.names != s0.names:
            raise BetseSimException("clusters of a batch need the same ions")
        for f in SHARED_FIELDS:
            if getattr(p, f) != getattr(p0, f):
                raise BetseSimException(f"clusters of a batch need the same {f}")
        if st.T != s0.T or st.NaKATP_block != s0.NaKATP_block:
            raise BetseSimException("clusters of a batch need the same temperature and NaK block")
        if tuple(st.channel_specs) != tuple(s0.channel_specs):
            raise BetseSimException("clusters of a batch need the same channels")
        if list(st.c_env_bound) != list(s0.c_env_bound):
            raise BetseSimException("clusters of a batch need the same boundary concentrations")
        for i in range(len(st.names)):
            if not np.all(st.D_gj[i] == s0.D_gj[i][0]):
                raise BetseSimException("clusters of a batch need one D_gj per ion")
        if c.gj_len != c0.gj_len:
            raise BetseSimException("clusters of a batch need the same gap-junction length")
        if (st.network is None) != (s0.network is None) or (
                st.network is not None and bytes(st.network.pack(p.R, st.T)) != bytes(s0.network.pack(p0.R, s0.T))):
            raise BetseSimException("clusters of a batch need the same reaction network")
        if (p.alpha_NaK > 0.0) != (p0.alpha_NaK > 0.0):
            raise BetseSimException("the NaK pump must be on in all clusters or in none")


def make_params(states: Sequence, params: Sequence, cells_list: Sequence, flux_mode: str,
                channels_active: bool) -> "_native.BtParams":
    """``bt_params`` of a batch (constants of the first cluster: the batch has
    been checked compatible)."""
    from .sim import gaussian_taps
    s0, p0 = states[0], params[0]
    prm = _native.BtParams()
    prm.n_cells = int(sum(c.n_cells for c in cells_list))
    prm.n_mems = int(sum(c.n_mems for c in cells_list))
    prm.n_squares = int(sum(c.X.size for c in cells_list))
    prm.n_ions = len(s0.names)
    prm.i_na = s0.names.index("Na") if ("Na" in s0.names and "K" in s0.names and p0.alpha_NaK > 0.0) else -1
    prm.i_k = s0.names.index("K") if prm.i_na >= 0 else -1
    prm.i_ca = s0.names.index("Ca") if ("Ca" in s0.names and p0.alpha_Ca > 0.0) else -1
    prm.flux_mode = _native.BT_FLUX_FP32 if flux_mode == "fp32" else _native.BT_FLUX_FP64
    prm.n_chan = len(s0.channel_specs) if channels_active else 0
    for k, spec in enumerate(s0.channel_specs):
        prm.chan_kind[k] = _native.BT_CHAN_KINDS[spec.kind]
        prm.chan_ion[k] = s0.names.index(CHANNEL_KINDS[spec.kind][0])
        prm.chan_max_dm[k] = float(spec.max_dm)
    prm.chan_dt = p0.dt * 1.0e3        # time_unit of the supported channel models (ms)
    for i in range(prm.n_ions):
        prm.z[i] = float(s0.zs[i])
        if not np.all(s0.D_gj[i] == s0.D_gj[i][0]):
            raise BetseSimException("the GPU step needs a uniform D_gj per ion")
        prm.d_gj[i] = float(s0.D_gj[i][0])
        prm.c_bound[i] = float(s0.c_env_bound[i])
    prm.dt, prm.T, prm.T_gj = p0.dt, s0.T, p0.T
    prm.F, prm.R, prm.q, prm.kb, prm.cm, prm.tm = p0.F, p0.R, p0.q, p0.kb, p0.cm, p0.tm
    prm.gj_len, prm.gj_surface = cells_list[0].gj_len, p0.gj_surface
    prm.NaK_block, prm.deltaGATP = float(s0.NaKATP_block), p0.deltaGATP
    prm.cATP, prm.cADP, prm.cPi = p0.cATP, p0.cADP, p0.cPi
    prm.KmNK_Na, prm.KmNK_K, prm.KmNK_ATP = p0.KmNK_Na, p0.KmNK_K, p0.KmNK_ATP
    prm.alpha_Ca, prm.KmCa_Ca, prm.KmCa_ATP = p0.alpha_Ca, p0.KmCa_Ca, p0.KmCa_ATP
    prm.cell_height, prm.eo, prm.er = p0.cell_height, p0.eo, p0.er
    prm.cell_radius, prm.true_cell_size = p0.cell_radius, p0.true_cell_size
    for j, w in enumerate(gaussian_taps(1.0)):
        prm.gauss_w[j] = float(w)
    prm.smooth_cells, prm.mu = p0.smooth_cells, p0.mu
    return prm


def make_clusters(states: Sequence, params: Sequence, cells_list: Sequence, soff: np.ndarray):
    """The ``bt_cluster`` table of a batch."""
    cls = (_native.BtCluster * len(states))()
    for i, (st, p, c) in enumerate(zip(states, params, cells_list)):
        cls[i].grid_ny, cls[i].grid_nx = c.X.shape
        cls[i].sq_off = int(soff[i])
        cls[i].delta = c.delta
        cls[i].mean_sa = float(c.memSa_per_envSquare[c.map_mem2ecm].mean())
        cls[i].ko_env = float(st.ko_env)
        cls[i].alpha_NaK = p.alpha_NaK
    return cls


BUFFER_FIELDS: List[str] = [n for n, _ in _native.BtBuffers._fields_[:-2]]


def host_buffers(arrays: Dict[str, np.ndarray], n_tiles: int) -> "_native.BtBuffers":
    """A ``bt_buffers`` whose pointers are the host arrays given (for
    ``bt_validate``); fields without an array stay NULL."""
    buf = _native.BtBuffers()
    for name in BUFFER_FIELDS:
        a = arrays.get(name)
        if a is not None:
            setattr(buf, name, a.ctypes.data)
    buf.n_tiles = int(n_tiles)
    buf.tile_m = _native.BT_TILE_MEMBRANES
    return buf


def validate(params, clusters, arrays: Dict[str, np.ndarray]) -> Tuple[int, str]:
    """``bt_validate`` on host arrays: (return code, message)."""
    lib = _native.load()
    buf = host_buffers(arrays, len(arrays["tile_ptr"]) - 1)
    rc = lib.bt_validate(params, clusters, len(clusters), buf)
    return rc, (lib.bt_last_error(None) or b"").decode()
