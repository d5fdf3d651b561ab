"""Conversion between the reference's objects and this package's.

These are the helpers a maintainer's dispatch branch in the reference's
``Simulator.run_sim_core`` (``betse/science/sim.py:1064-1102``) calls to hand
one phase's loop to the GPU and take the results back:

* ``gpu_cells_from(cells)``    -- the Cells contract of ``betse_amd.cells.Cells``
  from the reference's ``Cells`` (``betse/science/cells.py``; the arrays of
  its ``make_world``, ``cells.py:414-526``), deriving the CSR forms the
  kernels use (``mem_ptr = cumsum(num_mems)``, the square-sorted membrane
  list of ``M_divmap_mem2ecm``);
* ``gpu_params_from(p)``       -- ``betse_amd.parameters.Parameters`` from the
  reference's ``Parameters`` (``betse/science/parameters.py``), which keeps
  the same attribute names;
* ``copy_state(src, dst)``     -- the simulation state in either direction.
  The reference holds ``cc_at_mem`` and ``vm`` per membrane
  (``sim.py:744-745, 2027-2029``); on this path both are constant over a
  cell's membranes (``cc_at_mem = cc_cells[mem_to_cells]`` after
  ``update_intra``; ``vm + Phi_b[map_mem2ecm] = rho_surf[mem_to_cells]``), so
  the GPU state keeps them per cell (``c_lag``, ``vm_cell`` = Vmem plus the
  applied boundary potential ``Phi_b`` at the membrane's square) beside
  ``Phi_b`` per square. Going to the GPU, a per-membrane field that is not
  constant within a cell raises instead of being averaged away.

Every function reads its input by attribute name only (duck typing), so it
works on the reference's live objects and on test stand-ins alike.
"""

from __future__ import annotations

from typing import Optional, Sequence

import numpy as np

from .cells import Cells, _ragged
from .exceptions import BetseSimException
from .parameters import Parameters
from .sim import HostState

# arrays of the reference's Cells that the GPU path consumes (App. A names)
CELL_ARRAYS = ("cell_centres", "mem_mids_flat", "mem_vects_flat", "mem_to_cells", "mem_sa", "num_mems",
               "cell_sa", "R_rads", "mem_vol", "cell_vol", "diviterm", "nn_i", "bflags_mems", "X", "Y", "xypts",
               "map_cell2ecm", "map_mem2ecm", "memSa_per_envSquare", "envInds_inClust", "all_bound_mem_inds",
               "interior_bound_mem_inds", "ecm_inds_bound_cell")
CELL_SCALARS = ("delta", "gj_len", "xmin", "xmax", "ymin", "ymax")

# Parameters attributes copied by name (the reference and this package share them)
PARAM_FIELDS = ("F", "R", "eo", "kb", "q", "mu", "NAv", "er", "deltaGATP", "cATP", "cADP", "cPi", "KmNK_Na",
                "KmNK_K", "KmNK_ATP", "KmCa_Ca", "KmCa_ATP", "alpha_NaK", "alpha_Ca", "scale_cell", "cm", "tm",
                "cell_height", "cell_space", "cell_radius", "true_cell_size", "gj_surface", "T", "is_ecm",
                "cluster_open", "cell_polarizability", "sharpness", "smooth_cells", "fast_update_ecm",
                "v_sensitive_gj", "channel_noise_level", "dynamic_noise", "D_adh", "D_tj", "init_time_step",
                "init_time_total", "init_time_sampling", "sim_time_step", "sim_time_total", "sim_time_sampling",
                "dt", "resample")
PARAM_DICTS = ("Dtj_rel", "ions_dict", "cell_concs", "env_concs", "mem_perms", "free_diff", "ion_charge")


def gpu_cells_from(ref_cells) -> Cells:
    """``Cells`` from the reference's ``Cells`` object (attributes of App. A)."""
    c = Cells()
    for k in CELL_ARRAYS:
        setattr(c, k, np.array(getattr(ref_cells, k)))
    for k in CELL_SCALARS:
        setattr(c, k, float(getattr(ref_cells, k)))
    c.mem_to_cells = c.mem_to_cells.astype(np.int64)
    c.nn_i = c.nn_i.astype(np.int64)
    c.map_mem2ecm = c.map_mem2ecm.astype(np.int64)
    c.map_cell2ecm = c.map_cell2ecm.astype(np.int64)
    c.num_mems = c.num_mems.astype(np.int64)
    n_cells, n_mems = len(c.cell_centres), len(c.mem_to_cells)
    # membranes are flattened cell by cell (math/toolbox.py:21-64, cells.py:1095)
    c.mem_ptr = np.zeros(n_cells + 1, dtype=np.int64)
    c.mem_ptr[1:] = np.cumsum(c.num_mems)
    if c.mem_ptr[-1] != n_mems or not np.array_equal(np.repeat(np.arange(n_cells), c.num_mems), c.mem_to_cells):
        raise BetseSimException("mem_to_cells is not the cell-by-cell flattening of num_mems")
    c.cell_i = np.arange(n_cells)
    c.mem_i = np.arange(n_mems)
    c.cell_to_mems = _ragged([np.arange(c.mem_ptr[q], c.mem_ptr[q + 1]) for q in range(n_cells)])
    c.mem_nn = np.column_stack((c.mem_i, c.nn_i))
    c.bflags_mems = np.asarray(c.bflags_mems, dtype=np.int64)
    c.bflags_cells = c.mem_to_cells[c.bflags_mems]
    c.cell_nn_i = np.column_stack((c.mem_to_cells, c.mem_to_cells[c.nn_i]))
    c.gj_default_weights = np.array(getattr(ref_cells, "gj_default_weights", np.ones(n_mems)), dtype=np.float64)
    c._set_cell_nn()
    # the grid side of M_divmap_mem2ecm: membranes sorted by square, then by index
    G = c.X.size
    c.grid_n = c.X.shape[0]
    counts = np.bincount(c.map_mem2ecm, minlength=G)
    c.sq_ptr = np.zeros(G + 1, dtype=np.int64)
    c.sq_ptr[1:] = np.cumsum(counts)
    c.sq_mems = np.argsort(c.map_mem2ecm, kind="stable").astype(np.int64)
    c.mems_per_envSquare = counts.astype(np.float64)
    return c


def gpu_params_from(ref_p, *, channels: Sequence = (), network=None, **overrides) -> Parameters:
    """``Parameters`` from the reference's ``Parameters``: every constant, the
    ion profile and the time settings by name. The voltage-gated channels
    (``ChannelSpec`` per channel, from the molecules network) and the
    intracellular reaction network are given explicitly."""
    p = Parameters()
    for k in PARAM_FIELDS:
        if hasattr(ref_p, k):
            setattr(p, k, getattr(ref_p, k))
    for k in PARAM_DICTS:
        if hasattr(ref_p, k):
            setattr(p, k, dict(getattr(ref_p, k)))
    p.cbnd = dict(ref_p.cbnd) if getattr(ref_p, "cbnd", None) else None
    p.t_resample = p.resample / p.dt
    p.channels = tuple(channels)
    p.network = network
    for k, v in overrides.items():
        setattr(p, k, v)
    return p


def _per_cell(x: np.ndarray, mem_ptr: np.ndarray, name: str, atol: float = 0.0) -> np.ndarray:
    """Per-cell values (each cell's first membrane's) of a per-membrane field
    ([..., M]) that is constant over each cell's membranes, exactly or, where
    it was rebuilt from a difference, within ``atol``."""
    first = x[..., mem_ptr[:-1]]
    counts = np.diff(mem_ptr)
    dev = np.abs(np.repeat(first, counts, axis=-1) - x)
    if not (np.array_equal(np.repeat(first, counts, axis=-1), x) if atol == 0.0 else bool(np.all(dev <= atol))):
        raise BetseSimException(f"{name} differs between the membranes of a cell: the GPU state keeps it per cell")
    return first.copy()


STATE_OPTIONAL = ("Dm_cells", "D_env")


def copy_state(src, dst, cells: Optional[object] = None) -> None:
    """Copy the simulation state from ``src`` to ``dst``, where one of them is a
    ``HostState`` (``Simulator``/``BatchSimulator`` state) of this package and
    the other a reference ``Simulator`` (or anything with its attribute names).
    ``cells`` is the reference's Cells (default: ``dst.cells``/``src.cells``)."""
    if isinstance(dst, HostState) and not isinstance(src, HostState):
        _to_gpu(src, dst, cells if cells is not None else dst.cells)
    elif isinstance(src, HostState) and not isinstance(dst, HostState):
        _from_gpu(src, dst, cells if cells is not None else src.cells)
    else:
        raise TypeError("copy_state converts between a reference Simulator and a betse_amd HostState")


def _to_gpu(ref, st: HostState, cells) -> None:
    mem_ptr = np.zeros(len(cells.num_mems) + 1, dtype=np.int64)
    mem_ptr[1:] = np.cumsum(cells.num_mems)
    st.cc_cells = np.array(ref.cc_cells, dtype=np.float64)
    st.cc_env = np.array(ref.cc_env, dtype=np.float64)
    st.c_lag = _per_cell(np.asarray(ref.cc_at_mem, dtype=np.float64), mem_ptr, "cc_at_mem")
    shp = np.asarray(cells.X).shape
    Phi_b = np.array(getattr(ref, "Phi_b", np.zeros(int(np.prod(shp)))), dtype=np.float64).ravel()
    vm = np.asarray(ref.vm, dtype=np.float64)
    if np.any(Phi_b != 0):
        # rho_surf/cm = vm + Phi_b per membrane, up to the rounding of the
        # reference's subtraction (a few ulps of |vm| + |Phi_b|)
        pb = Phi_b[np.asarray(cells.map_mem2ecm)]
        tol = 8 * np.finfo(np.float64).eps * float(np.abs(vm).max() + np.abs(pb).max())
        st.vm_cell = _per_cell(vm + pb, mem_ptr, "vm + Phi_b[map_mem2ecm]", atol=tol)
    else:
        st.vm_cell = _per_cell(vm, mem_ptr, "vm")
    st.Phi_b = Phi_b
    bound_V = getattr(ref, "bound_V", None)
    if bound_V is not None:
        st.bound_V = {k: float(bound_V[k]) for k in ("T", "B", "L", "R")}
    st.E_env_x = np.array(ref.E_env_x, dtype=np.float64).reshape(shp)
    st.E_env_y = np.array(ref.E_env_y, dtype=np.float64).reshape(shp)
    st.v_env = np.array(ref.v_env, dtype=np.float64).ravel()
    for k in STATE_OPTIONAL:
        if hasattr(ref, k):
            setattr(st, k, np.array(getattr(ref, k), dtype=np.float64))
    gates = getattr(ref, "chan_gates", None)
    if gates is not None:
        st.chan_gates = np.array(gates, dtype=np.float64)


def _from_gpu(st: HostState, ref, cells) -> None:
    m2c = np.asarray(cells.mem_to_cells)
    ref.cc_cells = st.cc_cells.copy()
    ref.cc_env = st.cc_env.copy()
    ref.cc_at_mem = st.c_lag[:, m2c].copy()
    Phi_b = np.asarray(st.Phi_b, dtype=np.float64)
    ref.vm = st.vm_cell[m2c] - Phi_b[np.asarray(cells.map_mem2ecm)] if np.any(Phi_b != 0) else st.vm_cell[m2c].copy()
    ref.vm_ave = np.add.reduceat(ref.vm, np.concatenate([[0], np.cumsum(cells.num_mems)[:-1]])) / cells.num_mems
    ref.rho_cells = np.dot(st.zs * st.p.F, st.cc_cells)
    ref.E_env_x = st.E_env_x.copy()
    ref.E_env_y = st.E_env_y.copy()
    ref.v_env = st.v_env.copy()
    ref.Phi_b = Phi_b.copy()
    ref.bound_V = dict(st.bound_V)
    for k in STATE_OPTIONAL:
        setattr(ref, k, getattr(st, k).copy())
    if len(getattr(st, "channel_specs", ())):
        ref.chan_gates = st.chan_gates.copy()
    # the sampled time series of write2storage (sim.py:1789-1884), for the exporters
    for k, v in vars(st).items():
        if k.endswith("_time") and isinstance(v, list):
            setattr(ref, k, list(v))
    if hasattr(st, "time"):
        ref.time = list(st.time)
