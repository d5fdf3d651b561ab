"""The reference-side adapter (gpu_cells_from, gpu_params_from, copy_state) on
stand-ins that carry only the reference's attribute names (CPU only)."""

from types import SimpleNamespace

import numpy as np
import pytest

from betse_amd.adapter import CELL_ARRAYS, CELL_SCALARS, PARAM_DICTS, PARAM_FIELDS, copy_state, gpu_cells_from, \
    gpu_params_from
from betse_amd.cells import Cells
from betse_amd.exceptions import BetseSimException
from betse_amd.parameters import Parameters
from betse_amd.sim import HostState, SimPhase, SimPhaseKind


@pytest.fixture(scope="module")
def world():
    p = Parameters.preset("tiny")
    return p, Cells.make_world(p)


def reference_cells(cells):
    """A stand-in with the reference Cells' names only (no CSR arrays)."""
    ns = SimpleNamespace(**{k: np.array(getattr(cells, k)) for k in CELL_ARRAYS})
    for k in CELL_SCALARS:
        setattr(ns, k, getattr(cells, k))
    ns.gj_default_weights = np.ones(cells.n_mems)
    ns.cell_nn = cells.cell_nn
    return ns


def test_cells_round_trip(world):
    p, cells = world
    c = gpu_cells_from(reference_cells(cells))
    for k in ("mem_ptr", "sq_ptr", "sq_mems", "cell_nn_ptr", "cell_nn_idx", "bflags_cells", "mems_per_envSquare"):
        np.testing.assert_array_equal(getattr(c, k), getattr(cells, k), err_msg=k)
    assert c.n_cells == cells.n_cells and c.n_mems == cells.n_mems and c.X.shape == cells.X.shape


def test_cells_flattening_is_checked(world):
    p, cells = world
    ref = reference_cells(cells)
    ref.mem_to_cells = ref.mem_to_cells[::-1].copy()
    with pytest.raises(BetseSimException):
        gpu_cells_from(ref)


def test_params_by_name(world):
    p, _ = world
    ref = SimpleNamespace(**{k: getattr(p, k) for k in PARAM_FIELDS + PARAM_DICTS}, cbnd=None)
    q = gpu_params_from(ref)
    for k in PARAM_FIELDS + PARAM_DICTS:
        assert getattr(q, k) == getattr(p, k), k
    assert q.ion_names() == p.ion_names()
    assert q.t_resample == p.resample / p.dt


def host_state(p, cells):
    st = HostState()
    ph = SimPhase(SimPhaseKind.INIT, p, cells)
    st.init_core(ph)
    st.init_dynamics(ph)
    return st


def test_state_round_trip_per_membrane_to_per_cell(world):
    p, cells = world
    st = host_state(p.copy(), cells)
    rng = np.random.RandomState(0)
    ni, C, M, G = st.cc_cells.shape[0], cells.n_cells, cells.n_mems, cells.X.size
    cc = st.cc_cells * rng.uniform(0.9, 1.1, (ni, C))
    ref = SimpleNamespace(cc_cells=cc, cc_env=st.cc_env * rng.uniform(0.9, 1.1, (ni, G)),
                          cc_at_mem=(cc * 1.01)[:, cells.mem_to_cells], vm=rng.randn(C)[cells.mem_to_cells] * 0.01,
                          E_env_x=rng.randn(G), E_env_y=rng.randn(G), v_env=rng.randn(G), Phi_b=np.zeros(G))
    copy_state(ref, st, cells)
    np.testing.assert_array_equal(st.c_lag, cc * 1.01)
    np.testing.assert_array_equal(st.vm, ref.vm)           # per membrane again through mem_to_cells
    back = SimpleNamespace()
    copy_state(st, back, cells)
    for k in ("cc_cells", "cc_env", "cc_at_mem", "vm"):
        np.testing.assert_array_equal(getattr(back, k), getattr(ref, k), err_msg=k)
    np.testing.assert_array_equal(back.E_env_x.ravel(), ref.E_env_x)
    np.testing.assert_allclose(back.vm_ave, np.add.reduceat(ref.vm, cells.mem_ptr[:-1]) / cells.num_mems)


def test_state_not_constant_over_a_cell_raises(world):
    p, cells = world
    st = host_state(p.copy(), cells)
    G = cells.X.size
    vm = np.zeros(cells.n_mems)
    vm[1] = 1e-3                                        # two membranes of cell 0 disagree
    ref = SimpleNamespace(cc_cells=st.cc_cells, cc_env=st.cc_env, cc_at_mem=st.cc_cells[:, cells.mem_to_cells],
                          vm=vm, E_env_x=np.zeros(G), E_env_y=np.zeros(G), v_env=np.zeros(G), Phi_b=np.zeros(G))
    with pytest.raises(BetseSimException, match="vm differs"):
        copy_state(ref, st, cells)
    # with an applied boundary potential Vmem differs within a cell by Phi_b at
    # each membrane's square, and that is what the per-cell state gives back
    rng = np.random.RandomState(3)
    ref.Phi_b = rng.uniform(-1e-3, 1e-3, G)
    ref.bound_V = {"T": 1e-3, "B": -1e-3, "L": 0.0, "R": 0.0}
    rho = rng.randn(cells.n_cells) * 0.01
    ref.vm = rho[cells.mem_to_cells] - ref.Phi_b[cells.map_mem2ecm]
    copy_state(ref, st, cells)
    np.testing.assert_allclose(st.vm_cell, rho, rtol=0, atol=1e-17)
    np.testing.assert_array_equal(st.Phi_b, ref.Phi_b)
    back = SimpleNamespace()
    copy_state(st, back, cells)
    np.testing.assert_allclose(back.vm, ref.vm, rtol=0, atol=1e-17)
    np.testing.assert_array_equal(back.Phi_b, ref.Phi_b)
    assert back.bound_V == ref.bound_V
    np.testing.assert_allclose(st.vm, ref.vm, rtol=0, atol=1e-17)
    ref.vm[0] += 1e-3                                   # not rho - Phi_b: no per-cell value
    with pytest.raises(BetseSimException, match="differs"):
        copy_state(ref, st, cells)
