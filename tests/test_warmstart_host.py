"""Host side of the learned warm start: features, graph batching and the
charge-consistent injection (no GPU)."""

import numpy as np
import pytest

from betse_amd.batch import batched_worlds
from betse_amd.parameters import Parameters
from betse_amd.sim import HostState, SimPhase, SimPhaseKind
from betse_amd import warmstart as ws


@pytest.fixture(scope="module")
def worlds():
    return batched_worlds(3, cells_range=(60, 150), grid_range=(16, 24), seed=5, decade=2.0)


def _state(p, cells):
    st = HostState()
    ph = SimPhase(SimPhaseKind.INIT, p, cells)
    st.init_core(ph)
    st.init_dynamics(ph)
    return st


def test_features_and_graph_batch(worlds):
    base = Parameters.preset("small")
    X, ptr, idx, coff = ws.graph_batch(worlds, base)
    assert X.shape == (coff[-1], ws.N_FEATURES)
    assert np.all(np.isfinite(X))
    assert ptr[0] == 0 and ptr[-1] == len(idx) and np.all(np.diff(ptr) >= 1)
    for i, (p, c) in enumerate(worlds):
        a, b = coff[i], coff[i + 1]
        # neighbours of a cluster stay inside it and the graph is symmetric
        seg = idx[ptr[a]:ptr[b]]
        assert seg.min() >= a and seg.max() < b
        src = np.repeat(np.arange(a, b), np.diff(ptr[a:b + 1]))
        pairs = set(zip(src.tolist(), seg.tolist()))
        assert all((j, i2) in pairs for i2, j in pairs)
        # parameter features are the log10 factors the draw applied
        assert np.allclose(X[a:b, 4], np.log10(p.mem_perms["Na"] / base.mem_perms["Na"]))
        assert np.allclose(X[a:b, 7], np.log10(p.alpha_NaK / base.alpha_NaK))


def test_targets_of_default_state_are_zero_conc_and_zero_vm(worlds):
    p, c = worlds[0]
    st = _state(p, c)
    Y = ws.state_targets(st, st.names)
    assert Y.shape == (c.n_cells, 4)
    assert np.allclose(Y, 0.0)


def test_inject_reproduces_predicted_state_and_vm(worlds):
    p, c = worlds[1]
    st = _state(p, c)
    rng = np.random.RandomState(0)
    Y = np.zeros((c.n_cells, 4))
    Y[:, 0] = rng.uniform(-0.5, 0.5, c.n_cells)      # Na
    Y[:, 1] = rng.uniform(-0.1, 0.1, c.n_cells)      # K (M stays positive)
    Y[:, 2] = rng.uniform(-0.3, 0.3, c.n_cells)      # Cl
    Y[:, 3] = rng.uniform(-0.08, -0.02, c.n_cells) / ws.V_SCALE
    assert ws.inject(st, Y)
    # the charge of the injected state gives the predicted Vmem (sim.py:2007-2083 algebra)
    rho = np.dot(st.zs * p.F, st.cc_cells)
    vm = rho * c.diviterm / p.cm
    assert np.allclose(vm, Y[:, 3] * ws.V_SCALE, rtol=1e-9, atol=1e-12)
    back = ws.state_targets(st, st.names)
    assert np.allclose(back[:, :3], Y[:, :3], atol=1e-12)
    assert np.array_equal(st.c_lag, st.cc_cells)


def test_inject_refuses_unusable_predictions(worlds):
    p, c = worlds[2]
    st = _state(p, c)
    cc0 = st.cc_cells.copy()
    Y = np.zeros((c.n_cells, 4))
    Y[0, 0] = np.nan
    assert not ws.inject(st, Y)
    Y = np.zeros((c.n_cells, 4))
    Y[:, 3] = 1.0 / ws.V_SCALE          # 1 V
    assert not ws.inject(st, Y)
    Y = np.zeros((c.n_cells, 4))
    Y[:, 2] = 5.0                       # Cl x e^5: M would go negative
    assert not ws.inject(st, Y)
    assert np.array_equal(st.cc_cells, cc0)


def test_inject_sets_channel_gates_to_their_steady_state():
    from betse_amd.channels import draw_channel_state, init_gates
    p = Parameters.preset("tiny_channels")
    cells = __import__("betse_amd.cells", fromlist=["Cells"]).Cells.make_world(p)
    draw_channel_state(p, cells.n_cells, cells.n_mems)
    st = _state(p, cells)
    Y = np.zeros((cells.n_cells, 5))                  # Na, K, Cl, Ca, Vm
    Y[:, 4] = -0.065 / ws.V_SCALE
    assert ws.inject(st, Y)
    vm_mem = np.full(cells.n_mems, -0.065)
    for k, spec in enumerate(st.channel_specs):
        m, h = init_gates(spec.kind, vm_mem)
        assert np.allclose(st.chan_gates[k, 0], m) and np.allclose(st.chan_gates[k, 1], h)
    # features carry the per-cell channel multipliers
    X = ws.cell_features(p, cells, Parameters.preset("tiny_channels"))
    assert np.allclose(X[:, 10], np.log(p.channel_cell_mult[0]))
    assert np.allclose(X[:, 11], np.log(p.channel_cell_mult[1]))


def test_extrapolation_recovers_the_limit_of_a_geometric_relaxation():
    """x_n = x* + A q^n per cluster (one ratio per cluster, any amplitudes):
    three snapshots give x* exactly; a non-contracting cluster keeps x3."""
    rng = np.random.RandomState(1)
    coff = np.array([0, 5, 12, 20])
    xs = rng.uniform(1.0, 10.0, (3, 20))
    A = rng.uniform(-0.5, 0.5, (3, 20))
    vs, Av = rng.uniform(-0.07, -0.05, 20), rng.uniform(-0.01, 0.01, 20)
    q = np.array([0.9, 0.5, 1.2])                        # last cluster grows: no extrapolation
    qc = np.repeat(q, np.diff(coff))
    snaps = [(xs + A * qc ** n, vs + Av * qc ** n) for n in (1, 2, 3)]
    cx, vx, qs = ws._extrapolate(snaps, coff)
    assert np.allclose(qs, q)
    assert np.allclose(cx[:, :12], xs[:, :12], rtol=1e-12)
    assert np.allclose(vx[:12], vs[:12], rtol=1e-12)
    assert np.array_equal(cx[:, 12:], snaps[2][0][:, 12:])
