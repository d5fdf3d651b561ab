"""The applied boundary voltage on the GPU: the reference's voltage event
(``TissueEventVoltage.fire``, ``tisevevolt.py:75-89``) sets ``bound_V`` at
the start of a step, ``get_current`` (``ion_current.py:84-107``) turns it into
``Phi_b`` at that step's end, and from then on every membrane's Vmem is its
cell's less ``Phi_b`` at its square and ``v_env`` adds ``Phi_b``. The GPU run
(``set_boundary_voltage`` -> ``bt_stage_boundary``) is checked against the
oracle stepped through the same changes, at the FP64 bars: a ramp that
changes every step, a plateau, the voltage taken away again; the sampled
observables of the step a change fires in (E_gj from the Vmem before it, the
rest after it); one cluster of a batch, under every launch variant."""

import numpy as np
import pytest

from betse_amd.batch import batched_worlds
from test_gpu_parity import check, event_params, gpu_run, jenv_bound, jenv_rel, oracle_stepper, say  # noqa: F401

pytestmark = pytest.mark.gpu

RAMP = [dict(T=0.01, B=-0.01, L=0.0, R=0.0), dict(T=0.025, B=-0.025, L=0.0, R=0.0),
        dict(T=0.04, B=-0.04, L=0.005, R=-0.005)]
ZERO = dict(T=0.0, B=0.0, L=0.0, R=0.0)


def check_phib(st, o, tag):
    dp = float(np.abs(st.Phi_b - o.Phi_b).max())
    dve = float(np.abs(st.v_env - o.v_env).max())
    say(f"{tag}: |dPhi_b|={dp:.3e} |dv_env|={dve:.3e}")
    assert dp <= 1e-13 * max(np.abs(o.Phi_b).max(), 1e-3)
    assert dve <= 1e-9


@pytest.fixture(scope="module")
def tiny():
    from betse_amd.cells import Cells
    from betse_amd.parameters import Parameters
    p = Parameters.preset("tiny")
    return p, Cells.make_world(p)


def test_boundary_voltage_ramp_plateau_release(tiny):
    p, cells = tiny
    from betse_amd.sim import SimPhase, SimPhaseKind, Simulator
    say("case: bound_V ramp (3 steps), plateau (12), release (10), tiny fp64")
    sim = Simulator(p.copy(), device="cuda:0")
    ph = SimPhase(SimPhaseKind.INIT, sim.p, cells, sim)
    sim.init_core(ph)
    sim.prepare(ph)
    o = oracle_stepper(p, cells)
    sim.step_n(5)
    for _ in range(5):
        o.step("init", observables=False)
    for b in RAMP:                       # a change every step: each staged before its step
        sim.set_boundary_voltage(b)
        o.bound_V = dict(b)
        sim.step_n(1)
        o.step("init", observables=False)
    sim.step_n(12)
    for _ in range(12):
        o.step("init", observables=False)
    sim.pull()
    assert np.abs(o.Phi_b).max() > 0.01
    check(sim, o, "plateau")
    check_phib(sim, o, "plateau")
    # the voltage moved Vmem: without it the run differs
    ref = gpu_run(p, cells, 20)
    assert np.abs(ref.vm - sim.vm).max() > 1e-4
    sim.set_boundary_voltage(ZERO)
    o.bound_V = dict(ZERO)
    sim.step_n(10)
    for _ in range(10):
        o.step("init", observables=False)
    sim.pull()
    assert not o.Phi_b.any() and not sim.Phi_b.any()
    check(sim, o, "released")
    check_phib(sim, o, "released")


def observables_check(sim, o, cells, p, tag):
    rec = {"J_env_x": o.J_env_x, "J_env_y": o.J_env_y, "fluxes_env_x": o.fluxes_env_x,
           "fluxes_env_y": o.fluxes_env_y}
    egj = max(np.abs(sim.E_gj_x - o.E_gj_x).max(), np.abs(sim.E_gj_y - o.E_gj_y).max())
    assert egj <= max(1e-8 * np.abs(o.E_gj_x).max(), 2e-9 / cells.gj_len), (tag, "E_gj", egj)
    dva = np.abs(sim.vm_ave_obs - o.vm_ave).max()
    assert dva <= 1e-9, (tag, "vm_ave", dva)
    scale = max(np.abs(o.J_env_x).max(), np.abs(o.J_env_y).max())
    jerr = max(np.abs(sim.J_env_x.ravel() - o.J_env_x.ravel()).max(),
               np.abs(sim.J_env_y.ravel() - o.J_env_y.ravel()).max())
    assert jerr <= jenv_bound(rec, p, scale), (tag, "J_env", jerr)
    rel = jenv_rel(rec, p)
    for k in ("B_field", "Jtx", "Jty"):
        a, b = getattr(sim, k).ravel(), getattr(o, k).ravel()
        assert np.abs(a - b).max() <= rel * np.abs(b).max() + 1e-300, (tag, k)
    say(f"{tag}: |dE_gj|={egj:.3e} |dvm_ave|={dva:.3e} |dJ_env|={jerr:.3e}")


def test_observables_of_the_step_a_change_fires_in(tiny):
    """Sampled observables of the step the event fires in: its E_gj comes from
    the Vmem (and Phi_b) before the step, vm_ave and the env currents' boundary
    term from after it; then a change staged three steps before a sample."""
    p, cells = tiny
    from betse_amd.sim import SimPhase, SimPhaseKind, Simulator
    say("case: observables at a bound_V change, tiny fp64")
    sim = Simulator(p.copy(), device="cuda:0")
    ph = SimPhase(SimPhaseKind.INIT, sim.p, cells, sim)
    sim.init_core(ph)
    sim.prepare(ph)
    o = oracle_stepper(p, cells)
    sim.step_n(6)
    for _ in range(6):
        o.step("init", observables=False)
    b1, b2 = RAMP[2], dict(T=-0.03, B=0.02, L=0.01, R=0.0)
    sim.set_boundary_voltage(b1)
    o.bound_V = dict(b1)
    sim.step_observed(1)
    o.step("init", observables=True)
    assert np.abs(o.Phi_b).max() > 0.01
    observables_check(sim, o, cells, p, "change step")
    check(sim, o, "change step")
    sim.set_boundary_voltage(b2)
    o.bound_V = dict(b2)
    sim.step_observed(3)
    for s in range(3):
        o.step("init", observables=(s == 2))
    observables_check(sim, o, cells, p, "two steps after a change")
    check(sim, o, "two steps after a change")
    check_phib(sim, o, "two steps after a change")


VARIANTS = [dict(), dict(persistent=1), dict(streams=2), dict(mem_waves=2, exchange_form=1, env_ions_per_block=1),
            dict(exchange_form=2, env_ions_per_block=5)]


def test_batch_voltage_on_one_cluster_every_variant():
    """A voltage on cluster 1 of a batch of 3: that cluster follows the oracle
    stepped through it, the others are untouched, and every launch variant
    (persistent kernel, stream groups, wavefront splits, the exchange in the
    env kernel) gives the same bits."""
    from betse_amd.batch import BatchSimulator
    worlds = batched_worlds(n_clusters=3, cells_range=(30, 300), grid_range=(16, 32), seed=11)
    say(f"case: bound_V on cluster 1 of a batch of 3: cells {[c.n_cells for _, c in worlds]}")
    runs = []
    for v in VARIANTS:
        b = BatchSimulator([(p.copy(), c) for p, c in worlds], device="cuda:0", flux_mode="fp64")
        b.init("init")
        if v:
            b.set_tuning(v)
        b.step_n(8)
        b.set_boundary_voltage(1, RAMP[1])
        b.step_n(17)
        b.pull()
        runs.append([(st.cc_cells.copy(), st.vm.copy(), st.cc_env.copy()) for st in b.states])
        say(f"variant {v or 'default'}: tuning {b.get_tuning()}")
    for r in runs[1:]:
        for (c0, v0, e0), (c1, v1, e1) in zip(runs[0], r):
            np.testing.assert_array_equal(c0, c1)
            np.testing.assert_array_equal(v0, v1)
            np.testing.assert_array_equal(e0, e1)
    orcs = [oracle_stepper(p, c) for p, c in worlds]
    for o in orcs:
        for _ in range(8):
            o.step("init", observables=False)
    orcs[1].bound_V = dict(RAMP[1])
    for o in orcs:
        for _ in range(17):
            o.step("init", observables=False)
    for k, (o, (cc, vm, ce)) in enumerate(zip(orcs, runs[0])):
        dv = np.abs(vm - o.vm).max()
        dc = (np.abs(cc - o.cc_cells) / np.abs(o.cc_cells)).max()
        de = (np.abs(ce - o.cc_env) / np.abs(o.cc_env)).max()
        say(f"cluster {k}: |dVm|={dv:.3e} rel dc={dc:.3e} rel denv={de:.3e}")
        assert dv <= 1e-9 and dc <= 1e-10 and de <= 1e-10
    assert np.abs(orcs[1].Phi_b).max() > 0.01 and not orcs[0].Phi_b.any()
