"""A batch of independent clusters stepped together, each against its own
oracle run (per-cluster grids, spacing, diffusivities, pump rate)."""

import sys

import numpy as np
import pytest

from betse_amd.batch import batched_worlds
from betse_amd.channels import draw_channel_state
from oracle.betse_oracle import OracleSim
from test_gpu_parity import jenv_bound, jenv_rel

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def run_batch(worlds, n, flux_mode):
    from betse_amd.batch import BatchSimulator
    b = BatchSimulator([(p.copy(), c) for p, c in worlds], device="cuda:0", flux_mode=flux_mode)
    b.init("init")
    b.step_n(n)
    b.pull()
    return b


@pytest.mark.parametrize("flux_mode,vm_tol,rel_tol", [("fp64", 1e-9, 1e-10), ("fp32", 1e-6, 1e-6)])
def test_batch_matches_per_cluster_oracle(flux_mode, vm_tol, rel_tol):
    worlds = batched_worlds(n_clusters=5, cells_range=(30, 400), grid_range=(16, 40), seed=3)
    n = 60
    say(f"case: batch of 5 clusters, {flux_mode}, {n} steps: cells "
        f"{[c.n_cells for _, c in worlds]} grids {[c.X.shape[0] for _, c in worlds]}")
    b = run_batch(worlds, n, flux_mode)
    for k, ((p, cells), st) in enumerate(zip(worlds, b.states)):
        o = OracleSim(cells, p.copy(), mode="sparse")
        o.init_core()
        o.run_sim_core("init", n_steps=n)
        dv = np.abs(st.vm - o.vm).max()
        dc = (np.abs(st.cc_cells - o.cc_cells) / np.abs(o.cc_cells)).max()
        de = (np.abs(st.cc_env - o.cc_env) / np.abs(o.cc_env)).max()
        say(f"cluster {k}: |dVm|={dv:.3e} rel dc={dc:.3e} rel denv={de:.3e}")
        assert dv <= vm_tol and dc <= rel_tol and de <= rel_tol


def test_batch_with_channels():
    worlds = batched_worlds(n_clusters=3, cells_range=(30, 200), grid_range=(16, 24), seed=5,
                            base="tiny_channels")
    for p, c in worlds:
        draw_channel_state(p, c.n_cells, c.n_mems)
    say("case: batch of 3 clusters with Nav/Kv, fp64, 40 steps")
    b = run_batch(worlds, 40, "fp64")
    for k, ((p, cells), st) in enumerate(zip(worlds, b.states)):
        o = OracleSim(cells, p.copy(), mode="sparse")
        o.init_core()
        o.run_sim_core("init", n_steps=40)
        dv = np.abs(st.vm - o.vm).max()
        say(f"cluster {k}: |dVm|={dv:.3e}")
        assert dv <= 1e-9
        np.testing.assert_allclose(st.cc_cells, o.cc_cells, rtol=1e-10)
        np.testing.assert_allclose(st.chan_gates[0, 0], o.channels[0].m, rtol=1e-10, atol=1e-14)


def test_batch_with_full_physics():
    """Nav/Kv/Cav and the reaction network in a batch of clusters of
    different sizes, each against its own oracle run."""
    worlds = batched_worlds(n_clusters=3, cells_range=(30, 200), grid_range=(16, 24), seed=9, base="tiny_full",
                            decade=0.0)
    say("case: batch of 3 clusters with Nav/Kv/Cav and the network, fp64, 60 steps")
    b = run_batch(worlds, 60, "fp64")
    for k, ((p, cells), st) in enumerate(zip(worlds, b.states)):
        o = OracleSim(cells, p.copy(), mode="sparse")
        o.init_core()
        o.run_sim_core("init", n_steps=60)
        dv = np.abs(st.vm - o.vm).max()
        dn = (np.abs(st.c_net - o.c_net) / np.maximum(np.abs(o.c_net), 1e-30)).max()
        say(f"cluster {k}: |dVm|={dv:.3e} rel dc_net={dn:.3e}")
        assert dv <= 1e-9
        np.testing.assert_allclose(st.cc_cells, o.cc_cells, rtol=1e-10)
        for ch in range(len(o.channels)):
            np.testing.assert_allclose(st.chan_gates[ch, 0], o.channels[ch].m, rtol=1e-10, atol=1e-14)
        assert dn < 1e-10


def test_batch_sampled_observables_match_oracle():
    """The batch's loop with sampled steps: get_current's observables for every
    cluster (each Helmholtz-Hodge decomposition on its own grid, the
    cell/membrane ones over the whole batch) and each cluster's
    write2storage, against per-cluster oracle runs with the same sampling."""
    from betse_amd.batch import BatchSimulator
    worlds = batched_worlds(n_clusters=3, cells_range=(30, 200), grid_range=(16, 30), seed=5,
                            base="tiny_channels")
    n = 40
    for p, c in worlds:
        draw_channel_state(p, c.n_cells, c.n_mems)
        p.init_time_sampling = 10 * p.init_time_step
        p.set_phase_time("init")
    say(f"case: batch sampled observables, 3 clusters with Nav/Kv, grids {[c.X.shape[0] for _, c in worlds]}")
    b = BatchSimulator([(p.copy(), c) for p, c in worlds], device="cuda:0", flux_mode="fp64")
    b.init("init")
    n_samp = b.run_loop(n)
    assert n_samp == 3
    pairs = {"I_mem_time": "I_mem", "I_cell_x_time": "J_cell_x", "I_cell_y_time": "J_cell_y",
             "efield_gj_x_time": "E_gj_x", "rate_NaKATP_time": "rate_NaKATP", "I_tot_x_time": "J_env_x",
             "I_tot_y_time": "J_env_y", "vm_time": "vm", "cc_time": "cc_cells"}
    for k, ((p, cells), st) in enumerate(zip(worlds, b.states)):
        o = OracleSim(cells, p.copy(), mode="sparse")
        o.init_core()
        o.run_sim_core("init", n_steps=n)
        assert len(o.storage) == len(st.I_mem_time) == n_samp
        for rec_i, rec in enumerate(o.storage):
            jn_scale = np.abs(np.asarray(rec["I_mem"]) / cells.mem_sa).max()
            for mine, theirs in pairs.items():
                a = np.asarray(getattr(st, mine)[rec_i]).ravel()
                r = np.asarray(rec[theirs]).ravel()
                scale = max(np.abs(r).max(), 1e-300)
                if mine.startswith("I_cell"):
                    scale = max(scale, jn_scale)
                err = np.abs(a - r).max()
                if mine.startswith("efield_gj"):
                    assert err <= max(1e-8 * scale, 2e-9 / cells.gj_len), (k, mine, rec_i, err, scale)
                elif mine.startswith("I_tot"):
                    bound = jenv_bound(rec, p, scale)
                    if rec_i == 0:
                        say(f"  cluster {k} {mine}: err {err:.3e}, |J| {scale:.3e}, bar {bound:.3e}")
                    assert err <= bound, (k, mine, rec_i, err, scale, bound)
                elif mine == "vm_time":
                    assert err <= 1e-9, (k, mine, rec_i, err)
                else:
                    assert err / scale < 1e-8, (k, mine, rec_i, err / scale)
        rel = jenv_rel(o.storage[-1], p)
        np.testing.assert_allclose(st.B_field.ravel(), o.B_field.ravel(), rtol=0, atol=rel * np.abs(o.B_field).max())
        say(f"cluster {k}: {n_samp} sampled steps match")


def test_batch_event_on_one_cluster():
    """An event on one cluster of a batch (set_membrane_params): that cluster
    follows the oracle stepped through the change, the others are untouched."""
    from test_gpu_parity import event_params, oracle_stepper
    worlds = batched_worlds(n_clusters=3, cells_range=(30, 300), grid_range=(16, 32), seed=11)
    say(f"case: event on cluster 1 of a batch of 3: cells {[c.n_cells for _, c in worlds]}")
    from betse_amd.batch import BatchSimulator
    b = BatchSimulator([(p.copy(), c) for p, c in worlds], device="cuda:0", flux_mode="fp64")
    b.init("init")
    b.step_n(10)
    orcs = [oracle_stepper(p, c) for p, c in worlds]
    for o in orcs:
        for _ in range(10):
            o.step("init", observables=False)
    Dm, gjopen = event_params(orcs[1], worlds[1][1])
    orcs[1].Dm_cells = Dm.copy()
    orcs[1].gj_block = 0.5
    b.set_membrane_params(1, Dm=Dm, gjopen=gjopen)
    for o in orcs:
        for _ in range(15):
            o.step("init", observables=False)
    b.step_n(15)
    b.pull()
    for k, (o, st) in enumerate(zip(orcs, b.states)):
        dv = np.abs(st.vm - o.vm).max()
        dc = (np.abs(st.cc_cells - o.cc_cells) / np.abs(o.cc_cells)).max()
        de = (np.abs(st.cc_env - o.cc_env) / np.abs(o.cc_env)).max()
        say(f"cluster {k}: |dVm|={dv:.3e} rel dc={dc:.3e} rel denv={de:.3e}")
        assert dv <= 1e-9 and dc <= 1e-10 and de <= 1e-10


def test_sampled_steps_with_the_fused_exchange_are_bit_identical():
    """The loop with sampled steps (observed steps, observables, write2storage)
    with the exchange in the env kernel against the exchange kernel: every
    stored series of every cluster bit for bit (odd and even step counts
    between samples, the observed step's env fluxes taken from the env
    kernel that also does the exchange)."""
    from betse_amd.batch import BatchSimulator
    worlds = batched_worlds(n_clusters=4, cells_range=(30, 300), grid_range=(16, 32), seed=9)
    n = 47
    for p, c in worlds:
        p.init_time_sampling = 9 * p.init_time_step
        p.set_phase_time("init")
    runs = []
    for tuning in ({"exchange_form": 1}, {"exchange_form": 2, "env_ions_per_block": 5}):
        say(f"case: sampled loop, 4 clusters, tuning {tuning}")
        b = BatchSimulator([(p.copy(), c) for p, c in worlds], device="cuda:0", flux_mode="fp64", tuning=tuning)
        b.init("init")
        assert b.get_tuning()["exchange_form"] == tuning["exchange_form"]
        n_samp = b.run_loop(n)
        assert n_samp >= 5
        runs.append(b.states)
        b.close()          # the stored series outlive the context; a second close does nothing
        b.close()
        assert "_worker" not in vars(b) and "_dev" not in vars(b)
    keys = [k for k in vars(runs[0][0]) if k.endswith("_time") and isinstance(getattr(runs[0][0], k), list)]
    assert {"vm_time", "cc_env_time", "I_tot_x_time", "venv_time"} <= set(keys)
    for k_cl, (sa, sb) in enumerate(zip(*runs)):
        for key in keys:
            la, lb = getattr(sa, key), getattr(sb, key)
            assert len(la) == len(lb) > 0, key
            for x, y in zip(la, lb):
                assert np.array_equal(np.asarray(x), np.asarray(y)), (k_cl, key)
    say(f"sampled loop: {len(keys)} series of {len(runs[0])} clusters bit-identical")
