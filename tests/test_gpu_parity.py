"""Parity of the HIP step (through the C ABI) with the FP64 CPU oracle.

Bar (FP64 path): |dVmem| <= 1e-9 V and concentrations within 1e-10 relative
after the whole run; the FP32-flux path: |dVmem| < 1e-6 V.
"""

import os
import sys

import numpy as np
import pytest

from betse_amd.cells import Cells
from betse_amd.parameters import Parameters
from oracle.betse_oracle import OracleSim

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
VM_TOL = 1e-9          # V
REL_TOL = 1e-10        # relative, concentrations


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def gpu_run(p, cells, n, kind="init", flux_mode="fp64", use_graph=True, prior=None):
    from betse_amd.sim import SimPhase, SimPhaseKind, Simulator
    sim = Simulator(p.copy(), device="cuda:0", flux_mode=flux_mode, use_graph=use_graph)
    k = SimPhaseKind.INIT if kind == "init" else SimPhaseKind.SIM
    phase = SimPhase(k, sim.p, cells, sim)
    sim.init_core(phase)
    if prior is not None:
        prior(sim)
    sim.prepare(phase)
    sim.step_n(n)
    sim.pull()
    return sim


def oracle_run(p, cells, n, kind="init", prior=None):
    o = OracleSim(cells, p.copy(), mode="sparse")
    o.init_core()
    if prior is not None:
        prior(o)
    o.run_sim_core(kind, n_steps=n)
    return o


# E_env (both components) relative to the field's largest magnitude, floored
# at 1 uV/m: the env field comes from the env charge, which cancels like the
# cell charge, so its bar follows the Vmem bar (1e-9 V against fields of mV)
E_FLOOR = 1e-6         # V/m
E_TOL = 1e-8           # relative, FP64 path


def errors(sim, o):
    dv = float(np.abs(sim.vm - o.vm).max())
    dc = float((np.abs(sim.cc_cells - o.cc_cells) / np.abs(o.cc_cells).clip(1e-300)).max())
    de = float((np.abs(sim.cc_env - o.cc_env) / np.abs(o.cc_env).clip(1e-300)).max())
    lag = float((np.abs(sim.cc_at_mem - o.cc_at_mem) / np.abs(o.cc_at_mem).clip(1e-300)).max())
    emax = max(np.abs(o.E_env_x).max(), np.abs(o.E_env_y).max(), E_FLOOR)
    eex = float(max(np.abs(sim.E_env_x - o.E_env_x).max(), np.abs(sim.E_env_y - o.E_env_y).max()) / emax)
    return dv, dc, de, lag, eex


def check(sim, o, tag, vm_tol=VM_TOL, rel_tol=REL_TOL, e_tol=E_TOL):
    dv, dc, de, lag, eex = errors(sim, o)
    say(f"{tag}: |dVm|={dv:.3e} V  rel dc={dc:.3e}  rel denv={de:.3e}  rel dlag={lag:.3e}  rel dE={eex:.3e}")
    assert dv <= vm_tol
    assert dc <= rel_tol and de <= rel_tol and lag <= rel_tol
    assert eex <= e_tol


@pytest.fixture(scope="module")
def tiny():
    p = Parameters.preset("tiny")
    return p, Cells.make_world(p)


@pytest.fixture(scope="module")
def small():
    p = Parameters.preset("small")
    return p, Cells.load(os.path.join(ROOT, "tests", "golden", "world_small.npz"))


def test_tiny_fp64_parity(tiny):
    p, cells = tiny
    say("case: tiny fp64 50 steps")
    check(gpu_run(p, cells, 50), oracle_run(p, cells, 50), "tiny fp64")


def test_graph_equals_eager_and_odd_counts(tiny):
    p, cells = tiny
    say("case: graph vs eager, 37 steps")
    a = gpu_run(p, cells, 37, use_graph=True)
    b = gpu_run(p, cells, 37, use_graph=False)
    np.testing.assert_array_equal(a.cc_cells, b.cc_cells)
    np.testing.assert_array_equal(a.cc_env, b.cc_env)
    np.testing.assert_array_equal(a.vm_cell, b.vm_cell)
    check(a, oracle_run(p, cells, 37), "tiny odd")


def test_small_1000_steps_fp64(small):
    """The config-1 cluster (362 cells, 50x50 grid, Na/K/Cl/P/M), dt 1e-3, 1000 steps."""
    p, cells = small
    say("case: small fp64 1000 steps")
    check(gpu_run(p, cells, 1000), oracle_run(p, cells, 1000), "small fp64 1000")


def test_ragged_cells_and_tissue_profile():
    """Voronoi cells of random points (3..10 membranes) and a second tissue with other Dm."""
    p = Parameters.preset("tiny")
    rng = np.random.RandomState(11)
    pts = rng.uniform(-60e-6, 60e-6, (400, 2))
    cells = Cells.make_world(p, points=pts)
    assert len(set(cells.num_mems.tolist())) > 3
    p.tissue_profiles = [(np.arange(0, cells.n_cells, 3), {"Na": 5e-18, "K": 3e-17, "Cl": 1e-19})]
    say(f"case: ragged {cells.n_cells} cells, tissue profile, 80 steps")
    check(gpu_run(p, cells, 80), oracle_run(p, cells, 80), "ragged")


def test_calcium_and_sim_phase(tiny):
    """Ca ions (Ca-ATPase on post-diffusion env Ca) in the SIM phase after an INIT phase."""
    p = Parameters.preset("tiny")
    p.set_ion_profile({"Na": 8.0, "K": 125.0, "Cl": 10.0, "Ca": 1e-4, "P": 100.0},
                      {"Na": 145.0, "K": 5.0, "Cl": 105.0, "Ca": 1.0, "P": 10.0},
                      {"Na": 1e-18, "K": 15e-18, "Cl": 2e-18})
    cells = tiny[1]
    say("case: Ca init 30 + sim 30")
    o = oracle_run(p, cells, 30, "init")
    o.run_sim_core("sim", n_steps=30)

    from betse_amd.sim import SimPhase, SimPhaseKind, Simulator
    sim = Simulator(p.copy(), device="cuda:0")
    ph = SimPhase(SimPhaseKind.INIT, sim.p, cells, sim)
    sim.init_core(ph)
    sim.prepare(ph)
    sim.step_n(30)
    sim.pull()
    ph2 = SimPhase(SimPhaseKind.SIM, sim.p, cells, sim)
    sim.prepare(ph2)
    sim.step_n(30)
    sim.pull()
    check(sim, o, "Ca init+sim")


def test_run_sim_core_sampling(tiny):
    """run_sim_core drives sampled steps and write2storage like the reference loop."""
    p, cells = tiny
    from betse_amd.sim import SimPhase, SimPhaseKind, Simulator
    say("case: run_sim_core with sampling")
    sim = Simulator(p.copy(), device="cuda:0")
    ph = SimPhase(SimPhaseKind.INIT, sim.p, cells, sim)
    sim.init_core(ph)
    sim.run_sim_core(ph)
    o = OracleSim(cells, p.copy(), mode="sparse")
    o.init_core()
    idx = o.run_sim_core("init")
    assert len(sim.time) == len(idx) == len(o.storage) == ph.callbacks.progress
    for a, b in zip(sim.vm_time, o.storage):
        assert np.abs(a - b["vm"]).max() <= VM_TOL
    for a, b in zip(sim.cc_time, o.storage):
        np.testing.assert_allclose(a, b["cc_cells"], rtol=REL_TOL)
    check(sim, o, "run_sim_core")


@pytest.mark.parametrize("tuning", [None, {"persistent": 1}], ids=["graphs", "persistent"])
def test_nan_raises_unstable(tiny, tuning):
    """A NaN written into one device cell concentration stops the run with
    BetseSimUnstableException at that step (replayed graphs and the
    persistent step kernel)."""
    p, cells = tiny
    from betse_amd.exceptions import BetseSimUnstableException
    from betse_amd.sim import SimPhase, SimPhaseKind, Simulator
    say(f"case: NaN -> unstable, tuning {tuning}")
    sim = Simulator(p.copy(), device="cuda:0", tuning=tuning)
    ph = SimPhase(SimPhaseKind.INIT, sim.p, cells, sim)
    sim.init_core(ph)
    sim.prepare(ph)
    sim.step_n(5)
    import torch
    with torch.cuda.stream(sim.stream):
        sim._dev["c_cell"][1, 7] = float("nan")
    sim.stream.synchronize()
    with pytest.raises(BetseSimUnstableException) as ei:
        sim.step_n(10)
    assert ei.value.step == 5
    # the state is frozen at the failing step: further steps do nothing
    with torch.cuda.stream(sim.stream):
        snap = sim._dev["c_env"].clone()
    sim.stream.synchronize()
    with pytest.raises(BetseSimUnstableException):
        sim.step_n(4)
    assert bool((sim._dev["c_env"] == snap).all())


def test_small_fp32_flux_error(small):
    """FP32 flux laws, FP64 state/charge: Vmem within 1 uV of the FP64 oracle
    after 1000 steps; the per-field errors are printed."""
    p, cells = small
    say("case: small fp32 flux 1000 steps")
    sim = gpu_run(p, cells, 1000, flux_mode="fp32")
    o = oracle_run(p, cells, 1000)
    dv, dc, de, lag, eex = errors(sim, o)
    say(f"fp32 flux: |dVm|={dv:.3e} V rel dc={dc:.3e} rel denv={de:.3e} rel dE={eex:.3e}")
    assert dv < 1e-6


def test_channels_fp64_parity():
    """Nav1p2 + Kv1p1 on every membrane, per-cell multipliers, U[0,1] gates, Ca
    (the large-cluster physics on the small test cluster), init then sim."""
    from betse_amd.channels import draw_channel_state
    p = Parameters.preset("tiny_channels")
    cells = Cells.make_world(p)
    draw_channel_state(p, cells.n_cells, cells.n_mems)
    say("case: channels init 150 + sim 50")
    o = oracle_run(p, cells, 150, "init")
    o.run_sim_core("sim", n_steps=50)
    from betse_amd.sim import SimPhase, SimPhaseKind, Simulator
    sim = Simulator(p.copy(), device="cuda:0")
    ph = SimPhase(SimPhaseKind.INIT, sim.p, cells, sim)
    sim.init_core(ph)
    sim.prepare(ph)
    sim.step_n(150)
    sim.pull()
    ph2 = SimPhase(SimPhaseKind.SIM, sim.p, cells, sim)
    sim.prepare(ph2)
    sim.step_n(50)
    sim.pull()
    check(sim, o, "channels")
    for k, ch in enumerate(o.channels):
        np.testing.assert_allclose(sim.chan_gates[k, 0], ch.m, rtol=1e-10, atol=1e-14)
        np.testing.assert_allclose(sim.chan_gates[k, 1], ch.h, rtol=1e-10, atol=1e-14)


def test_many_tiles_single_wavefront_path():
    """A cluster with more than 4096 cell tiles runs the membrane kernel with one
    wavefront per tile (smaller ones split the ions over two): both FP64 and
    FP32 flux paths against the oracle over 20 steps."""
    from betse_amd.channels import draw_channel_state
    from betse_amd.sim import make_tiles
    p = Parameters.preset("large")
    p.cluster_radius = 1.1e-3
    cells = Cells.make_world(p)
    assert len(make_tiles(cells.mem_ptr)) - 1 > 4096
    draw_channel_state(p, cells.n_cells, cells.n_mems)
    say(f"case: {cells.n_cells} cells, single-wavefront tiles, 20 steps")
    o = oracle_run(p, cells, 20)
    for mode, vm_tol, rel in (("fp64", VM_TOL, REL_TOL), ("fp32", 1e-6, 1e-6)):
        sim = gpu_run(p, cells, 20, flux_mode=mode)
        dv, dc, de, lag, eex = errors(sim, o)
        say(f"  {mode}: |dVm|={dv:.3e} rel dc={dc:.3e} rel denv={de:.3e}")
        assert dv <= vm_tol and dc <= rel and de <= rel


def test_env_one_block_per_tile_path():
    """A grid of 1024 or more 16x16 env tiles runs the env kernel with one block
    per tile looping over all ions, each next ion's loads issued during the
    previous ion's stages (smaller grids split the ions into pairs): the large
    cluster on a 512x512 grid, FP64 and FP32 flux paths against the oracle
    over 16 steps (one replay of the eight-pair step graph)."""
    from betse_amd.channels import draw_channel_state
    p = Parameters.preset("large")
    p.grid_size = 512
    cells = Cells.make_world(p)
    ny, nx = cells.X.shape
    assert ((ny + 15) // 16) * ((nx + 15) // 16) >= 1024
    draw_channel_state(p, cells.n_cells, cells.n_mems)
    say(f"case: {cells.n_cells} cells on a {ny}x{nx} grid, one env block per tile, 16 steps")
    o = oracle_run(p, cells, 16)
    for mode, vm_tol, rel in (("fp64", VM_TOL, REL_TOL), ("fp32", 1e-6, 1e-6)):
        sim = gpu_run(p, cells, 16, flux_mode=mode)
        dv, dc, de, lag, eex = errors(sim, o)
        say(f"  {mode}: |dVm|={dv:.3e} rel dc={dc:.3e} rel denv={de:.3e}")
        assert dv <= vm_tol and dc <= rel and de <= rel


def test_channels_fp32_flux_parity():
    """The mixed-precision path with Nav1p2 + Kv1p1 (FP32 flux laws, FP64
    gating and state): |dVm| < 1 uV and gates within 1e-5 of the FP64 oracle
    after init 150 + sim 100 steps."""
    from betse_amd.channels import draw_channel_state
    p = Parameters.preset("tiny_channels")
    cells = Cells.make_world(p)
    draw_channel_state(p, cells.n_cells, cells.n_mems)
    say("case: channels fp32 flux init 150 + sim 100")
    o = oracle_run(p, cells, 150, "init")
    o.run_sim_core("sim", n_steps=100)
    from betse_amd.sim import SimPhase, SimPhaseKind, Simulator
    sim = Simulator(p.copy(), device="cuda:0", flux_mode="fp32")
    ph = SimPhase(SimPhaseKind.INIT, sim.p, cells, sim)
    sim.init_core(ph)
    sim.prepare(ph)
    sim.step_n(150)
    sim.pull()
    ph2 = SimPhase(SimPhaseKind.SIM, sim.p, cells, sim)
    sim.prepare(ph2)
    sim.step_n(100)
    sim.pull()
    dv, dc, de, lag, eex = errors(sim, o)
    say(f"channels fp32: |dVm|={dv:.3e} V rel dc={dc:.3e} rel denv={de:.3e}")
    assert dv < 1e-6
    assert dc < 1e-6 and de < 1e-6
    for k, ch in enumerate(o.channels):
        gm = np.abs(sim.chan_gates[k, 0] - ch.m).max()
        gh = np.abs(sim.chan_gates[k, 1] - ch.h).max()
        say(f"  gate {ch.kind}: max |dm| {gm:.2e} |dh| {gh:.2e}")
        assert gm < 1e-5 and gh < 1e-5


def test_full_physics_fp64_parity():
    """Nav1p2 + Kv1p1 + Cav1p2 and the 8-substance reaction network on the test
    cluster, init then sim, against the oracle (fp64 bars, substances 1e-10)."""
    from betse_amd.channels import draw_channel_state
    p = Parameters.preset("tiny_full")
    cells = Cells.make_world(p)
    draw_channel_state(p, cells.n_cells, cells.n_mems)
    say("case: full physics (Nav/Kv/Cav + network) init 150 + sim 100")
    o = oracle_run(p, cells, 150, "init")
    o.run_sim_core("sim", n_steps=100)
    from betse_amd.sim import SimPhase, SimPhaseKind, Simulator
    sim = Simulator(p.copy(), device="cuda:0")
    ph = SimPhase(SimPhaseKind.INIT, sim.p, cells, sim)
    sim.init_core(ph)
    sim.prepare(ph)
    sim.step_n(150)
    sim.pull()
    ph2 = SimPhase(SimPhaseKind.SIM, sim.p, cells, sim)
    sim.prepare(ph2)
    sim.step_n(100)
    sim.pull()
    check(sim, o, "full physics")
    for k, ch in enumerate(o.channels):
        np.testing.assert_allclose(sim.chan_gates[k, 0], ch.m, rtol=1e-10, atol=1e-14)
        np.testing.assert_allclose(sim.chan_gates[k, 1], ch.h, rtol=1e-10, atol=1e-14)
    rel = np.abs(sim.c_net - o.c_net) / np.maximum(np.abs(o.c_net), 1e-30)
    say(f"network substances: max rel err {rel.max():.3e}, moved by {np.abs(o.c_net - p.network.c0[:, None]).max():.3e} mM")
    assert rel.max() < 1e-10
    assert np.abs(o.c_net - p.network.c0[:, None]).max() > 1e-6     # the network did something


def test_network_kernel_alone_matches_oracle():
    """The network kernel over many steps with substances spread per cell, on a
    drawn network that stays finite (the oracle's no_negs raises on NaN)."""
    import ctypes
    import torch
    from betse_amd import _native
    from betse_amd.network import draw_network
    from oracle import network_oracle as no
    from betse_amd.sim import SimPhase, SimPhaseKind, Simulator
    p = Parameters.preset("tiny")
    net = draw_network(seed=13, n_species=8, n_reactions=20)
    p.network = net
    cells = Cells.make_world(p)
    sim = Simulator(p.copy(), device="cuda:0")
    ph = SimPhase(SimPhaseKind.INIT, sim.p, cells, sim)
    sim.init_core(ph)
    rng = np.random.RandomState(3)
    sim.c_net = sim.c_net * rng.uniform(0.2, 5.0, sim.c_net.shape)
    C = sim.c_net.copy()
    sim.prepare(ph)
    say("case: network kernel alone, 8 substances, 20 reactions, 2000 launches")
    # bt_time_kernel runs one untimed launch and then the timed ones: 2000 network steps of p.dt
    ms = ctypes.c_double()
    assert sim._lib.bt_time_kernel(sim._ctx, _native.KERNEL_IDS["network"], 1999, ctypes.byref(ms)) == _native.BT_OK
    sim.stream.synchronize()
    got = sim._dev["c_net"].cpu().numpy()
    for _ in range(2000):
        C = no.step(net, C, p.dt, p.R, sim.T)
    rel = np.abs(got - C) / np.maximum(np.abs(C), 1e-30)
    say(f"network alone: max rel err {rel.max():.3e} over {C.size} values")
    assert np.all(np.isfinite(got))
    assert rel.max() < 1e-10
    assert np.abs(C - sim.c_net).max() > 1e-3       # the substances moved


def test_network_nan_halts_like_reference():
    """A substance that turns NaN raises in the reference's no_negs
    (sim_toolbox.py:475-503, networks.py:2939): here a reversible reaction
    whose reactant and product are both exactly zero in one cell (Q = 0/0)
    after three ordinary steps. The GPU run halts with
    BetseSimUnstableException at the step where the oracle raises, and the
    state stays frozen there."""
    import torch
    import warnings
    from betse_amd.exceptions import BetseSimUnstableException
    from betse_amd.network import draw_network
    from betse_amd.sim import SimPhase, SimPhaseKind, Simulator
    from oracle.betse_oracle import OracleUnstable
    p = Parameters.preset("tiny")
    p.network = draw_network(seed=13, n_species=8, n_reactions=20)
    rev = [r for r in p.network.reactions if r.delta_Go is not None]
    assert rev, "the drawn network needs a reversible reaction"
    a_sp, b_sp = rev[0].reactants[0][0], rev[0].products[0][0]
    cells = Cells.make_world(p)
    cell = 7
    o = OracleSim(cells, p.copy(), mode="sparse")
    o.init_core()
    o.init_dynamics("init")
    o.p.set_phase_time("init")
    o.update_V(observables=False)
    for _ in range(3):
        o.step("init", observables=False)
    o.c_net[a_sp, cell] = 0.0
    o.c_net[b_sp, cell] = 0.0
    bad_at = None
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        for s in range(3, 20):
            try:
                o.step("init", observables=False)
            except OracleUnstable:
                bad_at = s
                break
    assert bad_at == 3
    say(f"case: network NaN (Q = 0/0), oracle raises at step {bad_at}")
    sim = Simulator(p.copy(), device="cuda:0")
    ph = SimPhase(SimPhaseKind.INIT, sim.p, cells, sim)
    sim.init_core(ph)
    sim.prepare(ph)
    sim.step_n(3)
    with torch.cuda.stream(sim.stream):
        sim._dev["c_net"][a_sp, cell] = 0.0
        sim._dev["c_net"][b_sp, cell] = 0.0
    sim.stream.synchronize()
    with pytest.raises(BetseSimUnstableException) as ei:
        sim.step_n(10)
    assert ei.value.step == bad_at
    frozen = sim.cc_cells.copy()
    with pytest.raises(BetseSimUnstableException):
        sim.step_n(4)
    assert np.array_equal(sim.cc_cells, frozen)


def test_mfma_gemm_f64():
    """The FP64 MFMA kernel behind the Poisson observables against numpy."""
    import ctypes
    import torch
    from betse_amd import _native
    from betse_amd.sim import SimPhase, SimPhaseKind, Simulator
    p = Parameters.preset("tiny")
    cells = Cells.make_world(p)
    sim = Simulator(p.copy(), device="cuda:0")
    ph = SimPhase(SimPhaseKind.INIT, sim.p, cells, sim)
    sim.init_core(ph)
    sim.prepare(ph)
    rng = np.random.RandomState(2)
    # inner dimensions whose quarters take the tail loop only (48), one block of
    # sixteen k-steps (256), two blocks (512), and one block plus the tail
    # (272, 304: quarters of 68 and 76); 256 and 512 are the padded DST sizes
    # of the 256x256 and 512x512 grids
    for m, k, n in ((32, 48, 64), (256, 256, 256), (64, 512, 32), (32, 272, 48), (304, 304, 304),
                    (512, 512, 512)):
        say(f"case: mfma gemm {m}x{k}x{n}")
        a = rng.randn(m, k)
        b = rng.randn(k, n)
        with torch.cuda.stream(sim.stream):
            A = torch.tensor(a, device="cuda:0")
            B = torch.tensor(b, device="cuda:0")
            Cm = torch.full((m, n), float("nan"), dtype=torch.float64, device="cuda:0")
        sim.stream.synchronize()
        rc = sim._lib.bt_gemm_f64(sim._ctx, ctypes.c_void_p(A.data_ptr()), ctypes.c_void_p(B.data_ptr()),
                                  ctypes.c_void_p(Cm.data_ptr()), m, n, k)
        assert rc == _native.BT_OK
        sim.stream.synchronize()
        np.testing.assert_allclose(Cm.cpu().numpy(), a @ b, rtol=1e-12, atol=1e-11)


def jenv_bound(rec, p, scale):
    """Error bar of the env current J_env = F sum_i z_i fluxes_env_i (and of its
    Helmholtz-Hodge parts): the ion currents cancel to a small fraction of
    their terms in the near-neutral environment, so an ulp or two of each
    ion's flux (the env kernel multiplies by 1/delta where the reference
    divides) is a large part of J_env. The bar is 1e-8 of |J_env| plus four
    ulps of the term scale F sum_i |z_i| max|fluxes_env_i|."""
    fx, fy = np.asarray(rec["fluxes_env_x"]), np.asarray(rec["fluxes_env_y"])
    z = np.array([p.ion_charge[n] for n in p.ion_names()], dtype=float)
    terms = p.F * float(np.sum(np.abs(z) * np.maximum(np.abs(fx).max(axis=1), np.abs(fy).max(axis=1))))
    return 1e-8 * scale + 4 * np.finfo(float).eps * terms


def jenv_rel(rec, p):
    """The env current's bar relative to its largest magnitude (``jenv_bound``
    over max|J_env|): the bar of the fields linear in J_env (its
    Helmholtz-Hodge parts, B_field)."""
    scale = max(np.abs(np.asarray(rec["J_env_x"])).max(), np.abs(np.asarray(rec["J_env_y"])).max(), 1e-300)
    return jenv_bound(rec, p, scale) / scale


@pytest.mark.parametrize("preset,n_steps", [("tiny", 40), ("tiny_channels", 40), ("large", 20)])
def test_sampled_observables(preset, n_steps):
    """Currents, E_gj, pump rate and the Helmholtz-Hodge env currents at every
    sampled step against the oracle's get_current (sparse Poisson solve); the
    large cluster's 256x256 grid pads its 254x254 interior to 256 for the
    DST products."""
    from betse_amd.channels import draw_channel_state
    from betse_amd.sim import SimPhase, SimPhaseKind, Simulator
    p = Parameters.preset(preset)
    cells = Cells.make_world(p)
    if p.channels:
        draw_channel_state(p, cells.n_cells, cells.n_mems)
    p.init_time_total = n_steps * p.init_time_step
    p.init_time_sampling = 10 * p.init_time_step
    say(f"case: observables {preset}")
    sim = Simulator(p.copy(), device="cuda:0")
    ph = SimPhase(SimPhaseKind.INIT, sim.p, cells, sim)
    sim.init_core(ph)
    sim.run_sim_core(ph)
    o = OracleSim(cells, p.copy(), mode="sparse")
    o.init_core()
    o.run_sim_core("init")
    assert len(o.storage) == len(sim.I_mem_time) > 0
    pairs = {"I_mem_time": "I_mem", "I_cell_x_time": "J_cell_x", "I_cell_y_time": "J_cell_y",
             "efield_gj_x_time": "E_gj_x", "rate_NaKATP_time": "rate_NaKATP", "I_tot_x_time": "J_env_x",
             "I_tot_y_time": "J_env_y"}
    for rec_i, rec in enumerate(o.storage):
        # J_cell is a weighted mean of Jn*n over a cell's membranes (weights
        # |n| sa / cell_sa summing to at most 1) whose terms nearly cancel, so
        # its rounding scale is max|Jn|, not max|J_cell|; E_gj is a difference
        # of neighbouring Vmem, so its error is bounded by twice the Vmem bar
        # (1e-9 V) over gj_len.
        jn_scale = np.abs(np.asarray(rec["I_mem"]) / cells.mem_sa).max()
        for mine, theirs in pairs.items():
            a = np.asarray(getattr(sim, mine)[rec_i]).ravel()
            b = np.asarray(rec[theirs]).ravel()
            scale = max(np.abs(b).max(), 1e-300)
            if mine.startswith("I_cell"):
                scale = max(scale, jn_scale)
            err = np.abs(a - b).max()
            if mine.startswith("efield_gj"):
                assert err <= max(1e-8 * scale, 2e-9 / cells.gj_len), (mine, rec_i, err, scale)
            elif mine.startswith("I_tot"):
                bound = jenv_bound(rec, p, scale)
                if rec_i == 0:
                    say(f"  {mine}: err {err:.3e}, |J| {scale:.3e}, bar {bound:.3e}")
                assert err <= bound, (mine, rec_i, err, scale, bound)
            else:
                assert err / scale < 1e-8, (mine, rec_i, err / scale)
    rel = jenv_rel(o.storage[-1], p)
    np.testing.assert_allclose(sim.B_field.ravel(), o.B_field.ravel(), rtol=0, atol=rel * np.abs(o.B_field).max())
    np.testing.assert_allclose(sim.Jtx.ravel(), o.Jtx.ravel(), rtol=0, atol=rel * np.abs(o.Jtx).max())
    say(f"observables {preset}: ok over {len(o.storage)} sampled steps")


def test_large_fp32_flux_full_horizon():
    """The headline workload over its whole run: the large cluster (Nav1p2 +
    Kv1p1, Ca) for 10,000 steps on the FP32-flux path stays within 1 uV of the
    FP64 path (which the tests above hold to the oracle; the oracle itself
    would take minutes here), and both stay finite and charge-consistent."""
    from betse_amd.channels import draw_channel_state
    p = Parameters.preset("large")
    cells = Cells.make_world(p)
    draw_channel_state(p, cells.n_cells, cells.n_mems)
    say(f"case: large {cells.n_cells} cells, fp32 vs fp64, 10000 steps")
    s64 = gpu_run(p, cells, 10000, flux_mode="fp64")
    s32 = gpu_run(p, cells, 10000, flux_mode="fp32")
    dv = float(np.abs(s32.vm - s64.vm).max())
    dc = float((np.abs(s32.cc_cells - s64.cc_cells) / np.abs(s64.cc_cells).clip(1e-300)).max())
    say(f"  fp32 vs fp64 after 10000 steps: |dVm|={dv:.3e} V rel dc={dc:.3e}")
    assert np.isfinite(s64.vm).all() and np.isfinite(s32.vm).all()
    assert dv < 1e-6


def test_coarse_grid_many_membranes_per_square():
    """An 8x8 grid under the 35-cell cluster: up to 17 membranes share a square,
    so the exchange sums a square's fluxes over several batches of loads, and
    most squares carry membranes. FP64 against the oracle over 300 steps."""
    p = Parameters.preset("tiny")
    p.grid_size = 8
    cells = Cells.make_world(p)
    counts = np.bincount(cells.map_mem2ecm, minlength=cells.X.size)
    assert counts.max() > 16
    say(f"case: coarse grid {cells.X.shape}, up to {counts.max()} membranes per square, 300 steps")
    sim = gpu_run(p, cells, 300)
    o = oracle_run(p, cells, 300)
    check(sim, o, "coarse grid")


def test_nonuniform_env_diffusivity():
    """Adherens and tight junctions lower D_env inside the cluster and at its
    boundary (initDenv, sim.py:2347-2400), so D_env is not one value per ion and
    the env kernel loads it per square. FP64 and the FP32-flux path against the
    oracle over 200 steps."""
    p = Parameters.preset("tiny")
    p.D_adh = 0.5
    p.D_tj = 0.3
    cells = Cells.make_world(p)
    say("case: non-uniform D_env (D_adh 0.5, D_tj 0.3), 200 steps")
    o = oracle_run(p, cells, 200)
    assert len(np.unique(o.D_env[0])) > 1
    check(gpu_run(p, cells, 200), o, "non-uniform D_env fp64")
    dv, dc, de, lag, eex = errors(gpu_run(p, cells, 200, flux_mode="fp32"), o)
    say(f"  fp32: |dVm|={dv:.3e} rel denv={de:.3e}")
    assert dv < 1e-6 and de < 1e-6


def test_static_noise_kept_in_init_dropped_in_sim():
    """A nonzero static K channel noise (channel_noise_level): init_dynamics
    draws np.random.random(M) in each phase and scales Dm_K by it, and the SIM
    phase's makeAllChanges resets Dm to Dm_scheduled + Dm_base, the copy taken
    before the noise (Q10: sim.py:825, 837-839; tishandler.py:1321-1332). The
    GPU run and the oracle start from the same global numpy seed."""
    p = Parameters.preset("tiny")
    p.channel_noise_level = 0.8
    cells = Cells.make_world(p)
    say("case: static noise 0.8, init 30 + sim 30")
    np.random.seed(1234)
    o = oracle_run(p, cells, 30, "init")
    dm_init = o.Dm_cells.copy()
    o.run_sim_core("sim", n_steps=30)
    iK = o.names.index("K") if hasattr(o, "names") else 1
    assert np.abs(dm_init[iK] / o.Dm_base[iK] - 1).max() > 0.1      # the init phase was noisy
    assert np.array_equal(o.Dm_cells, o.Dm_scheduled + o.Dm_base)    # the sim phase is not

    from betse_amd.sim import SimPhase, SimPhaseKind, Simulator
    np.random.seed(1234)
    sim = Simulator(p.copy(), device="cuda:0")
    ph = SimPhase(SimPhaseKind.INIT, sim.p, cells, sim)
    sim.init_core(ph)
    sim.prepare(ph)
    np.testing.assert_array_equal(sim.Dm_cells, dm_init)
    sim.step_n(30)
    sim.pull()
    ph2 = SimPhase(SimPhaseKind.SIM, sim.p, cells, sim)
    sim.prepare(ph2)
    sim.step_n(30)
    sim.pull()
    check(sim, o, "static noise init+sim")


def oracle_stepper(p, cells, kind="init"):
    """The oracle set up as run_sim_core sets it up, stepped one step at a time."""
    o = OracleSim(cells, p.copy(), mode="sparse")
    o.init_core()
    o.init_dynamics(kind)
    o.p.set_phase_time(kind)
    o.update_V(observables=False)
    return o


def event_params(o, cells):
    """An event's new parameters: Na Dm tripled on every other membrane (no
    longer one value over the cluster), K Dm halved everywhere, gap junctions
    half blocked (gj_block = 0.5)."""
    iNa, iK = o.names.index("Na"), o.names.index("K")
    Dm = o.Dm_cells.copy()
    Dm[iNa, ::2] *= 3.0
    Dm[iK] *= 0.5
    gjopen = 0.5 * np.ones(cells.n_mems) * cells.gj_default_weights
    return Dm, gjopen


def test_event_changes_membrane_params(tiny):
    """An event in the middle of a phase (TissueHandler.fire_events ->
    makeAllChanges changing Dm and the gap-junction block): the GPU run
    (set_membrane_params -> bt_refresh_params, replayed graphs kept) matches
    the oracle stepped through the same change, at the FP64 bars."""
    p, cells = tiny
    from betse_amd.sim import SimPhase, SimPhaseKind, Simulator
    say("case: event mid-phase (Dm, gap-junction block), 10 + 15 steps")
    sim = Simulator(p.copy(), device="cuda:0")
    ph = SimPhase(SimPhaseKind.INIT, sim.p, cells, sim)
    sim.init_core(ph)
    sim.prepare(ph)
    sim.step_n(10)
    o = oracle_stepper(p, cells)
    for _ in range(10):
        o.step("init", observables=False)
    Dm, gjopen = event_params(o, cells)
    o.Dm_cells = Dm.copy()
    o.gj_block = 0.5
    sim.set_membrane_params(Dm=Dm, gjopen=gjopen)
    for _ in range(15):
        o.step("init", observables=False)
    sim.step_n(15)
    sim.pull()
    check(sim, o, "after the event")
    # the event changed the trajectory: without it the run differs
    ref = gpu_run(p, cells, 25)
    assert np.abs(ref.cc_cells - sim.cc_cells).max() > 1e-6 * np.abs(sim.cc_cells).max()
