"""Benchmark of the electrodiffusion step on one MI355X.

Workload (default): the single large cluster -- jittered hex lattice (10 um
spacing, +-10% jitter, numpy seed 0) clipped to an 800 um disc (23,222 cells,
139,332 membranes), 256x256 environment grid, ions Na/K/Cl/Ca/P/M, dt 1e-4 s.
A step is one iteration of the reference loop over that cluster; the timed
region is ``--steps`` steps with all state resident in HBM.

Printed: one JSON line with cell-steps/s, the roofline of the dominant kernel
(algorithmic bytes per launch / HIP-event launch time), and the CPU baseline
(the FP64 NumPy oracle, timed on a bounded prefix of the same workload on this
host), plus |dVmem| of the GPU path against that oracle on the same prefix.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def _squares_with_membranes(cells) -> int:
    if hasattr(cells, "n_live"):
        return int(cells.n_live)
    return int(np.count_nonzero(np.diff(cells.sq_ptr)))


def algorithmic_bytes(cells, n_ions: int, ca: bool, flux_mode: str, n_ch: int = 0, n_sub: int = 0,
                      dm_loaded: int = None, gj_loaded: bool = True, denv_loaded: int = None) -> dict:
    """Compulsory bytes per launch of each step kernel (each array element the
    kernel must touch, counted once at its storage width). ``dm_loaded`` ions
    have per-membrane Dm (the others are one value per cluster and come with
    the cluster constants); ``gj_loaded`` likewise for gjopen, ``denv_loaded``
    for D_env over the grid."""
    C, M = cells.n_cells, cells.n_mems
    G2 = cells.X.size
    n_in = len(cells.envInds_inClust)
    ni = n_ions
    ndm = ni if dm_loaded is None else dm_loaded
    b = {}
    b["mem_cell"] = (M * (5 * 4 + 8 + (8 if gj_loaded else 0) + 8 * ndm)   # indices, mem_sa, gjopen, Dm
                     + M * 8 * ni                           # fsa write
                     + M * n_ch * (16 + 16 + 8)              # gates r/w, channel flux write
                     + C * (8 * ni * 2 + 8 * ni + 8 * ni     # c_cell r, c_lag r, c_cell w, c_lag w
                            + 16 + 16 + 4 + 8 * n_ch         # vm r/w, cell_vol, diviterm, mem_ptr, modulators
                            + 32 * n_ch)                     # gating rates r
                     + n_in * 8 * (ni + (1 if ca else 0) + n_ch))  # env gathers (pre- and post-diffusion)
    nde = ni if denv_loaded is None else denv_loaded
    b["env_field"] = (G2 * (8 + 8 * ni + 8 * nde + 8 * ni)  # u r, c_env r, D_env r, c_mid / next c_env w
                      + (C * (8 + 32 * n_ch) if n_ch else 0))  # vm r, gating rates w
    # the exchange runs over the squares with membranes (the others take their
    # new state from k_env_field): c_mid r, c_env w, u w, has_mem, sq_cluster, sq_ptr, list
    n_live = _squares_with_membranes(cells)
    b["env_exchange"] = n_live * (8 * ni + 8 * ni + 8 + 1 + 4 + 4 + 4) + M * 8 * (ni + n_ch)
    if n_sub:
        b["network"] = C * n_sub * 16                          # substances read and written
    b["step"] = sum(b.values())
    return b


class _Tot:
    """Totals of a batch of clusters, shaped like Cells for algorithmic_bytes."""

    def __init__(self, worlds):
        self.n_cells = sum(c.n_cells for _, c in worlds)
        self.n_mems = sum(c.n_mems for _, c in worlds)
        self.envInds_inClust = range(sum(len(c.envInds_inClust) for _, c in worlds))
        self.n_live = sum(int(np.count_nonzero(np.diff(c.sq_ptr))) for _, c in worlds)

        class _X:
            size = sum(c.X.size for _, c in worlds)
        self.X = _X()


def bench_batched(a):
    """Batched workload: 256 independent clusters of 100-5000 cells."""
    import torch
    from betse_amd.batch import BatchSimulator, batched_worlds
    from oracle.betse_oracle import OracleSim
    t0 = time.time()
    worlds = batched_worlds(a.clusters)
    tot = _Tot(worlds)
    C = tot.n_cells
    p0 = worlds[0][0]
    n_ions = len(p0.ion_names())
    print(f"# batch: {len(worlds)} clusters, {C} cells, {tot.n_mems} membranes, {tot.X.size} squares, "
          f"built in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    modes = ["fp32", "fp64"] if a.flux == "both" else [a.flux]
    res = {}
    for mode in modes:
        b = BatchSimulator([(p.copy(), c) for p, c in worlds], device="cuda:0", flux_mode=mode)
        b.init("init")
        uniform = all(np.all(st.Dm_cells == st.Dm_cells[:, :1]) for st in b.states)
        de_uniform = all(np.all((st.D_env * st.TJ_modulator) == (st.D_env * st.TJ_modulator)[:, :1])
                         for st in b.states)
        b.step_n(a.warmup)
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(b.stream)
        b.step_n(a.steps, sync=False)
        e1.record(b.stream)
        b.sync()
        ev_ms = e0.elapsed_time(e1)
        kern = {k: b.time_kernel(k, 20) for k in ("mem_cell", "env_field", "env_exchange")}
        res[mode] = dict(ev_ms=ev_ms, kernels=kern)
        print(f"# {mode}: {ev_ms / a.steps * 1e3:.1f} us/step, kernels(us): "
              + ", ".join(f"{k}={v * 1e3:.1f}" for k, v in kern.items()), file=sys.stderr, flush=True)
        del b
    cpu, parity = None, {}
    if not a.no_cpu:
        sub = worlds[:a.cpu_clusters]
        # time one step per cluster to size the sample
        n_cpu = 2
        t0 = time.perf_counter()
        oracles = []
        for p, c in sub:
            o = OracleSim(c, p.copy(), mode="sparse")
            o.init_core()
            o.init_dynamics("init")
            o.p.set_phase_time("init")
            o.update_V(observables=False)
            o.step("init", observables=False)
            oracles.append(o)
        one = time.perf_counter() - t0
        n_cpu = max(2, int(a.cpu_seconds / max(one, 1e-6)))
        t0 = time.perf_counter()
        for o in oracles:
            for _ in range(n_cpu - 1):
                o.step("init", observables=False)
        el = one + time.perf_counter() - t0
        csub = sum(c.n_cells for _, c in sub)
        cpu = {"value": csub * n_cpu / el, "unit": "cell-steps/s", "cores": 1, "kind": "port",
               "sample": f"FP64 NumPy oracle (sparse operators), the first {len(sub)} clusters "
                         f"({csub} cells) for {n_cpu} steps each, {el:.1f}s on this host, single thread"}
        for mode in modes:
            b = BatchSimulator([(p.copy(), c) for p, c in sub], device="cuda:0", flux_mode=mode)
            b.init("init")
            b.step_n(n_cpu)
            b.pull()
            parity[mode] = {"steps": n_cpu, "clusters": len(sub),
                            "max_abs_dVm_V": float(max(np.abs(st.vm - o.vm).max() for st, o in zip(b.states, oracles))),
                            "max_rel_dc_cell": float(max((np.abs(st.cc_cells - o.cc_cells) / np.abs(o.cc_cells)).max()
                                                         for st, o in zip(b.states, oracles)))}
            del b
        print(f"# cpu oracle sample: {cpu['sample']}; parity {parity}", file=sys.stderr, flush=True)
    ok = [m for m in modes if (not parity) or parity[m]["max_abs_dVm_V"] < 1e-6]
    best = min(ok or modes, key=lambda m: res[m]["ev_ms"])
    r = res[best]
    ms_step = r["ev_ms"] / a.steps
    # Dm (per-cluster factors) and gjopen are one value per cluster in this workload
    ab = algorithmic_bytes(tot, n_ions, "Ca" in p0.ion_names(), best, len(p0.channels), 0,
                           0 if uniform else n_ions, not uniform, 0 if de_uniform else n_ions)
    kern = r["kernels"]
    dom = max(kern, key=kern.get)
    achieved = ab[dom] / (kern[dom] * 1e-3) / 1e9
    line = {
        "metric": "cell-steps/s at |dVmem|<1uV vs FP64 oracle", "value": C * a.steps / (r["ev_ms"] / 1e3),
        "unit": "cell-steps/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": ms_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if best == "fp64" else "f32-flux/f64-state", "data": "synthetic",
        "config": {"workload": f"batched: {len(worlds)} independent clusters, {C} cells, {tot.n_mems} membranes, "
                               f"grids 32-128 ({tot.X.size} squares), ions {'/'.join(p0.ion_names())}, "
                               f"per-cluster Dm and NaK rate over one decade, dt {p0.dt:g} s",
                   "clusters": len(worlds), "cells": C, "flux_mode": best},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic("batched", dom, best),
                     "launch_us": kern[dom] * 1e3,
                     "algorithmic_bytes": ab[dom], "step_algorithmic_bytes": ab["step"],
                     "step_achieved_GBs": ab["step"] / (ms_step * 1e-3) / 1e9},
        "cpu_baseline": cpu, "parity_prefix": parity,
        "paths": {m: {"cell_steps_per_s": C * a.steps / (res[m]["ev_ms"] / 1e3),
                      "us_per_step": res[m]["ev_ms"] / a.steps * 1e3,
                      "kernels_us": {k: v * 1e3 for k, v in res[m]["kernels"].items()}} for m in modes},
    }
    print(json.dumps(line), flush=True)


def bench_warmstart(a):
    """Learned warm start: init-phase steady states of random clusters from the
    plain method (labels), the network trained on the GPU, then init time to
    the residual tolerance on held-out clusters, cold versus warm."""
    from betse_amd import warmstart as ws

    def log(msg):
        print("# " + msg, file=sys.stderr, flush=True)

    t0 = time.perf_counter()
    data, stats, base = ws.make_dataset(a.train_clusters, seed=1000, batch=a.clusters, horizon=a.label_horizon,
                                        flux_mode=a.ws_flux, log=log)
    t_data = time.perf_counter() - t0
    n_ok = sum(d["ok"] for d in data)
    log(f"dataset: {len(data)} clusters, {n_ok} labels accepted, {t_data:.1f}s")
    starter, hist, t_train = ws.train(data, epochs=a.epochs, log=log)
    if a.save_model:
        starter.save(a.save_model)
    ev = ws.evaluate(starter, a.eval_clusters, seed=999_000, max_steps=a.max_init_steps, flux_mode=a.ws_flux,
                     window=100, log=log)
    log(f"evaluation: {ev}")
    line = {
        "metric": "init seconds to steady state (residual < 1e-6), warm start vs default start",
        "value": ev["warm_seconds"],
        "unit": "s",
        "n_gpus": 1,
        "steps": int(ev["warm_max_steps"]),
        "warmup": 0,
        "ms_per_step": None,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("f64" if a.ws_flux == "fp64" else "f32-flux/f64-state") + " step; bf16-mfma/f32 network",
        "data": "synthetic",
        "config": {"workload": f"learned warm start: {len(data)} training clusters of 500-2000 cells "
                               f"({n_ok} steady-state labels accepted by the stepper), {a.eval_clusters} held-out "
                               f"clusters, Dm(Na/K/Cl) and NaK rate log-uniform over two decades, init capped at "
                               f"{a.max_init_steps} steps",
                   "train_clusters": len(data), "eval_clusters": a.eval_clusters,
                   "max_init_steps": a.max_init_steps, "label_horizon": a.label_horizon, "epochs": a.epochs},
        "cold_seconds": ev["cold_seconds"],
        "speedup": ev["cold_seconds"] / max(ev["warm_seconds"], 1e-9),
        "evaluation": ev,
        "dataset_seconds": t_data, "train_seconds": t_train,
        "train_loss_first_last": [hist[0], hist[-1]],
        "dataset_batches": stats,
    }
    print(json.dumps(line), flush=True)


def bench_full_pipeline(a):
    """The full workload as a pipeline: a warm-start network trained for its
    physics (random clusters of 500-2000 cells with the same ions, channels
    and multiplier statistics), the init phase of the 100k-cell cluster to the
    1e-6 residual from the default start and from the prediction, then the
    sim phase (channels, reaction network, mixed precision)."""
    from betse_amd import warmstart as ws
    from betse_amd.batch import BatchSimulator

    def log(msg):
        print("# " + msg, file=sys.stderr, flush=True)

    p, cells, t_world = make_problem("full")
    log(f"world: {cells.n_cells} cells, {cells.n_mems} membranes, built in {t_world:.1f}s")
    t0 = time.perf_counter()
    data, stats, _ = ws.make_dataset(a.train_clusters, seed=3000, base="full", cells_range=(500, 2000),
                                     grid_range=(36, 72), decades=0.0, batch=a.clusters, horizon=a.label_horizon,
                                     rounds=4, flux_mode="fp32", log=log)
    t_data = time.perf_counter() - t0
    # features are relative to the workload's own parameters (factors 0)
    starter, hist, t_train = ws.train(data, epochs=a.epochs, base=p, log=log)
    bs = BatchSimulator([(p, cells)], device="cuda:0", flux_mode="fp32")
    t0 = time.perf_counter()
    bs.init("init")
    t_setup = time.perf_counter() - t0
    cold = ws.run_to_residual(bs, 1e-6, 100, a.max_init_steps, log)
    warm, t_pred = ws.warm_start_batch(bs, starter, 1e-6, 100, a.max_init_steps, log)
    log(f"init: cold {cold.seconds + t_setup:.2f}s ({int(cold.steps[0])} steps), "
        f"warm {warm.seconds + t_pred:.2f}s ({int(warm.steps[0])} steps, injected {bool(warm.warm[0])})")
    # sim phase from the warm-started init state
    bs.prepare("sim")
    import torch
    bs.step_n(a.warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    bs.step_n(a.sim_steps)
    torch.cuda.synchronize()
    t_sim = time.perf_counter() - t0
    C = cells.n_cells
    line = {
        "metric": "full pipeline: init seconds to steady state (warm vs default start) and sim cell-steps/s",
        "value": warm.seconds + t_pred + t_sim, "unit": "s", "n_gpus": 1, "steps": a.sim_steps,
        "warmup": a.warmup, "ms_per_step": t_sim / a.sim_steps * 1e3, "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32-flux/f64-state", "data": "synthetic",
        "config": {"workload": f"full: {C} cells, {cells.n_mems} membranes, {cells.X.shape[0]}x{cells.X.shape[1]} "
                               f"grid, Nav1p2/Kv1p1/Cav1p2, {p.network.n_species}-substance reaction network; "
                               f"surrogate from {len(data)} clusters of 500-2000 cells, then {a.sim_steps} sim steps",
                   "cells": C},
        "init_cold_seconds": cold.seconds + t_setup, "init_cold_steps": int(cold.steps[0]),
        "init_warm_seconds": warm.seconds + t_pred, "init_warm_steps": int(warm.steps[0]),
        "warm_injected": bool(warm.warm[0]),
        "sim_seconds": t_sim, "sim_cell_steps_per_s": C * a.sim_steps / t_sim,
        "surrogate": {"train_clusters": len(data), "accepted": int(sum(d["ok"] for d in data)),
                      "dataset_seconds": t_data, "train_seconds": t_train,
                      "train_loss_first_last": [hist[0], hist[-1]]},
    }
    print(json.dumps(line), flush=True)


def make_problem(name: str):
    from betse_amd.cells import Cells
    from betse_amd.parameters import Parameters
    from betse_amd.channels import draw_channel_state
    p = Parameters.preset(name)
    t0 = time.time()
    cells = Cells.make_world(p)
    if p.channels and p.channel_random_state and p.channel_gates0 is None:
        # one seeded draw of gates and per-cell multipliers shared by the GPU runs and the oracle
        draw_channel_state(p, cells.n_cells, cells.n_mems)
    return p, cells, time.time() - t0


def gpu_session(p, cells, flux_mode):
    from betse_amd.sim import SimPhase, SimPhaseKind, Simulator
    sim = Simulator(p.copy(), device="cuda:0", flux_mode=flux_mode, use_graph=True)
    ph = SimPhase(SimPhaseKind.INIT, sim.p, cells, sim)
    sim.init_core(ph)
    sim.prepare(ph)
    return sim


def time_gpu(p, cells, flux_mode, steps, warmup):
    import torch
    sim = gpu_session(p, cells, flux_mode)
    sim.step_n(warmup)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(sim.stream)
    sim.step_n(steps, sync=False)
    e1.record(sim.stream)
    sim.pull()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ev_ms = e0.elapsed_time(e1)
    names = ("mem_cell", "env_field", "env_exchange") + (("network",) if sim.network is not None else ())
    kernels = {k: sim.time_kernel(k, 100) for k in names}
    return sim, wall, ev_ms, kernels


PROFILES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles")


def measured_traffic(workload: str, kernel: str, flux_mode: str):
    """HBM bytes per launch of a step kernel from the committed rocprofv3 PMC
    summary (profiles/hbm_traffic_<workload>.txt: FETCH_SIZE and WRITE_SIZE per
    dispatch, already scaled by the plain-copy calibration), or None."""
    path = os.path.join(PROFILES, f"hbm_traffic_{workload}.txt")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            t = line.split()
            if len(t) >= 6 and t[0] == "kernel" and t[1] == kernel and t[3] == flux_mode:
                return float(t[5])
    return None


def time_sampled_loop(p, cells, flux_mode, steps):
    """The init phase's step loop (_run_sim_core_loop) with the preset's
    sampling: sampled steps run bt_step_observed, derive the get_current
    observables (HH decomposition by DST on FP64 MFMA) and go through
    write2storage on the host. Set-up (run_sim_core's init_dynamics, upload,
    first update_V, observables workspace) is outside the timed loop."""
    import torch
    from betse_amd.sim import SimPhase, SimPhaseKind, Simulator, phase_time_steps
    q = p.copy()
    q.init_time_total = steps * q.init_time_step
    sim = Simulator(q, device="cuda:0", flux_mode=flux_mode, use_graph=True)
    ph = SimPhase(SimPhaseKind.INIT, sim.p, cells, sim)
    sim.init_core(ph)
    sim.prepare(ph)
    sim.step_observed(1)                 # allocates the observables workspace
    # an untimed pass over two sampled steps allocates the pinned and snapshot
    # buffers of the sampled-step copies
    warm, warm_sampled = phase_time_steps(int(3 * q.t_resample) + 1, q.dt, q.t_resample)
    sim._run_sim_core_loop(ph, warm, warm_sampled)
    time_steps, sampled = phase_time_steps(q.init_tsteps, q.dt, q.t_resample)
    walls = []
    for _ in range(3):                   # median of three passes: the host side varies
        sim.clear_storage()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sim._run_sim_core_loop(ph, time_steps, sampled)
        torch.cuda.synchronize()
        walls.append(time.perf_counter() - t0)
    return float(np.median(walls)), len(sim.time)


def cpu_baseline(p, cells, target_s: float):
    """FP64 oracle (sparse operators) on this host: a bounded prefix of the run."""
    from oracle.betse_oracle import OracleSim
    o = OracleSim(cells, p.copy(), mode="sparse")
    o.init_core()
    o.init_dynamics("init")
    o.p.set_phase_time("init")
    o.update_V(observables=False)
    t0 = time.perf_counter()
    o.step("init", observables=False)
    one = time.perf_counter() - t0
    n = max(2, int(target_s / max(one, 1e-6)))
    t0 = time.perf_counter()
    for _ in range(n - 1):
        o.step("init", observables=False)
    el = one + (time.perf_counter() - t0)
    return o, n, el


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--config", default="large")
    ap.add_argument("--flux", default="both", choices=["fp64", "fp32", "both"])
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--clusters", type=int, default=256, help="batched config: number of clusters")
    ap.add_argument("--cpu-clusters", type=int, default=8, help="batched config: clusters in the CPU sample")
    ap.add_argument("--train-clusters", type=int, default=4096, help="warmstart config: training clusters")
    ap.add_argument("--eval-clusters", type=int, default=256, help="warmstart config: held-out clusters")
    ap.add_argument("--max-init-steps", type=int, default=100_000, help="warmstart config: init step cap")
    ap.add_argument("--label-horizon", type=int, default=20_000,
                    help="warmstart config: stepper steps before the first extrapolation of a label")
    ap.add_argument("--epochs", type=int, default=40, help="warmstart config: training epochs")
    ap.add_argument("--save-model", default="", help="warmstart config: save the trained network (.npz)")
    ap.add_argument("--sim-steps", type=int, default=20_000, help="full_pipeline config: sim-phase steps")
    ap.add_argument("--ws-flux", default="fp32", choices=["fp64", "fp32"],
                    help="warmstart config: flux arithmetic of the init runs")
    ap.add_argument("--copy-calib-mib", type=int, default=0,
                    help="only run a plain device copy of this many MiB (HBM counter calibration)")
    a = ap.parse_args(argv)
    if a.copy_calib_mib:
        import torch
        n = a.copy_calib_mib * (1 << 20) // 8
        x = torch.ones(n, dtype=torch.float64, device="cuda:0")
        y = torch.empty_like(x)
        for _ in range(5):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calib_bytes_read": n * 8, "copy_calib_bytes_written": n * 8, "launches": 5}))
        return
    if a.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {a.gpus} is not supported "
              "(each GPU box has one GPU; the sharding design is in DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    if a.config == "batched":
        return bench_batched(a)
    if a.config == "warmstart":
        return bench_warmstart(a)
    if a.config == "full_pipeline":
        return bench_full_pipeline(a)
    p, cells, t_world = make_problem(a.config)
    n_ions = len(p.ion_names())
    has_ca = "Ca" in p.ion_names()
    C = cells.n_cells
    print(f"# world: {C} cells, {cells.n_mems} membranes, grid {cells.X.shape}, ions {p.ion_names()}, "
          f"built in {t_world:.1f}s", file=sys.stderr, flush=True)

    modes = ["fp64", "fp32"] if a.flux == "both" else [a.flux]
    res = {}
    for mode in modes:
        sim, wall, ev_ms, kern = time_gpu(p, cells, mode, a.steps, a.warmup)
        res[mode] = dict(sim=sim, wall=wall, ev_ms=ev_ms, kernels=kern)
        print(f"# {mode}: {a.steps} steps in {wall:.3f}s wall ({ev_ms/a.steps*1e3:.1f} us/step by events), "
              f"kernels(us): " + ", ".join(f"{k}={v*1e3:.2f}" for k, v in kern.items()),
              file=sys.stderr, flush=True)

    # the same steps through run_sim_core with the preset's sampling (observables + storage)
    sampled = {}
    for mode in modes:
        wall_s, n_samp = time_sampled_loop(p, cells, mode, a.steps)
        sampled[mode] = {"cell_steps_per_s": C * a.steps / wall_s, "wall_s": wall_s, "sampled_steps": n_samp}
        print(f"# {mode} run_sim_core with {n_samp} sampled steps: {wall_s / a.steps * 1e6:.1f} us/step",
              file=sys.stderr, flush=True)

    # CPU baseline + parity on the same prefix
    cpu = None
    parity = {}
    if not a.no_cpu:
        o, n_cpu, el = cpu_baseline(p, cells, a.cpu_seconds)
        cpu = {"value": C * n_cpu / el, "unit": "cell-steps/s", "cores": 1, "kind": "port",
               "sample": f"FP64 NumPy oracle (sparse operators), first {n_cpu} steps of the same "
                         f"workload, {el:.1f}s on this host, single thread"}
        for mode in modes:
            sim = gpu_session(p, cells, mode)
            sim.step_n(n_cpu)
            sim.pull()
            parity[mode] = {
                "steps": n_cpu,
                "max_abs_dVm_V": float(np.abs(sim.vm - o.vm).max()),
                "max_rel_dc_cell": float((np.abs(sim.cc_cells - o.cc_cells) / np.abs(o.cc_cells)).max()),
                "max_rel_dc_env": float((np.abs(sim.cc_env - o.cc_env) / np.abs(o.cc_env)).max()),
                "per_ion_rel_dc_cell": {n: float((np.abs(sim.cc_cells[i] - o.cc_cells[i]) / np.abs(o.cc_cells[i])).max())
                                        for i, n in enumerate(sim.names)},
                "per_ion_rel_dc_env": {n: float((np.abs(sim.cc_env[i] - o.cc_env[i]) / np.abs(o.cc_env[i])).max())
                                       for i, n in enumerate(sim.names)},
                "max_abs_dE_env_V_per_m": float(max(np.abs(sim.E_env_x - o.E_env_x).max(),
                                                    np.abs(sim.E_env_y - o.E_env_y).max())),
            }
            if len(p.channels):
                parity[mode]["max_abs_dgate"] = {ch.kind: float(max(np.abs(sim.chan_gates[k, 0] - ch.m).max(),
                                                                    np.abs(sim.chan_gates[k, 1] - ch.h).max()))
                                                 for k, ch in enumerate(o.channels)}
            if getattr(p, "network", None) is not None:
                parity[mode]["max_rel_dc_substance"] = float(
                    (np.abs(sim.c_net - o.c_net) / np.maximum(np.abs(o.c_net), 1e-300)).max())
        print(f"# cpu oracle: {n_cpu} steps in {el:.2f}s; parity {parity}", file=sys.stderr, flush=True)

    # headline: fastest path that holds |dVm| < 1 uV against the oracle prefix
    ok = [m for m in modes if (not parity) or parity[m]["max_abs_dVm_V"] < 1e-6]
    best = min(ok or modes, key=lambda m: res[m]["ev_ms"])
    r = res[best]
    ms_step = r["ev_ms"] / a.steps
    value = C * a.steps / (r["ev_ms"] / 1e3)
    n_sub = p.network.n_species if getattr(p, "network", None) is not None else 0
    Dm, gjo = r["sim"].Dm_cells, r["sim"].gjopen
    De = r["sim"].D_env * r["sim"].TJ_modulator
    dm_loaded = int(sum(not np.all(Dm[i] == Dm[i][0]) for i in range(n_ions)))
    de_loaded = int(sum(not np.all(De[i] == De[i][0]) for i in range(n_ions)))
    ab = algorithmic_bytes(cells, n_ions, has_ca, best, len(p.channels), n_sub, dm_loaded,
                           not bool(np.all(gjo == gjo[0])), de_loaded)
    kern = r["kernels"]
    dom = max(kern, key=kern.get)
    achieved = ab[dom] / (kern[dom] * 1e-3) / 1e9
    line = {
        "metric": "cell-steps/s at |dVmem|<1uV vs FP64 oracle",
        "value": value,
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if best == "fp64" else "f32-flux/f64-state",
        "data": "synthetic",
        "config": {"workload": f"single large cluster ({a.config}): {C} cells, {cells.n_mems} membranes, "
                               f"{cells.X.shape[0]}x{cells.X.shape[1]} grid, ions {'/'.join(p.ion_names())}, "
                               f"dt {p.dt:g} s; channels {'/'.join(c.kind for c in p.channels) or 'none'}",
                   "cells": C, "membranes": cells.n_mems, "grid": list(cells.X.shape), "n_ions": n_ions,
                   "flux_mode": best},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(a.config, dom, best),
                     "launch_us": kern[dom] * 1e3, "algorithmic_bytes": ab[dom],
                     "step_algorithmic_bytes": ab["step"],
                     "step_achieved_GBs": ab["step"] / (ms_step * 1e-3) / 1e9},
        "cpu_baseline": cpu,
        "parity_prefix": parity,
        "with_sampling": sampled,
        "paths": {m: {"cell_steps_per_s": C * a.steps / (res[m]["ev_ms"] / 1e3),
                      "us_per_step": res[m]["ev_ms"] / a.steps * 1e3,
                      "kernels_us": {k: v * 1e3 for k, v in res[m]["kernels"].items()}} for m in modes},
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
