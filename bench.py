"""Benchmark of the electrodiffusion step on one MI355X.

Default workload: the batched configuration -- 256 independent clusters of
100-5000 cells (jittered hex lattices of 10 um spacing, one numpy seed per
cluster; ~500k cells, ~3.0M membranes), environment grids 32x32 to 128x128
scaled to each cluster's radius, ions Na/K/Cl/P/M, per-cluster Dm(Na/K/Cl)
and NaK-ATPase rate log-uniform over one decade, dt 1e-3 s. It is the largest
configuration that fits one GPU. A step is one iteration of
``_run_sim_core_loop`` over every cluster; the timed region is ``--steps``
steps of that loop (``BatchSimulator.run_loop``: the sampled steps of a phase
of that length, by the reference's own rule of one every t_resample = 100
steps after the first 100, run the observables and ``write2storage``) with
the state resident in HBM. The default, 2,000 steps, is the workload's own
horizon with its 19 sampled steps. The headline path is FP64 throughout; the FP32-flux path
(FP32 flux laws, FP64 state, charge and voltage) is reported beside it.

Printed: one JSON line with cell-steps/s, the roofline of the step (the
compulsory bytes of one ideal fused pass per step, per membrane, cell and
env square as ``b_step`` states them, over the measured step time),
the measured HBM traffic per step from the committed rocprofv3 summaries, the
launch variants the autotuner chose, and the CPU baseline: the FP64 NumPy
oracle on the host's cores (clusters in parallel processes) and on one
thread, plus the reference's own dense-operator form on the small cluster.
"""

from __future__ import annotations

import argparse
import gc
import json
import multiprocessing
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
PROFILES = os.path.join(ROOT, "profiles")
MAX_CPU_WORKERS = 16          # the GPU box's CPU share for one GPU


def log(msg: str) -> None:
    print("# " + msg, file=sys.stderr, flush=True)


# ---------------------------------------------------------------------------
# algorithmic bytes of one step (one ideal fused pass)
# ---------------------------------------------------------------------------
def b_step(n_cells: int, n_mems: int, n_squares: int, n_ions: int, n_gates: int = 0, n_mult: int = 0,
           n_species: int = 0, p: int = 8) -> dict:
    """Compulsory HBM bytes of one step: every array element once per step at
    its storage width, reads and writes separately, gathers of counted arrays
    and the sampled-only observables excluded. Per membrane: two int32 indices
    (nn_i, map_mem2ecm), mem_sa, the GJ conductance and Dm per ion, each gate
    read and written; per cell: c_cell and c_lag read and written per ion,
    Vmem read and written, cell_vol and diviterm, the CSR row pointer, the
    channel multipliers and each substance read and written; per square:
    c_env read and written per ion, D_env per ion, E_env (x, y) read and
    written, v_env read and written, the charge scale and mask. ``p`` is the
    storage width of flux parameters, gates and E (8 here: this build stores
    them as FP64 on both flux paths)."""
    b_M = 8 + p * (2 + n_ions) + 2 * p * n_gates
    b_C = 2 * 2 * 8 * n_ions + 16 + 16 + 4 + p * n_mult + 32 * n_species
    b_G = 2 * 8 * n_ions + p * n_ions + 2 * 2 * p + 16 + 8
    total = n_mems * b_M + n_cells * b_C + n_squares * b_G
    return {"b_M": b_M, "b_C": b_C, "b_G": b_G, "bytes": total, "per_cell": total / max(n_cells, 1)}


def workload_b_step(worlds) -> dict:
    p0, _ = worlds[0]
    n_ions = len(p0.ion_names())
    n_ch = len(p0.channels)
    net = getattr(p0, "network", None)
    return b_step(sum(c.n_cells for _, c in worlds), sum(c.n_mems for _, c in worlds),
                  sum(c.X.size for _, c in worlds), n_ions, n_gates=2 * n_ch, n_mult=n_ch,
                  n_species=net.n_species if net is not None else 0)


def step_kernel_names(tuning: dict, network: bool):
    """The kernels of one step under a tuning: the env kernel does the exchange
    with exchange_form 2 (k_mem_cell runs first), else a kernel of its own."""
    names = ("mem_cell", "env_field_exchange") if tuning.get("exchange_form") == 2 else \
        ("mem_cell", "env_field", "env_exchange")
    return names + (("network",) if network else ())


def measured_traffic(workload: str, flux_mode: str, kernels=("mem_cell", "env_field", "env_exchange")):
    """HBM bytes per step of the step kernels from the committed rocprofv3 PMC
    summary (profiles/hbm_traffic_<workload>.txt: FETCH_SIZE and WRITE_SIZE per
    dispatch, scaled by the plain-copy calibration), or None when a kernel of
    the step has no figure for this flux path."""
    path = os.path.join(PROFILES, f"hbm_traffic_{workload}.txt")
    if not os.path.exists(path):
        return None
    per = {}
    with open(path) as f:
        for line in f:
            t = line.split()
            if len(t) >= 6 and t[0] == "kernel" and t[3] == flux_mode:
                per[t[1]] = float(t[5])
    need = set(kernels) - {"network"}
    if not need <= set(per):
        return None
    per = {k: v for k, v in per.items() if k in set(kernels)}
    return {"bytes": sum(per.values()), "kernels": per}


# ---------------------------------------------------------------------------
# workloads
# ---------------------------------------------------------------------------
def make_worlds(config: str, clusters: int):
    """[(Parameters, Cells)] of a configuration and its description."""
    from betse_amd.batch import batched_worlds
    t0 = time.time()
    if config == "batched":
        worlds = batched_worlds(clusters)
        p0 = worlds[0][0]
        desc = (f"batched: {len(worlds)} independent clusters of 100-5000 cells on 10 um jittered hex lattices, "
                f"grids 32-128 scaled to radius, ions {'/'.join(p0.ion_names())}, per-cluster Dm(Na/K/Cl) and "
                f"NaK rate log-uniform over one decade, dt {p0.dt:g} s")
    else:
        p, cells, _ = make_problem(config)
        worlds = [(p, cells)]
        desc = (f"single cluster ({config}): {cells.n_cells} cells, {cells.n_mems} membranes, "
                f"{cells.X.shape[0]}x{cells.X.shape[1]} grid, ions {'/'.join(p.ion_names())}, dt {p.dt:g} s; "
                f"channels {'/'.join(c.kind for c in p.channels) or 'none'}"
                + ("; 8-substance reaction network" if getattr(p, "network", None) is not None else ""))
    log(f"{desc}: {sum(c.n_cells for _, c in worlds)} cells, {sum(c.n_mems for _, c in worlds)} membranes, "
        f"{sum(c.X.size for _, c in worlds)} squares, built in {time.time() - t0:.1f}s")
    return worlds, desc


# ---------------------------------------------------------------------------
# CPU baseline: the FP64 oracle on the host (run before the GPU is touched:
# the workers are forked)
# ---------------------------------------------------------------------------
def _cpu_job(args):
    """Oracle steps of the given clusters in one process, one BLAS thread."""
    from threadpoolctl import threadpool_limits
    from oracle.betse_oracle import OracleSim
    worlds, n_steps, mode = args
    with threadpool_limits(1):
        t0 = time.perf_counter()
        out = []
        for p, cells in worlds:
            o = OracleSim(cells, p.copy(), mode=mode)
            o.init_core()
            o.init_dynamics("init")
            o.p.set_phase_time("init")
            o.update_V(observables=(mode == "dense"))
            t1 = time.perf_counter()
            for _ in range(n_steps):
                o.step("init", observables=(mode == "dense"))
            out.append({"cells": cells.n_cells, "seconds": time.perf_counter() - t1, "vm": o.vm.copy(),
                        "cc_cells": o.cc_cells.copy(), "cc_env": o.cc_env.copy()})
        return out


def host_info() -> dict:
    model = platform.processor()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    blas = None
    try:
        from threadpoolctl import threadpool_info
        info = [i for i in threadpool_info() if i.get("user_api") == "blas"]
        if info:
            blas = f"{info[0].get('internal_api')} {info[0].get('version')} ({info[0].get('num_threads')} threads)"
    except Exception:
        pass
    return {"cpu_model": model, "cores_available": len(os.sched_getaffinity(0)), "blas": blas}


def cpu_baseline(worlds, target_s: float, sample_clusters: int):
    """The sparse FP64 oracle on a sample of the workload's clusters: on all
    allotted cores (one process per core, clusters dealt out largest first)
    and on one thread. Returns (baseline dict, per-cluster final states, the
    sample's indices, steps)."""
    n_all = len(worlds)
    order = np.argsort([-c.n_cells for _, c in worlds])
    if n_all > sample_clusters:
        # every k-th cluster by size: the sample spans the size range
        pick = sorted(int(order[i]) for i in np.linspace(0, n_all - 1, sample_clusters).astype(int))
    else:
        pick = list(range(n_all))
    sample = [worlds[i] for i in pick]
    cores = max(1, min(len(os.sched_getaffinity(0)), MAX_CPU_WORKERS, len(sample)))
    # size the run: one step of the whole sample on one thread
    t_one = _cpu_job((sample, 1, "sparse"))
    one = sum(r["seconds"] for r in t_one)
    n_steps = max(2, int(target_s * min(cores, len(sample)) / max(one, 1e-6) / 2))
    # all cores: each process gets whole clusters (LPT on cells)
    loads = [[] for _ in range(cores)]
    tot = np.zeros(cores)
    for i in sorted(range(len(sample)), key=lambda j: -sample[j][1].n_cells):
        k = int(np.argmin(tot))
        loads[k].append(i)
        tot[k] += sample[i][1].n_cells
    ctx = multiprocessing.get_context("fork")
    t0 = time.perf_counter()
    with ctx.Pool(cores) as pool:
        res = pool.map(_cpu_job, [([sample[i] for i in ld], n_steps, "sparse") for ld in loads])
    wall = time.perf_counter() - t0
    states = {}
    for ld, r in zip(loads, res):
        for i, st in zip(ld, r):
            states[pick[i]] = st
    cells_total = sum(c.n_cells for _, c in sample)
    par = cells_total * n_steps / wall
    # one thread: the same sample for fewer steps
    n1 = max(2, int(target_s / 2 / max(one, 1e-6)))
    r1 = _cpu_job((sample, n1, "sparse"))
    single = cells_total * n1 / sum(r["seconds"] for r in r1)
    info = host_info()
    out = {
        "value": par, "unit": "cell-steps/s", "cores": cores, "kind": "port",
        "sample": (f"FP64 NumPy oracle (sparse operators: segment sums, scatter, sparse Poisson solve), "
                   f"{len(sample)} of the workload's {n_all} clusters ({cells_total} cells) for {n_steps} steps, "
                   f"one process per core with one BLAS thread, {wall:.1f}s wall"),
        "one_thread": {"value": single, "steps": n1},
        **info,
    }
    return out, states, pick, n_steps


def cpu_reference_form(target_s: float) -> dict:
    """The reference's own operator form on the small cluster (362 cells, 50x50
    grid): dense M_sum_mems, M_divmap_mem2ecm and the pinv Laplacian, every
    observable each step, as sim.py runs it; BLAS on all allotted cores and on
    one thread."""
    from threadpoolctl import threadpool_limits
    from betse_amd.cells import Cells
    from betse_amd.parameters import Parameters
    from oracle.betse_oracle import OracleSim
    p = Parameters.preset("small")
    cells = Cells.load(os.path.join(ROOT, "tests", "golden", "world_small.npz"))
    cores = max(1, min(len(os.sched_getaffinity(0)), MAX_CPU_WORKERS))
    out = {"workload": f"small cluster, {cells.n_cells} cells, {cells.n_mems} membranes, 50x50 grid, dt 1e-3 s",
           "operators": "dense (reference form)"}
    for tag, thr in (("all_cores", cores), ("one_thread", 1)):
        with threadpool_limits(thr):
            o = OracleSim(cells, p.copy(), mode="dense")
            o.init_core()
            o.init_dynamics("init")
            o.p.set_phase_time("init")
            o.update_V(observables=True)
            t0 = time.perf_counter()
            o.step("init", observables=True)
            one = time.perf_counter() - t0
            n = max(2, int(target_s / max(one, 1e-6)))
            t0 = time.perf_counter()
            for _ in range(n):
                o.step("init", observables=True)
            el = time.perf_counter() - t0
        out[tag] = {"threads": thr, "steps": n, "ms_per_step": el / n * 1e3,
                    "cell_steps_per_s": cells.n_cells * n / el}
    return out


# ---------------------------------------------------------------------------
# GPU
# ---------------------------------------------------------------------------
TUNING = None          # launch variants forced by --tuning
AUTOTUNE = True        # else the autotuner's choice (or, with --no-autotune, bt_create's by size)


def gpu_batch(worlds, mode, autotune=True):
    from betse_amd.batch import BatchSimulator
    b = BatchSimulator([(p.copy(), c) for p, c in worlds], device="cuda:0", flux_mode=mode,
                       autotune=autotune and AUTOTUNE and TUNING is None, tuning=TUNING)
    b.init("init")
    return b


def time_loop(worlds, mode, steps, warmup):
    """K steps of the batch's run_loop after W warm-up steps (the graphs are
    captured at set-up), bracketed by device synchronisations; HIP events on
    the simulation stream give the device time. The step alone (480 replayed
    steps on HIP events, about 0.1 s: the roofline's denominator) is measured
    before the warm-up steps: the set-up's host work (pinned snapshot buffers)
    leaves the GPU idle long enough for its clocks to drop, and its first
    milliseconds of steps then run up to 10% slower, which a 20-step timed
    region would otherwise take in. Each step kernel repeated alone
    (``bt_time_kernel``) is timed after the timed steps."""
    import torch
    b = gpu_batch(worlds, mode)
    b.prepare_sampling()              # set-up of the sampled steps, outside the timed region
    # the step alone (replayed graphs, no sampled step), for the roofline
    n_pure = 480
    p0 = torch.cuda.Event(enable_timing=True)
    p1 = torch.cuda.Event(enable_timing=True)
    p0.record(b.stream)
    b.step_n(n_pure, sync=False)
    p1.record(b.stream)
    b.sync()
    pure_us = p0.elapsed_time(p1) / n_pure * 1e3
    b.step_n(warmup)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(b.stream)
    n_samp = b.run_loop(steps)
    e1.record(b.stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ev_ms = e0.elapsed_time(e1)
    names = step_kernel_names(b.get_tuning(), "c_net" in b._dev)
    kernels = {k: b.time_kernel(k, 20) * 1e3 for k in names}
    res = {"wall_s": wall, "events_ms": ev_ms, "sampled_steps": n_samp, "kernels_us": kernels,
           "step_us": pure_us, "tuning": b.get_tuning(), "autotune": b.autotune_report}
    b.close()
    del b
    gc.collect()
    return res


def time_sampled(worlds, mode, steps):
    """The workload's own horizon through run_loop with its sampling (every
    t_resample steps: observables of get_current, HH decomposition on FP64
    MFMA, write2storage per cluster on the host)."""
    import torch
    b = gpu_batch(worlds, mode)
    b.prepare_sampling()
    b.step_n(2)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    n_samp = b.run_loop(steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    b.close()
    del b
    gc.collect()
    return {"steps": steps, "sampled_steps": n_samp, "wall_s": wall}


def gpu_parity(worlds, pick, states, n_steps, mode):
    b = gpu_batch([worlds[i] for i in pick], mode, autotune=False)
    b.step_n(n_steps)
    b.pull()
    dv, dc, de = 0.0, 0.0, 0.0
    for st, i in zip(b.states, pick):
        o = states[i]
        dv = max(dv, float(np.abs(st.vm - o["vm"]).max()))
        dc = max(dc, float((np.abs(st.cc_cells - o["cc_cells"]) / np.abs(o["cc_cells"]).clip(1e-300)).max()))
        de = max(de, float((np.abs(st.cc_env - o["cc_env"]) / np.abs(o["cc_env"]).clip(1e-300)).max()))
    b.close()
    del b
    gc.collect()
    return {"steps": n_steps, "clusters": len(pick), "max_abs_dVm_V": dv, "max_rel_dc_cell": dc,
            "max_rel_dc_env": de}


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


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--config", default="batched",
                    choices=["batched", "large", "full", "small", "warmstart", "full_pipeline"])
    ap.add_argument("--flux", default="both", choices=["fp64", "fp32", "both"])
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-sampled", action="store_true", help="skip the run over the workload's own horizon")
    ap.add_argument("--horizon", type=int, default=2000, help="steps of the sampled run (the workload's own)")
    ap.add_argument("--clusters", type=int, default=256, help="batched config: number of clusters")
    ap.add_argument("--cpu-clusters", type=int, default=32, help="batched config: clusters in the CPU sample")
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
    ap.add_argument("--tuning", default="", help="force the launch variants (JSON object of bt_tuning fields) "
                                                 "instead of the autotuner's choice")
    ap.add_argument("--no-autotune", action="store_true", help="keep the variants bt_create picks by size")
    ap.add_argument("--copy-calib-mib", type=int, default=0,
                    help="only run a plain device copy of this many MiB (HBM counter calibration)")
    a = ap.parse_args(argv)
    if a.gpus != 1:
        print(f"bench.py: one GPU only; --gpus {a.gpus} is not supported (the step runs on the one GPU; "
              "several GPUs would each take a strip of the grid with a halo exchange per step, not built)",
              file=sys.stderr)
        sys.exit(2)
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
    global TUNING, AUTOTUNE
    if a.tuning:
        TUNING = json.loads(a.tuning)
    AUTOTUNE = not a.no_autotune
    if a.config == "warmstart":
        return bench_warmstart(a)
    if a.config == "full_pipeline":
        return bench_full_pipeline(a)

    worlds, desc = make_worlds(a.config, a.clusters)
    C = sum(c.n_cells for _, c in worlds)
    # CPU first: its worker processes are forked before anything touches the GPU
    cpu, states, pick, n_cpu, ref_form = None, {}, [], 0, None
    if not a.no_cpu:
        cpu, states, pick, n_cpu = cpu_baseline(worlds, a.cpu_seconds, a.cpu_clusters)
        log(f"cpu oracle: {cpu['value']:.3g} cell-steps/s on {cpu['cores']} cores "
            f"({cpu['one_thread']['value']:.3g} on one thread); {cpu['sample']}")
        ref_form = cpu_reference_form(a.cpu_seconds / 3)
        log(f"reference form (dense operators), small cluster: {ref_form}")

    modes = ["fp64", "fp32"] if a.flux == "both" else [a.flux]
    res = {}
    for mode in modes:
        res[mode] = time_loop(worlds, mode, a.steps, a.warmup)
        r = res[mode]
        log(f"{mode}: {a.steps} steps in {r['wall_s'] * 1e3:.2f} ms wall ({r['events_ms'] / a.steps * 1e3:.1f} us/step "
            f"by events, {r['sampled_steps']} sampled; step alone {r['step_us']:.1f} us), kernels(us): "
            + ", ".join(f"{k}={v:.1f}" for k, v in r["kernels_us"].items()) + f"; variants {r['tuning']}")
    sampled = {}
    if not a.no_sampled:
        mode = modes[0]
        if a.steps == a.horizon:            # the timed run is the horizon itself
            s = {"steps": a.steps, "sampled_steps": res[mode]["sampled_steps"], "wall_s": res[mode]["wall_s"]}
        else:
            s = time_sampled(worlds, mode, a.horizon)
        s["cell_steps_per_s"] = C * s["steps"] / s["wall_s"]
        sampled[mode] = s
        log(f"{mode} over the workload's {a.horizon}-step horizon with {s['sampled_steps']} sampled steps: "
            f"{s['wall_s'] / a.horizon * 1e6:.1f} us/step")
    parity = {}
    if states:
        for mode in modes:
            parity[mode] = gpu_parity(worlds, pick, states, n_cpu, mode)
        log(f"parity against the cpu sample: {parity}")

    head = "fp64" if "fp64" in modes else modes[0]
    r = res[head]
    ms_step = r["wall_s"] * 1e3 / a.steps
    value = C * a.steps / r["wall_s"]
    bs = workload_b_step(worlds)
    ev_step_us = r["step_us"]
    achieved = bs["bytes"] / (ev_step_us * 1e-6) / 1e9
    traffic = measured_traffic(a.config, head, tuple(r["kernels_us"]))
    kern = r["kernels_us"]
    dom = max(kern, key=kern.get)
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
        "dtype": "f64" if head == "fp64" else "f32-flux/f64-state",
        "data": "synthetic",
        "config": {"workload": desc, "clusters": len(worlds), "cells": C,
                   "membranes": int(sum(c.n_mems for _, c in worlds)),
                   "squares": int(sum(c.X.size for _, c in worlds)), "flux_mode": head},
        "roofline": {"bound": "hbm", "kernel": "step (" + " + ".join(f"k_{k}" for k in kern) + ", graph-replayed)",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic["bytes"] if traffic else None,
                     "algorithmic_bytes": bs["bytes"], "bytes_per_unit": {k: bs[k] for k in ("b_M", "b_C", "b_G")},
                     "step_us": ev_step_us, "kernels_us": kern,
                     "dominant_kernel": {"name": dom, "launch_us": kern[dom]},
                     "traffic_per_kernel": traffic["kernels"] if traffic else None},
        "cpu_baseline": cpu,
        "cpu_reference_form": ref_form,
        "parity_prefix": parity,
        "with_sampling": sampled,
        "paths": {m: {"cell_steps_per_s": C * a.steps / res[m]["wall_s"],
                      "us_per_step": res[m]["wall_s"] / a.steps * 1e6,
                      "events_us_per_step": res[m]["events_ms"] / a.steps * 1e3,
                      "step_us": res[m]["step_us"],
                      "sampled_steps": res[m]["sampled_steps"],
                      "kernels_us": res[m]["kernels_us"], "tuning": res[m]["tuning"],
                      "autotune": res[m]["autotune"]} for m in modes},
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
