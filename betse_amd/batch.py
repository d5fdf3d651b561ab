"""Many independent clusters stepped together on one GPU.

Each cluster is a BETSE simulation of its own (its own cell cluster,
environment grid, membrane diffusivities and pump rate). The host sets every
cluster up exactly as ``Simulator`` does (``HostState.init_core`` /
``init_dynamics``, i.e. ``sim.py:452-1012``), then concatenates cells,
membranes and grid squares into one batch so that each step of the reference
loop (``sim.py:1169-1366``) is the same three kernels over all clusters: the
kernels read each cluster's grid and constants from a per-cluster table.

Clusters must share the ion profile, the channel models and the time step;
everything else may differ.
"""

from __future__ import annotations

import ctypes
from collections import deque
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import _native
from .boundary import SIDES, boundary_potential
from .cells import Cells
from .channels import draw_channel_state
from .exceptions import BetseNativeError, BetseSimException, BetseSimUnstableException
from .parameters import Parameters
from .assemble import (BUFFER_FIELDS, batch_geometry, batch_state, check_compatible, make_clusters,
                       make_params)
from .sim import DeviceFetch, HostState, SimPhase, SimPhaseKind, phase_time_steps


class BatchSimulator(DeviceFetch):
    """A batch of independent clusters on one device."""

    def __init__(self, worlds: Sequence[Tuple[Parameters, Cells]], *, device: Optional[str] = None,
                 flux_mode: str = "fp32", use_graph: bool = True, autotune: bool = False,
                 tuning: Optional[dict] = None):
        import torch
        if not torch.cuda.is_available():
            raise BetseSimException("no GPU visible: the GPU simulator has no CPU fallback")
        if flux_mode not in ("fp64", "fp32"):
            raise ValueError(flux_mode)
        self.torch = torch
        self.device = torch.device(device or "cuda")
        self.stream = torch.cuda.Stream(device=self.device)
        self.flux_mode = flux_mode
        self.use_graph = use_graph
        self._lib = _native.load()
        self._ctx = ctypes.c_void_p()
        self.worlds = [(p, c) for p, c in worlds]
        self.states: List[HostState] = []
        self._dev = {}
        # launch variants: forced (tuning), measured on the batch (autotune), or
        # bt_create's choice by problem size
        self.autotune = autotune
        self.tuning = tuning
        self.autotune_report = None

    # ------------------------------------------------------------------
    def init(self, kind: str = "init") -> None:
        """``init_core`` + ``init_dynamics`` for every cluster, then the device batch."""
        k = SimPhaseKind.INIT if kind == "init" else SimPhaseKind.SIM
        self.kind = k
        self.states = []
        for p, cells in self.worlds:
            st = HostState()
            ph = SimPhase(k, p, cells)
            st.init_core(ph)
            st.init_dynamics(ph)
            self.states.append(st)
        self._check_compatible()
        self._build()

    def prepare(self, kind: Optional[str] = None) -> None:
        """Re-run ``init_dynamics`` (a new phase) on the current states and rebuild."""
        k = self.kind if kind is None else (SimPhaseKind.INIT if kind == "init" else SimPhaseKind.SIM)
        self.kind = k
        for st, (p, cells) in zip(self.states, self.worlds):
            st.init_dynamics(SimPhase(k, p, cells))
        self._build()

    def subset(self, keep: Sequence[int]) -> "BatchSimulator":
        """A new batch of the clusters ``keep`` that continues from their current
        state (pulled from this batch); this batch is left as it is."""
        self.pull()
        sub = BatchSimulator([self.worlds[i] for i in keep], device=str(self.device), flux_mode=self.flux_mode,
                             use_graph=self.use_graph, autotune=self.autotune, tuning=self.tuning)
        sub.kind = self.kind
        sub.states = [self.states[i] for i in keep]
        sub._check_compatible()
        sub._build()
        return sub

    def rebuild(self) -> None:
        """Upload the host states again (after they were changed on the host)."""
        self._check_compatible()
        self._build()

    def _check_compatible(self) -> None:
        check_compatible(self.states, [p for p, _ in self.worlds], [c for _, c in self.worlds])

    # ------------------------------------------------------------------
    def _build(self) -> None:
        torch = self.torch
        self._destroy()
        sts = self.states
        params = [p for p, _ in self.worlds]
        cells_l = [c for _, c in self.worlds]
        geo, self.offsets = batch_geometry(cells_l)
        soff = self.offsets[2]
        self.n_tiles = len(geo["tile_ptr"]) - 1
        host = dict(geo)
        host.update(batch_state(sts, cells_l, self.kind is SimPhaseKind.SIM))
        net = sts[0].network
        self._dev = {}
        d = self._dev
        with torch.cuda.device(self.device), torch.cuda.stream(self.stream):
            for k, a in host.items():
                d[k] = torch.from_numpy(a).to(self.device)
            self.stream.synchronize()

        p0 = params[0]
        active = self.kind is SimPhaseKind.SIM or getattr(p0, "channels_init_active", True)
        prm = make_params(sts, params, cells_l, self.flux_mode, active)
        cls = make_clusters(sts, params, cells_l, soff)
        buf = _native.BtBuffers()
        for name in BUFFER_FIELDS:
            setattr(buf, name, d[name].data_ptr())
        buf.n_tiles = self.n_tiles
        buf.tile_m = _native.BT_TILE_MEMBRANES
        self._keep = (prm, cls, buf)
        with torch.cuda.device(self.device):
            rc = self._lib.bt_create(ctypes.byref(self._ctx), ctypes.byref(prm), cls, len(sts),
                                     ctypes.byref(buf), ctypes.c_void_p(self.stream.cuda_stream))
        if rc != _native.BT_OK:
            raise BetseNativeError(f"bt_create failed ({rc}): {self._lib.bt_last_error(None).decode()}")
        if net is not None:
            self._net = net.pack(p0.R, sts[0].T)
            self._check(self._lib.bt_set_network(self._ctx, ctypes.byref(self._net),
                                                 ctypes.c_void_p(d["c_net"].data_ptr())))
        self._check(self._lib.bt_update_v(self._ctx))
        if self.tuning is not None:
            self.set_tuning(self.tuning)
        elif self.autotune:
            self.run_autotune()
        if self.use_graph:
            self._check(self._lib.bt_capture_graphs(self._ctx))
        self._phib = host["Phi_b"].copy()      # the latest Phi_b staged or in effect
        self._phib_on = bool(np.any(self._phib != 0))

    # ------------------------------------------------------------------
    def set_boundary_voltage(self, cluster: int, bound_V: dict) -> None:
        """The applied-voltage event on one cluster of the batch
        (``Simulator.set_boundary_voltage``): its ``bound_V`` set at the start of
        the next step, its ``Phi_b`` from that step's end on; the other
        clusters keep theirs (``bt_stage_boundary`` takes the whole batch's)."""
        st = self.states[cluster]
        soff = self.offsets[2]
        bv = {k: float(bound_V.get(k, st.bound_V[k])) for k in SIDES}
        phi = boundary_potential(bv, self.worlds[cluster][1].X.shape)
        nxt = self._phib.copy()
        nxt[int(soff[cluster]):int(soff[cluster + 1])] = phi
        self._check(self._lib.bt_stage_boundary(self._ctx, nxt.ctypes.data))
        self._phib = nxt
        st.bound_V = bv
        self._phib_on = self._phib_on or bool(np.any(phi != 0))

    def set_membrane_params(self, cluster: int, Dm: Optional[np.ndarray] = None, gjopen: Optional[np.ndarray] = None,
                            D_env: Optional[np.ndarray] = None) -> None:
        """An event on one cluster of the batch (``Simulator.set_membrane_params``):
        its new ``Dm_cells`` [n_ions][M_k], ``gjopen`` [M_k] or ``D_env``
        [n_ions][G_k²] written into its slices of the batch's buffers, then
        ``bt_refresh_params``."""
        torch = self.torch
        coff, moff, soff = self.offsets
        m0, m1 = int(moff[cluster]), int(moff[cluster + 1])
        s0, s1 = int(soff[cluster]), int(soff[cluster + 1])
        with torch.cuda.device(self.device), torch.cuda.stream(self.stream):
            for key, a, lo, hi in (("Dm", Dm, m0, m1), ("gjopen", gjopen, m0, m1), ("D_env", D_env, s0, s1)):
                if a is None:
                    continue
                dst = self._dev[key][..., lo:hi]
                a = np.ascontiguousarray(a, dtype=np.float64)
                if a.shape != tuple(dst.shape):
                    raise BetseSimException(f"{key}: shape {a.shape}, cluster {cluster} needs {tuple(dst.shape)}")
                dst.copy_(torch.from_numpy(a).to(self.device))
        self.stream.synchronize()
        self._check(self._lib.bt_refresh_params(self._ctx))
        st = self.states[cluster]
        if Dm is not None:
            st.Dm_cells = np.array(Dm, dtype=np.float64)
        if gjopen is not None:
            st.gjopen = np.array(gjopen, dtype=np.float64)

    def get_tuning(self) -> dict:
        t = _native.BtTuning()
        self._check(self._lib.bt_get_tuning(self._ctx, ctypes.byref(t)))
        return t.as_dict()

    def set_tuning(self, tuning: dict) -> None:
        """Force the step kernels' launch variants (``bt_set_tuning``); keys of
        ``_native.BtTuning``, missing ones keep their current value."""
        cur = self.get_tuning()
        cur.update(tuning)
        t = _native.BtTuning(**cur)
        self._check(self._lib.bt_set_tuning(self._ctx, ctypes.byref(t)))

    def run_autotune(self, n_launches: int = 20) -> dict:
        """Time every launch variant on this batch and keep the fastest of each
        kernel (``bt_autotune``); returns the choice and the trial times (ms)."""
        t = _native.BtTuning()
        ms = (ctypes.c_double * len(_native.AUTOTUNE_TRIALS))()
        self._check(self._lib.bt_autotune(self._ctx, int(n_launches), ctypes.byref(t), ms, len(ms)))
        self.autotune_report = {"chosen": t.as_dict(),
                                "trial_ms": {k: (v if v >= 0 else None) for k, v in zip(_native.AUTOTUNE_TRIALS, ms)}}
        return self.autotune_report

    # ------------------------------------------------------------------
    def _check(self, rc: int) -> None:
        if rc != _native.BT_OK:
            raise BetseNativeError(f"libbetse_amd error {rc}: {self._lib.bt_last_error(self._ctx).decode()}")

    def _destroy(self) -> None:
        if self._ctx:
            self._lib.bt_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()
        w = self.__dict__.pop("_worker", None)
        if w is not None:
            w.shutdown(wait=True)

    def close(self) -> None:
        """Release the context, the sample worker thread, the device state and
        the pinned snapshot buffers now rather than at garbage collection
        (an instability the device recorded is not raised here)."""
        if self._ctx:
            self.stream.synchronize()
        self._destroy()
        for k in ("_dev", "_flat_dev", "_flat_host", "_pinned", "_obs"):
            self.__dict__.pop(k, None)

    def __del__(self):
        try:
            self._destroy()
        except Exception:
            pass

    def step_n(self, n: int, sync: bool = True) -> None:
        self._check(self._lib.bt_step(self._ctx, int(n), 1 if self.use_graph else 0))
        if sync:
            self.sync()

    def sync(self) -> None:
        bad = ctypes.c_int64(-1)
        rc = self._lib.bt_sync(self._ctx, ctypes.byref(bad))
        if rc == _native.BT_UNSTABLE:
            self.pull(check=False)
            raise BetseSimUnstableException("a cluster of the batch became unstable",
                                            step=None if bad.value < 0 else int(bad.value))
        self._check(rc)

    def pull(self, check: bool = True) -> None:
        """Copy the device state back into every cluster's host state."""
        if check:
            self.sync()
        self.stream.synchronize()
        d = self._dev
        coff, moff, soff = self.offsets
        keys = ("c_cell", "c_lag", "vm_cell", "c_env", "E_x", "E_y", "v_env", "gates", "Phi_b") + (
            ("c_net",) if "c_net" in d else ())
        h = {k: d[k].cpu().numpy() for k in keys}
        for i, (st, (p, c)) in enumerate(zip(self.states, self.worlds)):
            a, b = coff[i], coff[i + 1]
            sa, sb = soff[i], soff[i + 1]
            st.cc_cells = h["c_cell"][:, a:b].copy()
            st.c_lag = h["c_lag"][:, a:b].copy()
            st.vm_cell = h["vm_cell"][a:b].copy()
            st.cc_env = h["c_env"][:, sa:sb].copy()
            st.E_env_x = h["E_x"][sa:sb].reshape(c.X.shape).copy()
            st.E_env_y = h["E_y"][sa:sb].reshape(c.X.shape).copy()
            st.v_env = h["v_env"][sa:sb].copy()
            st.Phi_b = h["Phi_b"][sa:sb].copy()
            if len(st.channel_specs):
                st.chan_gates = h["gates"][:, :, moff[i]:moff[i + 1]].copy()
            if "c_net" in h:
                st.c_net = h["c_net"][:, a:b].copy()

    # ------------------------------------------------------------------
    # the loop with sampled steps (observables and write2storage per cluster)
    # ------------------------------------------------------------------
    def _obs_buffers(self):
        if getattr(self, "_obs", None) is not None and self._obs_key is self._ctx.value:
            return self._obs
        torch = self.torch
        coff, moff, soff = self.offsets
        ni, C, M, G2 = len(self.states[0].names), int(coff[-1]), int(moff[-1]), int(soff[-1])
        nch = max(len(self.states[0].channel_specs), 1)
        shapes = {"fluxes_mem": (ni, M), "fluxes_gj": (ni, M), "fluxes_env_x": (ni, G2), "fluxes_env_y": (ni, G2),
                  "chan_flux": (nch, M), "rate_NaKATP": (M,), "Jn": (M,), "I_mem": (M,), "E_gj_x": (M,),
                  "E_gj_y": (M,), "J_cell_x": (C,), "J_cell_y": (C,), "vm_ave": (C,), "J_env_x": (G2,),
                  "J_env_y": (G2,), "Jtx": (G2,), "Jty": (G2,), "B_field": (G2,), "vm": (M,)}
        with torch.cuda.device(self.device), torch.cuda.stream(self.stream):
            t = {k: torch.zeros(v, dtype=torch.float64, device=self.device) for k, v in shapes.items()}
        obs = _native.BtObs()
        for k in _native.OBS_FIELDS:
            setattr(obs, k, t[k].data_ptr())
        self._obs = (obs, t)
        self._obs_key = self._ctx.value
        return self._obs

    # per-cluster slicing of the batch's fields: by cell, membrane or square
    _CELL_FIELDS = ("c_cell", "vm_cell", "c_net", "J_cell_x", "J_cell_y", "vm_ave")
    _MEM_FIELDS = ("rate_NaKATP", "Jn", "I_mem", "E_gj_x", "E_gj_y", "vm")

    def _split(self, h: dict, i: int) -> dict:
        coff, moff, soff = self.offsets
        out = {}
        for k, a in h.items():
            lo, hi = ((coff[i], coff[i + 1]) if k in self._CELL_FIELDS else
                      (moff[i], moff[i + 1]) if k in self._MEM_FIELDS else (soff[i], soff[i + 1]))
            out[k] = a[..., lo:hi]
        return out

    def _sample_worker(self):
        w = self.__dict__.get("_worker")
        if w is None:
            from concurrent.futures import ThreadPoolExecutor
            w = self._worker = ThreadPoolExecutor(max_workers=1)
        return w

    def _sample_keys(self):
        d = self._dev
        state_keys = ["c_cell", "vm_cell", "c_env", "E_x", "E_y", "v_env"] + (["c_net"] if "c_net" in d else [])
        if self.__dict__.get("_phib_on"):
            state_keys.append("Phi_b")
        return state_keys, list(HostState._OBS_KEEP)

    def prepare_sampling(self) -> None:
        """Allocate what a sampled step uses -- the observation buffers, the
        observables' workspace (DST bases and eigenvalue tables) and the pinned
        snapshot buffers -- without stepping: the state is not touched (the
        observables are derived once from the zeroed flux buffers)."""
        obs, t_obs = self._obs_buffers()
        self._check(self._lib.bt_observables(self._ctx, ctypes.byref(obs)))
        state_keys, keep = self._sample_keys()
        d = self._dev
        for slot in (0, 1):
            self._fetch_start_async({**{k: d[k] for k in state_keys}, **{k: t_obs[k] for k in keep}}, slot)
            self._fetch_finish_async(state_keys + keep, slot)
        self._sample_worker()
        self.sync()

    def run_loop(self, n_steps: int, t_resample: Optional[float] = None, callbacks=None) -> int:
        """``_run_sim_core_loop`` (``sim.py:1132-1390``) of every cluster at once:
        ``n_steps`` steps, and at the sampled steps of the phase's time vector
        (``time_steps_sampled``, Q16) the observables of ``get_current`` for the
        whole batch (``bt_observables``) and each cluster's ``write2storage``
        (``sim.py:1789-1884``) into its own host state. The host side of a
        sample runs while the device steps on. Returns the number of sampled
        steps."""
        p0 = self.worlds[0][0]
        tr = p0.t_resample if t_resample is None else t_resample
        time_steps, sampled = phase_time_steps(int(n_steps), p0.dt, tr)
        sampled_idx = [i for i, t in enumerate(time_steps) if t in sampled]
        d = self._dev
        obs, t_obs = self._obs_buffers() if sampled_idx else (None, None)
        state_keys, keep = self._sample_keys()
        use_graph = 1 if self.use_graph else 0
        cleared = False

        def clear():
            # every cluster's series restart with the phase; cleared once the
            # first steps are queued, so that this host work overlaps them
            nonlocal cleared
            if not cleared:
                for st in self.states:
                    st.clear_storage()
                cleared = True

        def finish(t_sample, slot):
            h = self._fetch_finish_async(state_keys + keep, slot)
            for i, st in enumerate(self.states):
                hi = self._split(h, i)
                st._set_state(hi, False)
                st._set_observables(hi)
                st.write2storage(t_sample)

        # A sample's host side (copy-out and every cluster's write2storage) runs
        # on a worker thread, in sample order, while this thread keeps the
        # device stepping; two snapshot slots alternate, so a sample waits for
        # the one two before it. Progress callbacks run here, in order, once a
        # sample is stored.
        pending = deque()

        def drain(block_until: int):
            while pending and (len(pending) > block_until or pending[0].done()):
                pending.popleft().result()
                if callbacks is not None:
                    callbacks.progressed_next()

        # Per sample: wait for the device to reach it (the instability check),
        # queue its observables and its snapshot on the device, queue the
        # steps up to the next sample (or the rest), and only then wait for
        # the host side of the sample two before (its slot's pinned buffer)
        # and queue the snapshot's host copy: the device is not left idle
        # while this thread waits for the worker.
        pos = 0
        slot = 0

        def queue_steps_after(n):
            # the steps after sampled step n: up to the next sampled step, or the rest
            nonlocal pos
            nxt = sampled_idx[n + 1] if n + 1 < len(sampled_idx) else None
            if nxt is not None:
                self._check(self._lib.bt_step_observed(self._ctx, int(nxt + 1 - pos), use_graph, ctypes.byref(obs)))
                pos = nxt + 1
            elif pos < len(time_steps):
                self._check(self._lib.bt_step(self._ctx, len(time_steps) - pos, use_graph))
                pos = len(time_steps)

        try:
            if sampled_idx:
                s0 = sampled_idx[0]
                self._check(self._lib.bt_step_observed(self._ctx, int(s0 + 1), use_graph, ctypes.byref(obs)))
                pos = s0 + 1
            for n, s in enumerate(sampled_idx):
                clear()
                self.sync()
                self._check(self._lib.bt_observables(self._ctx, ctypes.byref(obs)))
                self._fetch_start_async({**{k: d[k] for k in state_keys}, **{k: t_obs[k] for k in keep}}, slot,
                                        host=False)
                queue_steps_after(n)
                drain(1)                          # this slot's previous sample is stored
                self._fetch_host_async(slot)
                pending.append(self._sample_worker().submit(finish, time_steps[s], slot))
                slot ^= 1
            if pos < len(time_steps):
                self._check(self._lib.bt_step(self._ctx, len(time_steps) - pos, use_graph))
            clear()
            drain(0)
            self.sync()
        except BaseException:
            # the samples taken before the failure are stored, as the
            # reference's loop has stored them when it raises
            for f in pending:
                f.result()
            raise
        return len(sampled_idx)

    def snapshot(self):
        """(c_cell [n_ions][C], vm_cell [C]) of the whole batch, on the host."""
        self.sync()
        self.stream.synchronize()
        return self._dev["c_cell"].cpu().numpy().copy(), self._dev["vm_cell"].cpu().numpy().copy()

    def residual_window(self, window: int) -> np.ndarray:
        """Run ``window`` steps and return, per cluster, max |c - c_prev| / (window |c|)
        over cells and ions (the per-step relative residual of the init phase)."""
        torch = self.torch
        d = self._dev
        with torch.cuda.device(self.device), torch.cuda.stream(self.stream):
            prev = d["c_cell"].clone()
            r = torch.empty(len(self.states), dtype=torch.float64, device=self.device)
        self._check(self._lib.bt_step(self._ctx, int(window), 1 if self.use_graph else 0))
        self._check(self._lib.bt_residual(self._ctx, ctypes.c_void_p(prev.data_ptr()),
                                          ctypes.c_void_p(d["cell_cluster"].data_ptr()),
                                          ctypes.c_void_p(r.data_ptr())))
        self.sync()
        return r.cpu().numpy() / float(window)

    def time_kernel(self, name: str, n_launches: int = 50) -> float:
        ms = ctypes.c_double(0.0)
        self._check(self._lib.bt_time_kernel(self._ctx, _native.KERNEL_IDS[name], int(n_launches),
                                             ctypes.byref(ms)))
        return ms.value

    @property
    def n_cells(self) -> int:
        return int(self.offsets[0][-1])


def batched_worlds(n_clusters: int = 256, cells_range=(100, 5000), grid_range=(32, 128), seed: int = 7,
                   base: str = "small", decade: float = 1.0):
    """Independent clusters of the batched workload.

    Radii are uniform between the radii that give ``cells_range`` cells on the
    10 um jittered hex lattice; grids scale linearly with the radius between
    ``grid_range``; per cluster the Na/K/Cl membrane diffusivities and the NaK
    pump rate are log-uniform over ``decade`` decades centred on the base
    values (factor ``10**U(-decade/2, decade/2)``); every cluster's lattice has
    its own numpy seed.
    """
    rng = np.random.RandomState(seed)
    # hex lattice: one cell per (sqrt(3)/2) spacing^2
    area = (np.sqrt(3.0) / 2.0) * (10.0e-6) ** 2
    r_lo = np.sqrt(cells_range[0] * area / np.pi)
    r_hi = np.sqrt(cells_range[1] * area / np.pi)
    out = []
    for k in range(n_clusters):
        p = Parameters.preset(base)
        r = rng.uniform(r_lo, r_hi)
        p.cluster_radius = float(r)
        p.grid_size = int(round(grid_range[0] + (r - r_lo) / (r_hi - r_lo) * (grid_range[1] - grid_range[0])))
        p.seed = int(rng.randint(0, 2 ** 31 - 1))
        f = 10.0 ** rng.uniform(-decade / 2, decade / 2, 4)
        p.mem_perms = dict(p.mem_perms)
        for j, ion in enumerate(("Na", "K", "Cl")):
            p.mem_perms[ion] = p.mem_perms[ion] * f[j]
        p.alpha_NaK = p.alpha_NaK * f[3]
        cells = Cells.make_world(p)
        if p.channels and p.channel_random_state:
            # every cluster its own seeded channel multipliers and gates
            p.channel_seed = int(rng.randint(0, 2 ** 31 - 1))
            draw_channel_state(p, cells.n_cells, cells.n_mems)
        out.append((p, cells))
    return out
