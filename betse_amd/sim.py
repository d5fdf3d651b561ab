"""GPU ``Simulator``: the drop-in for ``Simulator._run_sim_core_loop``.

Mirrors the reference's ``betse/science/sim.py`` interface for the full-solver
electrodiffusion path:

* ``init_core(phase)``      -- ``sim.py:452-754``: state arrays from the ion profile;
* ``init_dynamics(phase)``  -- ``sim.py:758-1012``: per-phase set-up (tissue Dm,
  D_env, ko_env, noise, blocks);
* ``run_sim_core(phase)``   -- ``sim.py:1016-1128``: init_dynamics, storage reset,
  ``update_V``, time vector, the loop, instability handling;
* ``_run_sim_core_loop(phase, time_steps, time_steps_sampled, anim_cells)`` --
  ``sim.py:1132-1390``: every step runs on the GPU through the C ABI of
  ``libbetse_amd.so``; the host only wakes at sampled steps to run
  ``write2storage`` (``sim.py:1789-1884``) and the progress callbacks.

State lives in device tensors (PyTorch holds the memory; the kernels are the
library's own HIP code). Host NumPy mirrors (``cc_cells``, ``cc_env``, ``vm`` ...)
are refreshed at sampled steps and at the end of a phase.
"""

from __future__ import annotations

import ctypes
import enum
import os
from typing import Dict, List, Optional

import numpy as np

from . import _native
from .boundary import SIDES, boundary_potential
from .cells import Cells
from .channels import CHANNEL_KINDS, draw_channel_state, init_gates
from .exceptions import BetseNativeError, BetseSimException, BetseSimUnstableException
from .parameters import Parameters
from .assemble import BUFFER_FIELDS, batch_geometry, batch_state, make_clusters, make_params, make_tiles  # noqa: F401


class SimPhaseKind(enum.Enum):
    """``betse/science/enum/enumphase.py:17``."""
    SEED = 1
    INIT = 2
    SIM = 3


class SolverType(enum.Enum):
    """``betse/science/enum/enumconf.py:120-125``; only FULL has a GPU loop."""
    FULL = 1
    FAST = 2


class SimCallbacks:
    """Progress callbacks (``phase/phasecallbacks.py:20-110``)."""

    def __init__(self):
        self.progress_max = 0
        self.progress = 0

    def progress_ranged(self, progress_max: int, progress_min: int = 0) -> None:
        self.progress_max = progress_max
        self.progress = progress_min

    def progressed_next(self) -> None:
        self.progress += 1


class SimPhase:
    """A seed/init/sim phase (``phase/phasecls.py:46-204``).

    Constructing an INIT or SIM phase sets ``p.dt``, ``p.resample`` and
    ``p.t_resample`` (``_init_kludge_time``, ``phasecls.py:315-330``).
    """

    def __init__(self, kind: SimPhaseKind, p: Parameters, cells: Optional[Cells] = None,
                 sim: Optional["Simulator"] = None, callbacks: Optional[SimCallbacks] = None,
                 save_path: Optional[str] = None):
        self.kind = kind
        self.p = p
        self.cells = cells
        self.sim = sim
        self.callbacks = callbacks or SimCallbacks()
        self.save_path = save_path
        if kind is SimPhaseKind.INIT:
            p.set_phase_time("init")
        elif kind is SimPhaseKind.SIM:
            p.set_phase_time("sim")

    def die_unless_kind_init_or_sim(self) -> None:
        if self.kind not in (SimPhaseKind.INIT, SimPhaseKind.SIM):
            raise BetseSimException(f"phase {self.kind} is neither init nor sim")


def gaussian_taps(sigma: float = 1.0, truncate: float = 4.0) -> np.ndarray:
    """Taps of ``scipy.ndimage.gaussian_filter1d`` (centre first), built with
    SciPy's published formula ``exp(-x^2/(2 sigma^2))`` normalised to sum 1."""
    radius = int(truncate * float(sigma) + 0.5)
    sigma2 = sigma * sigma
    x = np.arange(-radius, radius + 1)
    phi_x = np.exp(-0.5 / sigma2 * x ** 2)
    phi_x = phi_x / phi_x.sum()
    return phi_x[radius:]


def phase_time_steps(n_steps: int, dt: float, t_resample: float):
    """``time_steps`` and ``time_steps_sampled`` (``sim.py:2455-2462``, Q16).
    ``t_resample`` is the sampling interval in steps; below one step the
    reference's loop never advances, so it is refused here."""
    if not t_resample >= 1:
        raise ValueError(f"t_resample must be at least one step, got {t_resample!r}")
    time_steps = np.linspace(0, n_steps * dt, n_steps)
    sampled = set()
    i = 0
    while i < len(time_steps) - t_resample:
        i += t_resample
        i = int(i)
        sampled.add(time_steps[i])
    return time_steps, sampled


class HostState:
    """Host-side state of one cluster, set up as the reference's ``Simulator``
    does before its loop (``init_core``, ``init_dynamics``); no device work."""

    # ------------------------------------------------------------------
    # set-up mirrored from the reference
    # ------------------------------------------------------------------
    def init_core(self, phase: SimPhase) -> None:
        p, cells = phase.p, phase.cells
        self.p = p
        self.cells = cells
        self.mdl = cells.n_mems
        self.cdl = cells.n_cells
        self.edl = cells.X.size
        self.T = p.T
        self.names = p.ion_names()
        self.ionlabel = {}
        cc_cells, cc_env, zs, Dm, D_free, D_gj, D_env, cbound, Dtj = [], [], [], [], [], [], [], [], []
        for i, name in enumerate(self.names):
            setattr(self, "i" + name, i)
            self.ionlabel[i] = name
            cc_cells.append(np.full(self.cdl, p.cell_concs[name]))
            cc_env.append(np.full(self.edl, p.env_concs[name]))
            zs.append(p.ion_charge[name])
            Dm.append(np.full(self.mdl, p.mem_perms[name]))
            D_gj.append(np.full(self.mdl, p.free_diff[name]))
            D_env.append(np.full(self.edl, p.free_diff[name]))
            D_free.append(p.free_diff[name])
            cbound.append(p.env_concs[name])
            Dtj.append(p.Dtj_rel[name])
        self.movingIons = list(range(len(self.names)))
        self.cc_cells = np.asarray(cc_cells)
        self.cc_env = np.asarray(cc_env)
        self.zs = np.asarray(zs)
        self.Dm_cells = np.asarray(Dm)
        self.D_free = np.asarray(D_free)
        self.D_gj = np.asarray(D_gj)
        self.D_env = np.asarray(D_env)
        self.Dtj_rel = Dtj
        self.c_env_bound = cbound
        self.bound_V = {"T": 0, "B": 0, "L": 0, "R": 0}
        self.initDenv(cells, p)
        self.gj_block = 1
        self.gjopen = np.ones(self.mdl) * cells.gj_default_weights
        # cc_at_mem = cc_cells[mem_to_cells] (sim.py:744-745) is constant over a
        # cell's membranes on this path, so the device keeps it per cell.
        self.c_lag = self.cc_cells.copy()
        self.vm_cell = np.zeros(self.cdl)
        self.E_env_x = np.zeros(cells.X.shape)
        self.E_env_y = np.zeros(cells.X.shape)
        self.v_env = np.zeros(self.edl)
        self.Phi_b = np.zeros(self.edl)
        self.init_channels(cells, p)
        # reaction network substances (MasterOfNetworks cell_concs), uniform start
        self.network = getattr(p, "network", None)
        self.c_net = (np.tile(np.asarray(self.network.c0, dtype=np.float64)[:, None], (1, self.cdl))
                      if self.network is not None else None)

    def init_channels(self, cells: Cells, p: Parameters) -> None:
        """``Channel.init_channel`` (networks.py:6550-6628) for every configured
        channel, on all membranes: gates from ``_init_state`` at the current
        Vmem, or the workload's seeded draws."""
        specs = tuple(getattr(p, "channels", ()))
        if len(specs) > _native.BT_MAX_CHAN:
            raise BetseSimException("too many channels")
        if specs and p.channel_random_state and p.channel_gates0 is None:
            draw_channel_state(p, self.cdl, self.mdl)
        self.channel_specs = specs
        self.chan_gates = np.zeros((len(specs), 2, self.mdl))
        self.chan_mod = np.ones((len(specs), self.cdl))
        ions = []
        for k, spec in enumerate(specs):
            ion = CHANNEL_KINDS[spec.kind][0]
            if ion not in self.names:
                raise BetseSimException(f"channel {spec.kind} needs ion {ion}")
            if ion in ions:
                raise BetseSimException("the GPU step supports one channel per ion")
            ions.append(ion)
            m, h = init_gates(spec.kind, self.vm)
            if p.channel_gates0 is not None and p.channel_gates0[k] is not None:
                m, h = p.channel_gates0[k]
            self.chan_gates[k, 0] = m
            self.chan_gates[k, 1] = h
            if p.channel_cell_mult is not None and p.channel_cell_mult[k] is not None:
                self.chan_mod[k] = p.channel_cell_mult[k]

    def initDenv(self, cells: Cells, p: Parameters) -> None:
        """``sim.py:2347-2400`` (ECM branch)."""
        for i in range(len(self.D_env)):
            Denv_o = np.ones(self.edl) * self.D_free[i]
            Denv_o[cells.envInds_inClust] = self.D_free[i] * p.D_adh
            Denv_o[cells.all_bound_mem_inds] = self.D_free[i] * p.D_tj * self.Dtj_rel[i]
            Denv_o[cells.interior_bound_mem_inds] = self.D_free[i] * p.D_tj * self.Dtj_rel[i]
            Denv_o[cells.ecm_inds_bound_cell] = self.D_free[i] * p.D_tj * self.Dtj_rel[i]
            self.D_env[i] = Denv_o * 1.0
        self.TJ_modulator = np.ones(self.D_env.shape)

    def init_dynamics(self, phase: SimPhase) -> None:
        """``sim.py:758-1012`` for the electrodiffusion path."""
        p, cells = phase.p, phase.cells
        if p.v_sensitive_gj:
            raise BetseSimException("voltage-sensitive gap junctions are not supported by the GPU step")
        if p.sharpness < 1.0:
            raise BetseSimException("env smoothing (sharpness < 1) is not supported by the GPU step")
        if p.cell_polarizability != 0.0 or not p.is_ecm or p.fast_update_ecm:
            raise BetseSimException("the GPU step covers the ECM, non-polarizable, slow-update path")
        nfrac = p.smooth_cells
        nm = cells.num_mems[cells.mem_to_cells]
        self.smooth_weight_mem = (nfrac * nm - 1) / (nfrac * nm)
        self.smooth_weight_o = 1 / (nfrac * nm)
        self.initDenv(cells, p)
        if p.cbnd is not None:
            for i, name in enumerate(self.names):
                self.c_env_bound[i] = p.cbnd[name]
        # tissue profiles (tishandler.py:363-443): default tissue, then extras
        for i, name in enumerate(self.names):
            self.Dm_cells[i][:] = p.mem_perms[name]
        for cell_inds, dms in getattr(p, "tissue_profiles", ()):
            mems = np.concatenate([cells.cell_to_mems[c] for c in cell_inds]).astype(np.int64)
            for name, val in dms.items():
                if name in self.names:
                    self.Dm_cells[self.names.index(name)][mems] = val
        self.D_env_base = np.copy(self.D_env)
        self.Dm_base = np.copy(self.Dm_cells)
        self.Dm_scheduled = np.zeros_like(self.Dm_cells)
        self.channel_noise_factor = np.random.random(self.mdl)
        if "K" in self.names:
            iK = self.iK
            self.Dm_cells[iK] = (p.channel_noise_level * self.channel_noise_factor + 1) * self.Dm_cells[iK]
        self.gj_block = 1
        self.NaKATP_block = p.NaKATP_block
        self.CaATP_block = np.ones(self.mdl)
        self.rho_pump = 1
        self.rho_channel = 1
        self.ko_env = (np.sqrt(np.dot((p.NAv * (p.q ** 2) * self.zs ** 2) / (p.er * p.eo * p.kb * p.T),
                                      self.cc_env))).mean()
        self.ko_cell = (np.sqrt(np.dot((p.NAv * (p.q ** 2) * self.zs ** 2) / (p.er * p.eo * p.kb * p.T),
                                       self.cc_cells))).mean()

    def clear_storage(self, cells=None, p=None) -> None:
        """Time-series lists (``sim.py:1680-1786``, the fields this path fills)."""
        self.time: List[float] = []
        self.cc_time: List[np.ndarray] = []
        self.cc_env_time: List[np.ndarray] = []
        self.dd_time: List[np.ndarray] = []
        self.vm_time: List[np.ndarray] = []
        self.vm_ave_time: List[np.ndarray] = []
        self.rho_cells_time: List[np.ndarray] = []
        self.venv_time: List[np.ndarray] = []
        self.efield_ecm_x_time: List[np.ndarray] = []
        self.efield_ecm_y_time: List[np.ndarray] = []
        self.gjopen_time: List[np.ndarray] = []
        self.efield_gj_x_time: List[np.ndarray] = []
        self.efield_gj_y_time: List[np.ndarray] = []
        self.I_cell_x_time: List[np.ndarray] = []
        self.I_cell_y_time: List[np.ndarray] = []
        self.I_mem_time: List[np.ndarray] = []
        self.rate_NaKATP_time: List[np.ndarray] = []
        self.I_tot_x_time: List[np.ndarray] = []
        self.I_tot_y_time: List[np.ndarray] = []
        self.Jn = None

    # ------------------------------------------------------------------
    # host copies of the device state and the sampled-step storage
    # ------------------------------------------------------------------
    def _set_state(self, h: dict, full: bool) -> None:
        shp = self.cells.X.shape
        self.__dict__.pop("vm_ave_obs", None)       # observables of an earlier state
        self.__dict__.pop("vm_obs", None)
        self.cc_cells = h["c_cell"]
        self.vm_cell = h["vm_cell"]
        self.cc_env = h["c_env"]
        self.E_env_x = h["E_x"].reshape(shp)
        self.E_env_y = h["E_y"].reshape(shp)
        self.v_env = h["v_env"]
        if full:
            self.c_lag = h["c_lag"]
            if len(self.channel_specs):
                self.chan_gates = h["gates"]
        if self.network is not None:
            self.c_net = h["c_net"]
        if "Phi_b" in h:
            self.Phi_b = h["Phi_b"]

    # the observables write2storage keeps; the raw fluxes stay on the device (pull_raw_fluxes)
    _OBS_KEEP = ("rate_NaKATP", "Jn", "I_mem", "E_gj_x", "E_gj_y", "J_cell_x", "J_cell_y", "vm_ave", "vm", "J_env_x",
                 "J_env_y", "Jtx", "Jty", "B_field")

    def _set_observables(self, h: dict) -> None:
        shp = self.cells.X.shape
        self.rate_NaKATP = h["rate_NaKATP"]
        self.Jn, self.I_mem = h["Jn"], h["I_mem"]
        self.E_gj_x, self.E_gj_y = h["E_gj_x"], h["E_gj_y"]
        self.J_cell_x, self.J_cell_y = h["J_cell_x"], h["J_cell_y"]
        self.vm_ave_obs = h["vm_ave"]
        self.vm_obs = h["vm"]
        self.J_env_x, self.J_env_y = h["J_env_x"].reshape(shp), h["J_env_y"].reshape(shp)
        self.Jtx, self.Jty = h["Jtx"].reshape(shp), h["Jty"].reshape(shp)
        self.B_field = h["B_field"].reshape(shp)

    def write2storage(self, t, cells=None, p=None) -> None:
        """``sim.py:1789-1884`` for the fields of this path. Every sampled step
        pulls fresh host arrays from the device (nothing on the host modifies
        them afterwards), so they are stored without a further copy; Dm and
        gjopen do not change within a phase on this path and are stored as one
        snapshot per phase."""
        self.time.append(t)
        if getattr(self, "Jn", None) is not None:
            self.efield_gj_x_time.append(self.E_gj_x)
            self.efield_gj_y_time.append(self.E_gj_y)
            self.I_cell_x_time.append(self.J_cell_x)
            self.I_cell_y_time.append(self.J_cell_y)
            self.I_mem_time.append(self.I_mem)
            self.rate_NaKATP_time.append(self.rate_NaKATP)
            self.I_tot_x_time.append(self.J_env_x)
            self.I_tot_y_time.append(self.J_env_y)
        self.cc_time.append(self.cc_cells)
        self.cc_env_time.append(self.cc_env)
        snap = self.__dict__.get("_phase_snapshot")
        if snap is None or len(self.time) == 1:
            snap = (np.copy(self.Dm_cells), self.gjopen * 1)
            self._phase_snapshot = snap
        self.dd_time.append(snap[0])
        # update_V's vm per membrane: the sampled step derived it on the device
        vm = self.__dict__.get("vm_obs")
        if vm is None:
            vm = self.vm
        self.vm_time.append(vm)
        # update_V's vm_ave (sim.py:2038): the sampled step derived it on the device
        ave = self.__dict__.get("vm_ave_obs")
        self.vm_ave_time.append(ave if ave is not None else
                                np.add.reduceat(vm, self.cells.mem_ptr[:-1]) / self.cells.num_mems)
        self.rho_cells_time.append(self.rho_cells)
        self.venv_time.append(self.v_env)
        self.efield_ecm_x_time.append(self.E_env_x)
        self.efield_ecm_y_time.append(self.E_env_y)
        self.gjopen_time.append(snap[1])

    # ------------------------------------------------------------------
    # host views of the device state
    # ------------------------------------------------------------------
    @property
    def vm(self) -> np.ndarray:
        """Vmem per membrane: ``(1/cm)*rho_surf[mem_to_cells] - Phi_b[map_mem2ecm]``
        (x - 0.0 is x, so a zero Phi_b is not gathered)."""
        vm = self.vm_cell[self.cells.mem_to_cells]
        if self.Phi_b.any():
            vm = vm - self.Phi_b[self.cells.map_mem2ecm]
        return vm

    @property
    def cc_at_mem(self) -> np.ndarray:
        return self.c_lag[:, self.cells.mem_to_cells]

    @property
    def vm_ave(self) -> np.ndarray:
        cells = self.cells
        return np.add.reduceat(self.vm, cells.mem_ptr[:-1]) / cells.num_mems

    @property
    def rho_cells(self) -> np.ndarray:
        return np.dot(self.zs * self.p.F, self.cc_cells)


COPY_THREADS = max(1, min(16, len(os.sched_getaffinity(0))))  # host threads of a sampled step's copy-out


class DeviceFetch:
    """Device -> host copies through cached pinned buffers, for objects with
    ``torch``, ``device`` and ``stream`` attributes (``Simulator``,
    ``BatchSimulator``)."""

    def _fetch_start(self, tensors: dict) -> None:
        """Queue device -> pinned host copies on the simulation stream."""
        torch = self.torch
        pinned = self.__dict__.setdefault("_pinned", {})
        with torch.cuda.device(self.device), torch.cuda.stream(self.stream):
            for k, t in tensors.items():
                h = pinned.get(k)
                if h is None or h.shape != t.shape or h.dtype != t.dtype:
                    h = torch.empty(t.shape, dtype=t.dtype, pin_memory=True)
                    pinned[k] = h
                h.copy_(t, non_blocking=True)

    def _fetch_start_async(self, tensors: dict, slot: int = 0, host: bool = True) -> None:
        """Sampled steps: snapshot the (FP64) tensors into a flat device buffer
        on the simulation stream (a single multi-tensor copy), then copy that
        buffer to pinned host memory on a side stream (default priority), so the
        host transfer overlaps the next steps. ``_fetch_finish_async(keys,
        slot)`` waits for it. Each slot has its own pair of buffers, reused: the
        snapshot waits on the device for the slot's previous transfer, and a
        caller finishes a slot's fetch before it starts the slot's next one.
        ``host=False`` takes the device snapshot only; ``_fetch_host_async(slot)``
        then queues its host copy (once the slot's previous sample is read)."""
        torch = self.torch
        layout = [(k, tuple(t.shape)) for k, t in tensors.items()]
        if any(t.dtype != torch.float64 for t in tensors.values()):
            raise BetseSimException("sampled-step fetch expects FP64 tensors")
        if self.__dict__.get("_flat_layout") != layout:
            self._flat_dev, self._flat_host, self._copied = {}, {}, {}
            self._flat_at, o = {}, 0
            for k, shape in layout:
                n = int(np.prod(shape))
                self._flat_at[k] = (o, n, shape)
                o += n
            self._flat_total = o
            self._flat_layout = layout
        if slot not in self._flat_dev:
            self._flat_dev[slot] = torch.empty(self._flat_total, dtype=torch.float64, device=self.device)
            self._flat_host[slot] = torch.empty(self._flat_total, dtype=torch.float64, pin_memory=True)
        if self.__dict__.get("_copy_stream") is None:
            self._copy_stream = torch.cuda.Stream(device=self.device)     # default priority
        fdev = self._flat_dev[slot]
        with torch.cuda.device(self.device):
            with torch.cuda.stream(self.stream):
                prev = self._copied.get(slot)
                if prev is not None:
                    self.stream.wait_event(prev)
                dst = [fdev[o:o + n].view(shape) for o, n, shape in self._flat_at.values()]
                torch._foreach_copy_(dst, [tensors[k] for k in self._flat_at])
                taken = torch.cuda.Event()
                taken.record(self.stream)
        self.__dict__.setdefault("_taken", {})[slot] = taken
        if host:
            self._fetch_host_async(slot)

    def _fetch_host_async(self, slot: int = 0) -> None:
        """The host copy of the slot's device snapshot, on the side stream."""
        torch = self.torch
        fdev = self._flat_dev[slot]
        taken = self._taken.pop(slot)
        with torch.cuda.device(self.device):
            with torch.cuda.stream(self._copy_stream):
                self._copy_stream.wait_event(taken)
                self._flat_host[slot].copy_(fdev, non_blocking=True)
                done = torch.cuda.Event()
                done.record(self._copy_stream)
                self._copied[slot] = done

    def _fetch_finish_async(self, keys, slot: int = 0) -> dict:
        """Host arrays of the slot's last ``_fetch_start_async``: views into one
        fresh NumPy copy of the pinned buffer (which is reused), so each
        sample's arrays are its own. The copy runs in chunks on a few host
        threads (NumPy releases the GIL while it copies)."""
        self._copied[slot].synchronize()
        src = self._flat_host[slot].numpy()
        flat = np.empty_like(src)
        n = len(src)
        if n < (1 << 22):
            flat[:] = src
        else:
            pool = self.__dict__.get("_copy_pool")
            if pool is None:
                from concurrent.futures import ThreadPoolExecutor
                pool = self._copy_pool = ThreadPoolExecutor(max_workers=COPY_THREADS)
            step = (n + COPY_THREADS - 1) // COPY_THREADS
            list(pool.map(lambda a: np.copyto(flat[a:a + step], src[a:a + step]), range(0, n, step)))
        out = {}
        for k in keys:
            o, n, shape = self._flat_at[k]
            out[k] = flat[o:o + n].reshape(shape)
        return out

    def _fetch_finish(self, keys) -> dict:
        """Host arrays of pinned copies that have completed (NumPy copies out of
        the pinned buffers, which are reused; torch's CPU clone of a pinned
        tensor was measured ~14x slower on the largest cluster)."""
        pinned = self._pinned
        return {k: pinned[k].numpy().copy() for k in keys}

    def _fetch(self, tensors: dict) -> dict:
        """Device tensors -> host arrays through cached pinned buffers: all copies
        are queued on the simulation stream, then one synchronisation."""
        self._fetch_start(tensors)
        self.stream.synchronize()
        return self._fetch_finish(tensors.keys())


class Simulator(HostState, DeviceFetch):
    """GPU-backed BETSE ``Simulator`` for the full electrodiffusion solver.

    ``flux_mode`` is ``"fp64"`` (every operation in FP64, the parity path) or
    ``"fp32"`` (flux laws in FP32; concentrations, charge sums and the
    charge-to-voltage step in FP64).
    """

    def __init__(self, p: Optional[Parameters] = None, *, device: Optional[str] = None,
                 flux_mode: str = "fp64", use_graph: bool = True, tuning: Optional[dict] = None):
        import torch
        if not torch.cuda.is_available():
            raise BetseSimException("no GPU visible: the GPU simulator has no CPU fallback")
        if flux_mode not in ("fp64", "fp32"):
            raise ValueError(flux_mode)
        self.p = p
        self.flux_mode = flux_mode
        self.use_graph = use_graph
        self.torch = torch
        self.device = torch.device(device or "cuda")
        self.stream = torch.cuda.Stream(device=self.device)     # default priority
        self._lib = _native.load()
        self._ctx = ctypes.c_void_p()
        self._geom_key = None
        self._dev: Dict[str, "torch.Tensor"] = {}
        self.exception_instability = None
        self.tuning = tuning            # forced launch variants (bt_set_tuning) or None
        self.autotune_report = None

    # ------------------------------------------------------------------
    # device set-up
    # ------------------------------------------------------------------
    def _t(self, a, dtype):
        torch = self.torch
        return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype).to(self.device, non_blocking=False)

    def _upload_geometry(self, cells: Cells) -> None:
        if self._geom_key is cells:
            return
        torch = self.torch
        geo, _ = batch_geometry([cells])
        self.n_tiles = len(geo["tile_ptr"]) - 1
        for k, a in geo.items():
            self._dev[k] = torch.from_numpy(a).to(self.device)
        self._geom_key = cells

    def _device_setup(self, phase: SimPhase) -> None:
        p, cells = phase.p, phase.cells
        torch = self.torch
        self._destroy_ctx()
        with torch.cuda.device(self.device), torch.cuda.stream(self.stream):
            self._upload_geometry(cells)
            d = self._dev
            for k, a in batch_state([self], [cells], phase.kind is SimPhaseKind.SIM).items():
                d[k] = torch.from_numpy(a).to(self.device)
            self.stream.synchronize()

        active = phase.kind is SimPhaseKind.SIM or getattr(p, "channels_init_active", True)
        prm = make_params([self], [p], [cells], self.flux_mode, active)
        cl = make_clusters([self], [p], [cells], np.zeros(1, dtype=np.int64))
        buf = _native.BtBuffers()
        for name in BUFFER_FIELDS:
            setattr(buf, name, d[name].data_ptr())
        buf.n_tiles = self.n_tiles
        buf.tile_m = _native.BT_TILE_MEMBRANES
        self._prm, self._buf, self._cl = prm, buf, cl
        with torch.cuda.device(self.device):
            rc = self._lib.bt_create(ctypes.byref(self._ctx), ctypes.byref(prm), cl, 1,
                                     ctypes.byref(buf), ctypes.c_void_p(self.stream.cuda_stream))
        if rc != _native.BT_OK:
            raise BetseNativeError(f"bt_create failed ({rc}): {self._lib.bt_last_error(None).decode()}")
        if self.network is not None:
            self._net = self.network.pack(p.R, self.T)
            self._check(self._lib.bt_set_network(self._ctx, ctypes.byref(self._net),
                                                 ctypes.c_void_p(d["c_net"].data_ptr())))
        if self.tuning is not None:
            self.set_tuning(self.tuning)
        if self.use_graph:
            self._check(self._lib.bt_capture_graphs(self._ctx))
        self._phib_on = bool(np.any(self.Phi_b != 0))

    # ------------------------------------------------------------------
    # events
    # ------------------------------------------------------------------
    def set_boundary_voltage(self, bound_V: dict) -> None:
        """The applied-voltage event (``TissueEventVoltage.fire``,
        ``tisevevolt.py:75-89``): ``bound_V`` (keys ``T``, ``B``, ``L``, ``R``;
        missing ones keep their value) set at the start of the next step. That
        step still runs with the ``Phi_b`` in effect; ``get_current`` derives
        the new ``Phi_b`` at its end (``boundary_potential``), which then holds
        (``bt_stage_boundary``)."""
        bv = {k: float(bound_V.get(k, self.bound_V[k])) for k in SIDES}
        phi = boundary_potential(bv, self.cells.X.shape)
        self._check(self._lib.bt_stage_boundary(self._ctx, phi.ctypes.data))
        self.bound_V = bv
        self._phib_on = self._phib_on or bool(np.any(phi != 0))

    def set_membrane_params(self, Dm: Optional[np.ndarray] = None, gjopen: Optional[np.ndarray] = None,
                            D_env: Optional[np.ndarray] = None) -> None:
        """An event's new membrane diffusivities (``Dm_cells``, [n_ions][M]),
        gap-junction openings (``gjopen``, [M]) or env diffusivities (``D_env``,
        [n_ions][G²]) for the next steps, as ``TissueHandler.fire_events`` ->
        ``makeAllChanges`` sets them (``tishandler.py:709-724,1321-1332``):
        written into the device buffers, then ``bt_refresh_params``."""
        torch = self.torch
        with torch.cuda.device(self.device), torch.cuda.stream(self.stream):
            for key, a in (("Dm", Dm), ("gjopen", gjopen), ("D_env", D_env)):
                if a is None:
                    continue
                a = np.ascontiguousarray(a, dtype=np.float64)
                dst = self._dev[key]
                if a.size != dst.numel():
                    raise BetseSimException(f"{key}: {a.size} values for a buffer of {dst.numel()}")
                dst.copy_(torch.from_numpy(a.reshape(dst.shape)).to(self.device))
        self.stream.synchronize()
        self._check(self._lib.bt_refresh_params(self._ctx))
        if Dm is not None:
            self.Dm_cells = np.array(Dm, dtype=np.float64)
        if gjopen is not None:
            self.gjopen = np.array(gjopen, dtype=np.float64)

    # ------------------------------------------------------------------
    # launch variants of the step kernels
    # ------------------------------------------------------------------
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
        """Time every launch variant on this cluster and keep the fastest of
        each kernel (``bt_autotune``; the state is left as it was)."""
        t = _native.BtTuning()
        ms = (ctypes.c_double * len(_native.AUTOTUNE_TRIALS))()
        self._check(self._lib.bt_autotune(self._ctx, int(n_launches), ctypes.byref(t), ms, len(ms)))
        self.autotune_report = {"chosen": t.as_dict(),
                                "trial_ms": {k: (v if v >= 0 else None) for k, v in zip(_native.AUTOTUNE_TRIALS, ms)}}
        return self.autotune_report

    def _destroy_ctx(self) -> None:
        if self._ctx:
            self._lib.bt_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self._destroy_ctx()
        except Exception:
            pass

    def _check(self, rc: int) -> None:
        if rc != _native.BT_OK:
            raise BetseNativeError(f"libbetse_amd error {rc}: {self._lib.bt_last_error(self._ctx).decode()}")

    def _state_names(self, full: bool):
        names = ["c_cell", "vm_cell", "c_env", "E_x", "E_y", "v_env"]
        if self.__dict__.get("_phib_on"):
            names.append("Phi_b")
        if full:
            names.append("c_lag")
            if len(self.channel_specs):
                names.append("gates")
        if self.network is not None:
            names.append("c_net")
        return names

    def _pull_state(self, full: bool = True) -> None:
        """Device state -> host. ``full=False`` (sampled steps) skips what
        write2storage does not store: the lagged membrane concentrations and
        the channel gates."""
        d = self._dev
        self._set_state(self._fetch({k: d[k] for k in self._state_names(full)}), full)

    def _sync(self, phase_step0: int = 0) -> None:
        bad = ctypes.c_int64(-1)
        rc = self._lib.bt_sync(self._ctx, ctypes.byref(bad))
        if rc == _native.BT_UNSTABLE:
            self._pull_state()
            step = None if bad.value < 0 else int(bad.value) - phase_step0
            raise BetseSimUnstableException(
                "Your simulation has become unstable. Please try a smaller time step,"
                "reduce gap junction radius, and/or reduce pump rate coefficients.", step=step)
        self._check(rc)

    # ------------------------------------------------------------------
    # sampled-step observables
    # ------------------------------------------------------------------
    def _obs_buffers(self):
        if getattr(self, "_obs", None) is not None and self._obs_key == (self.cdl, self.mdl, self.edl):
            return self._obs
        torch = self.torch
        f64 = torch.float64
        ni, M, C, G2 = len(self.names), self.mdl, self.cdl, self.edl
        nch = max(len(self.channel_specs), 1)
        shapes = {"fluxes_mem": (ni, M), "fluxes_gj": (ni, M), "fluxes_env_x": (ni, G2), "fluxes_env_y": (ni, G2),
                  "chan_flux": (nch, M), "rate_NaKATP": (M,), "Jn": (M,), "I_mem": (M,), "E_gj_x": (M,),
                  "E_gj_y": (M,), "J_cell_x": (C,), "J_cell_y": (C,), "vm_ave": (C,), "J_env_x": (G2,),
                  "J_env_y": (G2,), "Jtx": (G2,), "Jty": (G2,), "B_field": (G2,), "vm": (M,)}
        with torch.cuda.device(self.device), torch.cuda.stream(self.stream):
            t = {k: torch.zeros(v, dtype=f64, device=self.device) for k, v in shapes.items()}
        obs = _native.BtObs()
        for k in _native.OBS_FIELDS:
            setattr(obs, k, t[k].data_ptr())
        self._obs = (obs, t)
        self._obs_key = (self.cdl, self.mdl, self.edl)
        return self._obs

    def _step_observed(self, n: int) -> None:
        obs, _ = self._obs_buffers()
        self._check(self._lib.bt_step_observed(self._ctx, int(n), 1 if self.use_graph else 0, ctypes.byref(obs)))

    def _pull_observables(self) -> None:
        """Observables of the last observed step (get_current, ion_current.py:13-109)."""
        obs, t = self._obs_buffers()
        self._check(self._lib.bt_observables(self._ctx, ctypes.byref(obs)))
        self._set_observables(self._fetch({k: t[k] for k in self._OBS_KEEP}))

    def pull_raw_fluxes(self) -> None:
        """The last observed step's raw fluxes (fluxes_mem, fluxes_gj,
        fluxes_env_x/y, channel fluxes) to the host."""
        _, t = self._obs_buffers()
        h = self._fetch({k: t[k] for k in ("fluxes_mem", "fluxes_gj", "fluxes_env_x", "fluxes_env_y", "chan_flux")})
        self.fluxes_mem, self.fluxes_gj = h["fluxes_mem"], h["fluxes_gj"]
        self.fluxes_env_x, self.fluxes_env_y = h["fluxes_env_x"], h["fluxes_env_y"]
        self.chan_flux = h["chan_flux"][:len(self.channel_specs)]

    def step_observed(self, n: int) -> None:
        """Run n steps; the last one's observables are pulled to the host."""
        self._step_observed(n)
        self._sync()
        self._pull_state()
        self._pull_observables()

    # ------------------------------------------------------------------
    # update_V and the loop
    # ------------------------------------------------------------------
    def update_V(self, cells=None, p=None) -> None:
        """Voltage part of ``update_V`` (``sim.py:2007-2038``) on the device."""
        self._check(self._lib.bt_update_v(self._ctx))

    def run_sim_core(self, phase: SimPhase) -> None:
        """``sim.py:1016-1128``."""
        phase.die_unless_kind_init_or_sim()
        self.cells = phase.cells
        self.init_dynamics(phase)
        self.clear_storage(phase.cells, phase.p)
        self._device_setup(phase)
        self.update_V(phase.cells, phase.p)
        n = phase.p.init_tsteps if phase.kind is SimPhaseKind.INIT else phase.p.sim_tsteps
        time_steps, time_steps_sampled = phase_time_steps(n, phase.p.dt, phase.p.t_resample)
        phase.callbacks.progress_ranged(progress_max=len(time_steps_sampled))
        self.exception_instability = None
        if getattr(phase.p, "solver_type", SolverType.FULL) is not SolverType.FULL:
            raise BetseSimException("only the full solver runs on the GPU")
        try:
            self._run_sim_core_loop(phase=phase, time_steps=time_steps,
                                    time_steps_sampled=time_steps_sampled, anim_cells=None)
        except BetseSimUnstableException as exception:
            self.exception_instability = exception
        if phase.save_path:
            self.save(phase.save_path)
        if self.exception_instability is not None:
            raise self.exception_instability

    def _run_sim_core_loop(self, phase: SimPhase, time_steps: np.ndarray, time_steps_sampled: set,
                           anim_cells=None) -> None:
        """Run ``len(time_steps)`` steps on the GPU; wake at sampled steps."""
        if anim_cells is not None:
            raise BetseSimException("in-loop animation is not supported by the GPU step")
        step0 = int(self._lib.bt_steps_done(self._ctx))
        sampled_idx = [i for i, t in enumerate(time_steps) if t in time_steps_sampled]
        obs, t_obs = self._obs_buffers() if sampled_idx else (None, None)
        d = self._dev
        state_keys = self._state_names(False)
        pending = None

        def finish(t_sample):
            # host side of a sampled step: runs while the device steps on
            h = self._fetch_finish_async(state_keys + list(self._OBS_KEEP))
            self._set_state(h, False)
            self._set_observables(h)
            phase.callbacks.progressed_next()
            self.write2storage(t_sample, phase.cells, phase.p)

        pos = 0
        for s in sampled_idx:
            self._step_observed(s + 1 - pos)
            pos = s + 1
            if pending is not None:
                finish(pending)
                pending = None
            self._sync(step0)
            # observables of the sampled step, then one batch of copies to pinned host
            # memory that runs beside the next steps
            self._check(self._lib.bt_observables(self._ctx, ctypes.byref(obs)))
            self._fetch_start_async({**{k: d[k] for k in state_keys}, **{k: t_obs[k] for k in self._OBS_KEEP}})
            pending = time_steps[s]
        if pos < len(time_steps):
            self._check(self._lib.bt_step(self._ctx, len(time_steps) - pos, 1 if self.use_graph else 0))
        if pending is not None:
            finish(pending)
        self._sync(step0)
        self._pull_state()

    # ------------------------------------------------------------------
    # raw stepping (bench / tests): no sampling, state stays on the device
    # ------------------------------------------------------------------
    def prepare(self, phase: SimPhase) -> None:
        """init_dynamics + device set-up + initial update_V, without running steps."""
        self.cells = phase.cells
        self.init_dynamics(phase)
        self.clear_storage()
        self._device_setup(phase)
        self.update_V()

    def step_n(self, n: int, sync: bool = True) -> None:
        self._check(self._lib.bt_step(self._ctx, int(n), 1 if self.use_graph else 0))
        if sync:
            self._sync()

    def pull(self) -> None:
        self._sync()
        self._pull_state()

    def time_kernel(self, name: str, n_launches: int = 50) -> float:
        """Average device ms of one launch of a step kernel (HIP events)."""
        ms = ctypes.c_double(0.0)
        self._check(self._lib.bt_time_kernel(self._ctx, _native.KERNEL_IDS[name], int(n_launches),
                                             ctypes.byref(ms)))
        return ms.value

    # ------------------------------------------------------------------
    def save(self, path: str) -> None:
        """Phase-end state (stand-in for ``_pickle_phase``, ``sim.py:1888-1950``)."""
        np.savez_compressed(path, cc_cells=self.cc_cells, c_lag=self.c_lag, vm_cell=self.vm_cell,
                            cc_env=self.cc_env, E_env_x=self.E_env_x, E_env_y=self.E_env_y,
                            v_env=self.v_env, names=np.asarray(self.names), Phi_b=self.Phi_b,
                            bound_V=np.asarray([self.bound_V[k] for k in SIDES], dtype=np.float64))

    def load(self, path: str) -> None:
        with np.load(path, allow_pickle=False) as z:
            self.cc_cells = z["cc_cells"]
            self.c_lag = z["c_lag"]
            self.vm_cell = z["vm_cell"]
            self.cc_env = z["cc_env"]
            self.E_env_x = z["E_env_x"]
            self.E_env_y = z["E_env_y"]
            self.v_env = z["v_env"]
            if "Phi_b" in z:
                self.Phi_b = z["Phi_b"]
                self.bound_V = dict(zip(SIDES, (float(v) for v in z["bound_V"])))
