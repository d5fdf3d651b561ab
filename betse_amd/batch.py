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
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import _native
from .cells import Cells
from .channels import CHANNEL_KINDS, draw_channel_state
from .exceptions import BetseNativeError, BetseSimException, BetseSimUnstableException
from .parameters import Parameters
from .sim import HostState, SimPhase, SimPhaseKind, gaussian_taps, make_tiles


class BatchSimulator:
    """A batch of independent clusters on one device."""

    def __init__(self, worlds: Sequence[Tuple[Parameters, Cells]], *, device: Optional[str] = None,
                 flux_mode: str = "fp32", use_graph: bool = True):
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
                             use_graph=self.use_graph)
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
        s0, p0 = self.states[0], self.worlds[0][0]
        for st, (p, _) in zip(self.states, self.worlds):
            if st.names != s0.names:
                raise BetseSimException("clusters of a batch need the same ions")
            if p.dt != p0.dt or p.T != p0.T or p.alpha_Ca != p0.alpha_Ca:
                raise BetseSimException("clusters of a batch need the same dt, T and Ca pump")
            if tuple(st.channel_specs) != tuple(s0.channel_specs):
                raise BetseSimException("clusters of a batch need the same channels")
            if st.c_env_bound != s0.c_env_bound:
                raise BetseSimException("clusters of a batch need the same boundary concentrations")
            if (st.network is None) != (s0.network is None) or (
                    st.network is not None and bytes(st.network.pack(p.R, st.T)) != bytes(s0.network.pack(p0.R, s0.T))):
                raise BetseSimException("clusters of a batch need the same reaction network")
            if (p.alpha_NaK > 0.0) != (p0.alpha_NaK > 0.0):
                raise BetseSimException("the NaK pump must be on in all clusters or in none")

    # ------------------------------------------------------------------
    def _build(self) -> None:
        torch = self.torch
        i32, f64 = torch.int32, torch.float64
        self._destroy()
        sts = self.states
        cells_l = [c for _, c in self.worlds]
        C = [c.n_cells for c in cells_l]
        M = [c.n_mems for c in cells_l]
        G = [c.X.size for c in cells_l]
        coff = np.concatenate([[0], np.cumsum(C)])
        moff = np.concatenate([[0], np.cumsum(M)])
        soff = np.concatenate([[0], np.cumsum(G)])
        self.offsets = (coff, moff, soff)
        nI = len(sts[0].names)
        nch = len(sts[0].channel_specs)

        def cat(f):
            return np.concatenate([f(i, c) for i, c in enumerate(cells_l)])

        mem_ptr = np.concatenate([cells_l[0].mem_ptr[:1]] + [c.mem_ptr[1:] + moff[i] for i, c in enumerate(cells_l)])
        sq_ptr = np.concatenate([cells_l[0].sq_ptr[:1]] + [c.sq_ptr[1:] + moff[i] for i, c in enumerate(cells_l)])
        sq_mems = cat(lambda i, c: c.sq_mems + moff[i])
        pos = np.empty(int(moff[-1]), dtype=np.int64)
        pos[sq_mems] = np.arange(int(moff[-1]))
        tiles, tcl = [np.zeros(1, dtype=np.int64)], []
        for i, c in enumerate(cells_l):
            t = make_tiles(c.mem_ptr)
            tiles.append(t[1:].astype(np.int64) + coff[i])
            tcl.append(np.full(len(t) - 1, i, dtype=np.int32))
        tile_ptr = np.concatenate(tiles)
        tile_cluster = np.concatenate(tcl)
        self.n_tiles = len(tile_cluster)
        sq_cluster = cat(lambda i, c: np.full(c.X.size, i, dtype=np.int32))
        counts = np.diff(sq_ptr)

        host = {
            "mem_ptr": (mem_ptr, i32), "mem_cell": (cat(lambda i, c: c.mem_to_cells + coff[i]), i32),
            "nn_mem": (cat(lambda i, c: c.nn_i + moff[i]), i32),
            "nn_cell": (cat(lambda i, c: c.mem_to_cells[c.nn_i] + coff[i]), i32),
            "mem_sq": (cat(lambda i, c: c.map_mem2ecm + soff[i]), i32),
            "sq_ptr": (sq_ptr, i32), "sq_mems": (sq_mems, i32), "mem_pos": (pos, i32),
            "tile_ptr": (tile_ptr, i32), "tile_cluster": (tile_cluster, i32), "sq_cluster": (sq_cluster, i32),
            "mem_sa": (cat(lambda i, c: c.mem_sa), f64), "cell_vol": (cat(lambda i, c: c.cell_vol), f64),
            "diviterm": (cat(lambda i, c: c.diviterm), f64), "cell_sa": (cat(lambda i, c: c.cell_sa), f64),
            "mem_nx": (cat(lambda i, c: c.mem_vects_flat[:, 2]), f64),
            "mem_ny": (cat(lambda i, c: c.mem_vects_flat[:, 3]), f64),
            "sq_has_mem": ((counts > 0).astype(np.uint8), torch.uint8),
            "gjopen": (np.concatenate([st.gj_block * np.ones(st.mdl) * c.gj_default_weights
                                       for st, c in zip(sts, cells_l)]), f64),
            "Dm": (np.concatenate([(st.Dm_cells if self.kind is SimPhaseKind.INIT
                                    else st.Dm_scheduled + st.Dm_base) for st in sts], axis=1), f64),
            "D_env": (np.concatenate([st.D_env * st.TJ_modulator for st in sts], axis=1), f64),
            "c_cell": (np.concatenate([st.cc_cells for st in sts], axis=1), f64),
            "c_lag": (np.concatenate([st.c_lag for st in sts], axis=1), f64),
            "vm_cell": (np.concatenate([st.vm_cell for st in sts]), f64),
            "c_env": (np.concatenate([st.cc_env for st in sts], axis=1), f64),
            "E_x": (np.concatenate([st.E_env_x.ravel() for st in sts]), f64),
            "E_y": (np.concatenate([st.E_env_y.ravel() for st in sts]), f64),
            "v_env": (np.concatenate([st.v_env for st in sts]), f64),
            "gates": (np.concatenate([st.chan_gates for st in sts], axis=2) if nch else np.zeros((1, 2, 1)), f64),
            "chan_mod": (np.concatenate([st.chan_mod for st in sts], axis=1) if nch else np.ones((1, 1)), f64),
        }
        host["cell_cluster"] = (np.concatenate([np.full(n, i, dtype=np.int32) for i, n in enumerate(C)]), i32)
        net = sts[0].network
        if net is not None:
            host["c_net"] = (np.concatenate([st.c_net for st in sts], axis=1), f64)
        self._dev = {}
        d = self._dev
        with torch.cuda.device(self.device), torch.cuda.stream(self.stream):
            for k, (a, dt) in host.items():
                d[k] = torch.as_tensor(np.ascontiguousarray(a), dtype=dt).to(self.device)
            self.stream.synchronize()

        s0, p0 = sts[0], self.worlds[0][0]
        prm = _native.BtParams()
        prm.n_cells, prm.n_mems, prm.n_squares = int(coff[-1]), int(moff[-1]), int(soff[-1])
        prm.n_ions = nI
        prm.i_na = s0.names.index("Na") if ("Na" in s0.names and "K" in s0.names and p0.alpha_NaK > 0.0) else -1
        prm.i_k = s0.names.index("K") if prm.i_na >= 0 else -1
        prm.i_ca = s0.names.index("Ca") if ("Ca" in s0.names and p0.alpha_Ca > 0.0) else -1
        prm.flux_mode = _native.BT_FLUX_FP32 if self.flux_mode == "fp32" else _native.BT_FLUX_FP64
        active = self.kind is SimPhaseKind.SIM or getattr(p0, "channels_init_active", True)
        prm.n_chan = nch if active else 0
        for k, spec in enumerate(s0.channel_specs):
            prm.chan_kind[k] = _native.BT_CHAN_KINDS[spec.kind]
            prm.chan_ion[k] = s0.names.index(CHANNEL_KINDS[spec.kind][0])
            prm.chan_max_dm[k] = float(spec.max_dm)
        prm.chan_dt = p0.dt * 1.0e3
        for i in range(nI):
            prm.z[i] = float(s0.zs[i])
            prm.d_gj[i] = float(s0.D_gj[i][0])
            prm.c_bound[i] = float(s0.c_env_bound[i])
        prm.dt, prm.T, prm.T_gj = p0.dt, s0.T, p0.T
        prm.F, prm.R, prm.q, prm.kb, prm.cm, prm.tm = p0.F, p0.R, p0.q, p0.kb, p0.cm, p0.tm
        prm.gj_len, prm.gj_surface = cells_l[0].gj_len, p0.gj_surface
        prm.NaK_block, prm.deltaGATP = float(s0.NaKATP_block), p0.deltaGATP
        prm.cATP, prm.cADP, prm.cPi = p0.cATP, p0.cADP, p0.cPi
        prm.KmNK_Na, prm.KmNK_K, prm.KmNK_ATP = p0.KmNK_Na, p0.KmNK_K, p0.KmNK_ATP
        prm.alpha_Ca, prm.KmCa_Ca, prm.KmCa_ATP = p0.alpha_Ca, p0.KmCa_Ca, p0.KmCa_ATP
        prm.cell_height, prm.eo, prm.er = p0.cell_height, p0.eo, p0.er
        prm.cell_radius, prm.true_cell_size = p0.cell_radius, p0.true_cell_size
        for j, w in enumerate(gaussian_taps(1.0)):
            prm.gauss_w[j] = float(w)
        prm.smooth_cells, prm.mu = p0.smooth_cells, p0.mu

        cls = (_native.BtCluster * len(sts))()
        for i, (st, (p, c)) in enumerate(zip(sts, self.worlds)):
            if c.gj_len != cells_l[0].gj_len:
                raise BetseSimException("clusters of a batch need the same gap-junction length")
            cls[i].grid_ny, cls[i].grid_nx = c.X.shape
            cls[i].sq_off = int(soff[i])
            cls[i].delta = c.delta
            cls[i].mean_sa = float(c.memSa_per_envSquare[c.map_mem2ecm].mean())
            cls[i].ko_env = float(st.ko_env)
            cls[i].alpha_NaK = p.alpha_NaK
        buf = _native.BtBuffers()
        for name, _ in _native.BtBuffers._fields_[:-2]:
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
            self._net = net.pack(p0.R, s0.T)
            self._check(self._lib.bt_set_network(self._ctx, ctypes.byref(self._net),
                                                 ctypes.c_void_p(d["c_net"].data_ptr())))
        self._check(self._lib.bt_update_v(self._ctx))

    # ------------------------------------------------------------------
    def _check(self, rc: int) -> None:
        if rc != _native.BT_OK:
            raise BetseNativeError(f"libbetse_amd error {rc}: {self._lib.bt_last_error(self._ctx).decode()}")

    def _destroy(self) -> None:
        if self._ctx:
            self._lib.bt_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

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
        keys = ("c_cell", "c_lag", "vm_cell", "c_env", "E_x", "E_y", "v_env", "gates") + (("c_net",) if "c_net" in d else ())
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
            if len(st.channel_specs):
                st.chan_gates = h["gates"][:, :, moff[i]:moff[i + 1]].copy()
            if "c_net" in h:
                st.c_net = h["c_net"][:, a:b].copy()

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
