"""Learned warm start of the init phase.

The reference runs its init phase for a fixed ``init_tsteps``
(``parameters.py:525``) from the ion profile's default concentrations
(``sim.py:452-754``) and has no convergence criterion. This module adds one,
and a graph network that predicts where the init phase ends up:

* **Residual.** After a window of ``W`` steps the per-cluster residual is
  ``r = max over cells and ions of |c_n - c_{n-W}| / (W |c_n|)``, the mean
  per-step relative change of the intracellular concentrations
  (``BatchSimulator.residual_window``). A cluster has reached its init steady
  state when ``r < tol`` (``tol = 1e-6``).
* **Prediction.** ``WarmStartGNN`` (message passing on the cell-adjacency
  graph ``Cells.cell_nn``, built like ``cells.py:1536-1564``) maps per-cell
  geometry and the cluster's parameters to ``log(c/c0)`` of Na, K, Cl (and Ca
  when the profile has it) and to Vmem.
* **Injection.** Vmem is a function of the charge: ``vm = F sum(z c)
  diviterm / cm`` (``sim.py:2007-2083``). Predicted concentrations and a
  predicted Vmem are made consistent by solving for the charge-balancing
  anion M, the ion the reference itself balances with
  (``parameters.py:1569-1605``). The lagged membrane concentrations start
  equal to the cell values, as after ``init_core`` (``sim.py:744-745``);
  voltage-gated channel gates start at their steady state for the predicted
  Vmem (the channels' ``_init_state``).
* **Fallback.** A prediction that is not finite, gives a non-positive M or
  an implausible Vmem is not injected: that cluster starts cold. A batch
  that turns unstable restarts with the clusters that went non-finite
  started cold. Either way the plain time-stepper runs from the start state
  to its own residual, so the network can only change how long the init
  phase takes, not the criterion it ends on.
"""

from __future__ import annotations

import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from .batch import BatchSimulator, batched_worlds
from .channels import init_gates
from .exceptions import BetseSimUnstableException
from .parameters import Parameters

CHANNEL_FEATURES = ("Nav1p2", "Kv1p1", "Cav1p2")
N_FEATURES = 10 + len(CHANNEL_FEATURES)
V_SCALE = 0.05        # Vmem target in units of 50 mV


def target_ions(names: Sequence[str]) -> List[str]:
    return [n for n in ("Na", "K", "Cl", "Ca") if n in names]


def cell_features(p: Parameters, cells, base: Parameters) -> np.ndarray:
    """Per-cell inputs: geometry, position in the cluster, the cluster's
    membrane diffusivities and pump rate relative to ``base``, and the cell's
    voltage-gated channel density multipliers (log; 0 without the channel).
    The cluster-size input saturates at 10^3.5 cells."""
    n = cells.n_cells
    X = np.zeros((n, N_FEATURES))
    vol = cells.cell_vol
    X[:, 0] = vol / np.median(vol) - 1.0
    X[:, 1] = cells.num_mems / 6.0 - 1.0
    bnd = np.zeros(n)
    bnd[np.asarray(cells.bflags_cells, dtype=np.int64)] = 1.0
    X[:, 2] = bnd
    ctr = cells.cell_centres - cells.cell_centres.mean(axis=0)
    rad = np.sqrt((ctr ** 2).sum(axis=1))
    X[:, 3] = rad / max(rad.max(), 1e-30)
    for j, ion in enumerate(("Na", "K", "Cl")):
        X[:, 4 + j] = np.log10(p.mem_perms[ion] / base.mem_perms[ion])
    X[:, 7] = np.log10(p.alpha_NaK / base.alpha_NaK)
    X[:, 8] = min(np.log10(n), 3.5) / 4.0 - 0.75
    X[:, 9] = cells.diviterm / cells.diviterm.mean() - 1.0
    mults = getattr(p, "channel_cell_mult", None)
    for k, spec in enumerate(getattr(p, "channels", ())):
        if spec.kind in CHANNEL_FEATURES and mults is not None and mults[k] is not None:
            X[:, 10 + CHANNEL_FEATURES.index(spec.kind)] = np.log(mults[k])
    return X


def state_targets(st, names: Sequence[str]) -> np.ndarray:
    """Per-cell targets of a cluster state: log(c/c0) per target ion, Vmem/50 mV."""
    ions = target_ions(names)
    Y = np.zeros((st.cdl, len(ions) + 1))
    for j, ion in enumerate(ions):
        i = names.index(ion)
        Y[:, j] = np.log(np.maximum(st.cc_cells[i], 1e-300) / st.p.cell_concs[ion])
    Y[:, -1] = st.vm_cell / V_SCALE
    return Y


def inject(st, Y: np.ndarray, vm_limit: float = 0.3) -> bool:
    """Set a cold-initialised cluster state to the prediction ``Y`` (rows per
    cell, columns as ``state_targets``); M is solved so that the charge gives
    the predicted Vmem. Returns False (state untouched) if the prediction is
    not usable."""
    names = st.names
    if "M" not in names or not np.all(np.isfinite(Y)):
        return False
    p, cells = st.p, st.cells
    ions = target_ions(names)
    vm = Y[:, -1] * V_SCALE
    if np.any(np.abs(vm) > vm_limit):
        return False
    cc = st.cc_cells.copy()
    for j, ion in enumerate(ions):
        cc[names.index(ion)] = p.cell_concs[ion] * np.exp(np.clip(Y[:, j], -30.0, 30.0))
    iM = names.index("M")
    zs = np.asarray(st.zs, dtype=np.float64)
    # sum(z c) = vm cm / (F diviterm)
    target = vm * p.cm / (p.F * cells.diviterm)
    others = zs @ cc - zs[iM] * cc[iM]
    cM = (target - others) / zs[iM]
    if not np.all(np.isfinite(cc)) or np.any(cM <= 0.0):
        return False
    cc[iM] = cM
    st.cc_cells = cc
    st.c_lag = cc.copy()
    # channel gates at their steady state for the predicted Vmem (the gates'
    # _init_state, channels/vg_*.py), which is where they relax to
    for k, spec in enumerate(getattr(st, "channel_specs", ())):
        m, h = init_gates(spec.kind, vm[cells.mem_to_cells])
        st.chan_gates[k, 0] = m
        st.chan_gates[k, 1] = h
    return True


# ----------------------------------------------------------------------
# running the init phase to its residual
# ----------------------------------------------------------------------
@dataclass
class InitResult:
    steps: np.ndarray             # steps each cluster needed (-1: not within max_steps)
    residual: np.ndarray          # last residual per cluster
    seconds: float                # wall seconds of the stepping loop (all clusters converged or capped)
    seconds_to_converge: np.ndarray   # wall seconds at which each cluster converged (nan: never)
    warm: np.ndarray              # cluster started from a prediction
    restarts: int = 0


def run_to_residual(bs: BatchSimulator, tol: float = 1e-6, window: int = 50, max_steps: int = 200_000,
                    log=None, run_all: bool = False, compact: bool = True) -> InitResult:
    """Step the batch in windows until every cluster's residual is below
    ``tol`` (or ``max_steps``); returns per-cluster step counts and the wall
    seconds at which each cluster got there. A converged cluster's init phase
    is over: with ``compact`` the batch is rebuilt without the converged
    clusters whenever they are at least half of it (their states are kept in
    ``bs.states``). With ``run_all`` every cluster runs ``max_steps``."""
    n_cl = len(bs.states)
    steps = np.full(n_cl, -1, dtype=np.int64)
    secs = np.full(n_cl, np.nan)
    last = np.full(n_cl, np.inf)
    active = np.arange(n_cl)            # clusters of `cur`, in order
    cur = bs
    n = 0
    t0 = time.perf_counter()
    next_log = t0 + 30.0
    while n < max_steps:
        r = cur.residual_window(window)
        n += window
        now = time.perf_counter() - t0
        last[active] = r
        new = (r < tol) & (steps[active] < 0)
        steps[active[new]] = n
        secs[active[new]] = now
        if np.all(steps >= 0) and not run_all:
            break
        if compact and not run_all:
            live = np.nonzero(steps[active] < 0)[0]
            if 0 < len(live) <= len(active) // 2:
                nxt = cur.subset(live)
                if cur is not bs:
                    del cur
                cur = nxt
                active = active[live]
        if log is not None and time.perf_counter() > next_log:
            next_log = time.perf_counter() + 30.0
            log(f"  step {n}: {int((steps >= 0).sum())}/{n_cl} converged, {len(active)} in the batch, "
                f"max r {np.max(r):.3e}")
    cur.pull()                          # every cluster's final state is now in bs.states
    return InitResult(steps, last, time.perf_counter() - t0, secs, np.zeros(n_cl, dtype=bool))


class WarmStarter:
    """A trained network plus the feature/target normalisation it was trained with."""

    def __init__(self, net, x_mean, x_std, y_mean, y_std, base: Parameters, names: Sequence[str]):
        self.net = net
        self.x_mean, self.x_std = x_mean, x_std
        self.y_mean, self.y_std = y_mean, y_std
        self.base = base
        self.names = list(names)

    def predict(self, worlds) -> List[np.ndarray]:
        """Per-cluster prediction (rows per cell, columns as ``state_targets``)."""
        torch = self.net.torch
        X, ptr, idx, coff = graph_batch([(p, c) for p, c in worlds], self.base)
        Xn = ((X - self.x_mean) / self.x_std).astype(np.float32)
        dev = self.net.device
        Xt = torch.tensor(Xn, device=dev)
        pt = torch.tensor(ptr.astype(np.int32), device=dev)
        it = torch.tensor(idx.astype(np.int32), device=dev)
        y = self.net.forward(Xt, pt, it, keep=False)
        self.net.sync()
        Y = y.cpu().numpy().astype(np.float64) * self.y_std + self.y_mean
        return [Y[coff[i]:coff[i + 1]] for i in range(len(worlds))]

    def save(self, path: str) -> None:
        d = {"p_" + k: v for k, v in self.net.params_np().items()}
        d.update(x_mean=self.x_mean, x_std=self.x_std, y_mean=self.y_mean, y_std=self.y_std,
                 shape=np.array([self.net.f_in, self.net.f_out, self.net.H, self.net.L]))
        np.savez(path, **d)


def graph_batch(worlds, base: Parameters):
    """Concatenated features and neighbour CSR of several clusters."""
    Xs, ptrs, idxs = [], [np.zeros(1, dtype=np.int64)], []
    coff = [0]
    eoff = 0
    for p, cells in worlds:
        Xs.append(cell_features(p, cells, base))
        ptrs.append(cells.cell_nn_ptr[1:] + eoff)
        idxs.append(cells.cell_nn_idx + coff[-1])
        eoff += int(cells.cell_nn_ptr[-1])
        coff.append(coff[-1] + cells.n_cells)
    return np.concatenate(Xs), np.concatenate(ptrs), np.concatenate(idxs), np.asarray(coff)


def warm_start_batch(bs: BatchSimulator, starter: Optional[WarmStarter], tol: float = 1e-6, window: int = 50,
                     max_steps: int = 200_000, log=None) -> Tuple[InitResult, float]:
    """Cold-initialise ``bs`` (``init``), inject predictions where usable, run to
    the residual. Returns the result and the seconds spent predicting and
    injecting (part of the init time)."""
    t0 = time.perf_counter()
    bs.init("init")
    warm = np.zeros(len(bs.states), dtype=bool)
    if starter is not None:
        preds = starter.predict(bs.worlds)
        for i, (st, Y) in enumerate(zip(bs.states, preds)):
            warm[i] = inject(st, Y)
        bs.rebuild()
    t_pred = time.perf_counter() - t0
    restarts = 0
    while True:
        try:
            res = run_to_residual(bs, tol, window, max_steps, log)
            break
        except BetseSimUnstableException:
            # clusters that went non-finite start cold; the rest from their prediction again
            bad = np.array([not (np.all(np.isfinite(st.cc_cells)) and np.all(np.isfinite(st.vm_cell)))
                            for st in bs.states])
            restarts += 1
            if restarts > 2 or starter is None:
                raise
            bs.init("init")
            preds = starter.predict(bs.worlds)
            for i, (st, Y) in enumerate(zip(bs.states, preds)):
                if warm[i] and not bad[i]:
                    warm[i] = inject(st, Y)
                else:
                    warm[i] = False
            bs.rebuild()
    res.warm = warm
    res.restarts = restarts
    return res, t_pred


# ----------------------------------------------------------------------
# training data and training
# ----------------------------------------------------------------------
def _extrapolate(snaps, coff):
    """One vector extrapolation step per cluster from three equally spaced
    snapshots (c, vm): with d1 = x2 - x1, d2 = x3 - x2 and the cluster's
    contraction ratio q = <d1, d2>/<d1, d1> (relative changes over all cells
    and ions), the limit of x_n = x* + A q^n is x3 + d2 q/(1 - q). Clusters
    without a contracting ratio keep x3."""
    (c1, v1), (c2, v2), (c3, v3) = snaps
    cx, vx = c3.copy(), v3.copy()
    qs = np.zeros(len(coff) - 1)
    for i in range(len(coff) - 1):
        a, b = coff[i], coff[i + 1]
        den = np.where(c3[:, a:b] != 0, np.abs(c3[:, a:b]), 1.0)
        d1 = (c2[:, a:b] - c1[:, a:b]) / den
        d2 = (c3[:, a:b] - c2[:, a:b]) / den
        q = float(np.sum(d1 * d2) / max(np.sum(d1 * d1), 1e-300))
        qs[i] = q
        if 0.0 < q < 0.999:
            g = q / (1.0 - q)
            cx[:, a:b] = c3[:, a:b] + (c3[:, a:b] - c2[:, a:b]) * g
            vx[a:b] = v3[a:b] + (v3[a:b] - v2[a:b]) * g
    return cx, vx, qs


def _targets_from(st, c, vm) -> np.ndarray:
    """state_targets of a cluster state given as arrays."""
    names = st.names
    ions = target_ions(names)
    Y = np.zeros((c.shape[1], len(ions) + 1))
    for j, ion in enumerate(ions):
        Y[:, j] = np.log(np.maximum(c[names.index(ion)], 1e-300) / st.p.cell_concs[ion])
    Y[:, -1] = vm / V_SCALE
    return Y


def _run_snapshots(bs: BatchSimulator, lead: int, gap: int):
    """Run ``lead`` steps, then take three snapshots ``gap`` steps apart."""
    if lead > 0:
        bs.step_n(lead)
    snaps = [bs.snapshot()]
    for _ in range(2):
        bs.step_n(gap)
        snaps.append(bs.snapshot())
    return snaps


def steady_state_labels(bs: BatchSimulator, horizon: int = 20_000, rounds: int = 2, tol: float = 1e-6,
                        window: int = 100, log=None):
    """Init-phase steady states of every cluster of ``bs``: the plain stepper
    runs ``horizon`` steps from the default start; its trajectory is
    extrapolated to its limit (``_extrapolate``); each further round restarts
    the stepper from the extrapolated state (charge made consistent through M,
    as ``inject``) and extrapolates again over ``horizon/2`` steps. A label is
    accepted only if the plain stepper, started from it, is below ``tol`` on
    its own residual. Returns (targets per cluster, accepted flags, residuals,
    contraction ratios, stepper steps spent)."""
    gap = max(horizon // 4, window)
    bs.init("init")
    coff = bs.offsets[0]
    snaps = _run_snapshots(bs, horizon - 2 * gap, gap)
    steps = horizon
    cx, vx, qs = _extrapolate(snaps, coff)
    Y = [_targets_from(st, cx[:, coff[i]:coff[i + 1]], vx[coff[i]:coff[i + 1]]) for i, st in enumerate(bs.states)]
    for rnd in range(1, rounds):
        bs.init("init")
        usable = np.array([inject(st, y) for st, y in zip(bs.states, Y)])
        bs.rebuild()
        snaps = _run_snapshots(bs, 0, gap)
        steps += 2 * gap
        cx, vx, qs2 = _extrapolate(snaps, coff)
        for i, st in enumerate(bs.states):
            if usable[i]:
                Y[i] = _targets_from(st, cx[:, coff[i]:coff[i + 1]], vx[coff[i]:coff[i + 1]])
                qs[i] = qs2[i]
        if log is not None:
            log(f"  extrapolation round {rnd + 1}: {int(usable.sum())}/{len(usable)} restarted from the limit")
    # verification by the plain stepper itself
    bs.init("init")
    usable = np.array([inject(st, y) for st, y in zip(bs.states, Y)])
    bs.rebuild()
    bs.residual_window(window)                       # let the injected state settle one window
    r = bs.residual_window(window)
    steps += 2 * window
    ok = usable & (r < tol)
    return Y, ok, r, qs, steps


def make_dataset(n_clusters: int, seed: int, *, cells_range=(500, 2000), grid_range=(48, 88),
                 decades: float = 2.0, batch: int = 256, tol: float = 1e-6, window: int = 100,
                 horizon: int = 20_000, rounds: int = 2, flux_mode: str = "fp64", device=None, log=None,
                 base: str = "small"):
    """Init-phase steady states of ``n_clusters`` random clusters
    (``steady_state_labels``), produced on the GPU in batches. Returns
    per-cluster (features, neighbour CSR, targets, ok) and run statistics."""
    base_name = base
    base = Parameters.preset(base_name)
    data = []
    stats = []
    for b0 in range(0, n_clusters, batch):
        nb = min(batch, n_clusters - b0)
        t0 = time.perf_counter()
        worlds = batched_worlds(nb, cells_range=cells_range, grid_range=grid_range, seed=seed + b0,
                                base=base_name, decade=decades)
        t_w = time.perf_counter() - t0
        bs = BatchSimulator(worlds, device=device, flux_mode=flux_mode)
        t1 = time.perf_counter()
        Y, ok, r, qs, steps = steady_state_labels(bs, horizon, rounds, tol, window, log)
        t_l = time.perf_counter() - t1
        for i, (p, cells) in enumerate(worlds):
            data.append(dict(X=cell_features(p, cells, base), ptr=cells.cell_nn_ptr.copy(),
                             idx=cells.cell_nn_idx.copy(), Y=Y[i], ok=bool(ok[i]), residual=float(r[i]),
                             q=float(qs[i]), names=bs.states[i].names))
        stats.append(dict(clusters=nb, cells=int(sum(c.n_cells for _, c in worlds)), world_seconds=t_w,
                          label_seconds=t_l, stepper_steps=int(steps), accepted=int(ok.sum()),
                          median_residual=float(np.median(r))))
        if log is not None:
            log(f"dataset batch {b0 // batch}: {stats[-1]}")
        del bs
    return data, stats, base


def train(data, *, hidden: int = 64, n_layers: int = 6, epochs: int = 30, clusters_per_batch: int = 32,
          lr: float = 2e-3, seed: int = 0, device=None, base: Optional[Parameters] = None, log=None):
    """Train ``WarmStartGNN`` on the accepted labels of ``data`` (GPU kernels
    only: BF16 MFMA products, FP32 accumulation, Adam)."""
    from .gnn import WarmStartGNN
    good = [d for d in data if d["ok"]]
    if not good:
        raise ValueError("no accepted label to train on")
    X = np.concatenate([d["X"] for d in good])
    Y = np.concatenate([d["Y"] for d in good])
    x_mean, x_std = X.mean(axis=0), X.std(axis=0) + 1e-6
    y_mean, y_std = Y.mean(axis=0), Y.std(axis=0) + 1e-6
    net = WarmStartGNN(X.shape[1], Y.shape[1], hidden=hidden, n_layers=n_layers, device=device, seed=seed)
    torch = net.torch
    dev = net.device
    rng = np.random.RandomState(seed)
    # fixed minibatches of whole clusters, uploaded once
    mbs = []
    order = rng.permutation(len(good))
    for k in range(0, len(order), clusters_per_batch):
        sel = [good[j] for j in order[k:k + clusters_per_batch]]
        Xb = np.concatenate([(d["X"] - x_mean) / x_std for d in sel]).astype(np.float32)
        Yb = np.concatenate([(d["Y"] - y_mean) / y_std for d in sel]).astype(np.float32)
        ptr, idx, off, eoff = [np.zeros(1, dtype=np.int64)], [], 0, 0
        for d in sel:
            ptr.append(d["ptr"][1:] + eoff)
            idx.append(d["idx"] + off)
            eoff += int(d["ptr"][-1])
            off += len(d["X"])
        with torch.cuda.device(dev), torch.cuda.stream(net.stream):
            mbs.append((torch.tensor(Xb, device=dev), torch.tensor(Yb, device=dev),
                        torch.ones(Yb.shape, dtype=torch.float32, device=dev),
                        torch.tensor(np.concatenate(ptr).astype(np.int32), device=dev),
                        torch.tensor(np.concatenate(idx).astype(np.int32), device=dev), float(Yb.size)))
    history = []
    t0 = time.perf_counter()
    n_steps = epochs * len(mbs)
    step = 0
    for ep in range(epochs):
        tot = 0.0
        for j in rng.permutation(len(mbs)):
            Xt, Yt, Mt, pt, it, nv = mbs[j]
            # cosine learning-rate decay
            lr_t = 0.5 * lr * (1.0 + np.cos(np.pi * step / max(n_steps, 1)))
            y = net.forward(Xt, pt, it)
            loss, dY = net.loss(y, Yt, Mt, nv)
            net.adam(net.backward(dY, pt, it), lr=lr_t)
            tot += loss
            step += 1
        history.append(tot / len(mbs))
        if log is not None and (ep % 5 == 0 or ep == epochs - 1):
            log(f"epoch {ep}: loss {history[-1]:.4e} ({time.perf_counter() - t0:.1f} s)")
    return WarmStarter(net, x_mean, x_std, y_mean, y_std, base or Parameters.preset("small"),
                       good[0]["names"]), history, time.perf_counter() - t0


def evaluate(starter: WarmStarter, n_clusters: int, seed: int, *, cells_range=(500, 2000), grid_range=(48, 88),
             decades: float = 2.0, tol: float = 1e-6, window: int = 50, max_steps: int = 200_000,
             flux_mode: str = "fp64", device=None, log=None) -> Dict[str, object]:
    """Init time to ``tol`` on held-out clusters: from the default start and
    from the prediction, same stepper, same batch (compacted as clusters
    converge). Seconds include the set-up of each start (cold: the default
    state; warm: the default state, the prediction and the injection)."""
    worlds = batched_worlds(n_clusters, cells_range=cells_range, grid_range=grid_range, seed=seed,
                            base="small", decade=decades)
    bs = BatchSimulator(worlds, device=device, flux_mode=flux_mode)
    t0 = time.perf_counter()
    bs.init("init")
    t_setup_cold = time.perf_counter() - t0
    cold = run_to_residual(bs, tol, window, max_steps, log)
    Y_cold = [state_targets(st, st.names) for st in bs.states]
    warm, t_pred = warm_start_batch(bs, starter, tol, window, max_steps, log)
    Y_warm = [state_targets(st, st.names) for st in bs.states]
    preds = starter.predict(worlds)
    err_pred = np.array([np.abs(P - Yc).max(axis=0) for P, Yc in zip(preds, Y_cold)])
    both = (cold.steps >= 0) & (warm.steps >= 0)
    return dict(
        clusters=n_clusters, cells=int(bs.n_cells), tol=tol, window=window, max_steps=max_steps,
        cold_seconds=cold.seconds + t_setup_cold, warm_seconds=warm.seconds + t_pred, predict_seconds=t_pred,
        setup_seconds=t_setup_cold,
        cold_converged=int((cold.steps >= 0).sum()), warm_converged=int((warm.steps >= 0).sum()),
        warm_injected=int(warm.warm.sum()), restarts=warm.restarts,
        cold_median_steps=float(np.median(np.where(cold.steps >= 0, cold.steps, max_steps))),
        warm_median_steps=float(np.median(np.where(warm.steps >= 0, warm.steps, max_steps))),
        cold_max_steps=int(np.where(cold.steps >= 0, cold.steps, max_steps).max()),
        cold_total_steps=int(np.where(cold.steps >= 0, cold.steps, max_steps).sum()),
        warm_total_steps=int(np.where(warm.steps >= 0, warm.steps, max_steps).sum()),
        cold_steps=cold.steps.tolist(), warm_steps=warm.steps.tolist(),
        warm_max_steps=int(np.where(warm.steps >= 0, warm.steps, max_steps).max()),
        speedup_median=float(np.median(cold.steps[both] / np.maximum(warm.steps[both], 1))) if both.any() else 0.0,
        pred_err_logc_max=err_pred[:, :-1].max(axis=0).tolist() if len(err_pred) else [],
        pred_err_vm_mV_median=float(np.median(err_pred[:, -1]) * V_SCALE * 1e3) if len(err_pred) else 0.0,
        warm_vs_cold_vm_mV_max=float(max(np.abs(a[:, -1] - b[:, -1]).max() for a, b in zip(Y_warm, Y_cold))
                                     * V_SCALE * 1e3),
    )
