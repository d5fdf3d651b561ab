"""FP64 CPU oracle for BETSE's electrodiffusion step -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg
may import this module, and only as the checker / CPU baseline. The product
path (``betse_amd``) never imports it.

This is a NumPy restatement, operation by operation, of
``Simulator._run_sim_core_loop`` (``betse/science/sim.py:1132-1390``) and of
everything it calls on the full-solver, ECM, non-polarizable path:

==========================  ===============================================
oracle function             reference
==========================  ===============================================
``electroflux``             ``sim_toolbox.py:18-69`` (Q5: in-place +1e-25 on vBA)
``pumpNaKATP``              ``sim_toolbox.py:71-122``
``pumpCaATP``               ``sim_toolbox.py:124-182`` (in-place no_negs on inputs)
``no_negs``                 ``sim_toolbox.py:475-503``
``check_v``                 ``sim_toolbox.py:332-344``
``nernst_planck_flux``      ``sim_toolbox.py:385-413``
``fd_gradient``             ``math/finitediff.py:1236-1266``
``fd_diff``                 ``math/finitediff.py:1268-1303`` (Q4 sign flip)
``fd_divergence``           ``math/finitediff.py:1305-1311``
``fd_integrator``           ``math/finitediff.py:1479-1512``
``make_laplacian_value``    ``math/finitediff.py:252-431`` ('value' bounds)
``gaussian_filter_const``   ``scipy.ndimage.gaussian_filter`` as called at ``physics/ion_current.py:104``
``OracleSim.init_core``     ``sim.py:452-754``
``OracleSim.init_dynamics`` ``sim.py:758-1012`` (+ ``tishandler.init_profiles`` ``:363-443``)
``OracleSim.initDenv``      ``sim.py:2347-2400``
``OracleSim.run_sim_core``  ``sim.py:1016-1128``
``OracleSim.step``          ``sim.py:1169-1390`` (one loop iteration)
``OracleSim.update_gj``     ``sim.py:2162-2206``
``OracleSim.update_ecm``    ``sim.py:2209-2254``
``OracleSim.update_intra``  ``sim.py:2257-2310``
``OracleSim.ca_handler``    ``sim.py:2126-2159``
``OracleSim.update_all_concs`` ``sim.py:2086-2111`` + ``update_Co`` ``sim_toolbox.py:1155-1207`` + ``div_env`` ``:1209-1234``
``OracleSim.update_V``      ``sim.py:2007-2083``
``OracleSim.get_current``   ``physics/ion_current.py:13-114``
``HH_Decomp``               ``sim_toolbox.py:1236-1290``
``sampled_steps``           ``sim.py:2455-2462`` with ``phase/phasecls.py:330``
``OracleChannel``           ``channels/channelsabc.py:40-60`` (update_mh), ``channels/vg_na.py:34-113,131-186``
                            (Nav1p2), ``channels/vg_k.py:34-115,133-178`` (Kv1p1),
                            ``channels/vg_ca.py:60-113,292-340`` (Cav1p2)
``OracleSim.run_loop_channels`` ``chemistry/networks.py:3115-3215`` (channels on all membranes)
reaction network            ``oracle/network_oracle.py`` (``chemistry/networks.py:2805-2982``),
                            called where ``sim.py:1286-1300`` runs the molecules network
==========================  ===============================================

Two operator forms are supported:

* ``mode="dense"`` keeps the reference's dense operators: ``M_sum_mems`` (C x M),
  ``M_divmap_mem2ecm`` (G^2 x M) and ``lapENVinv = pinv(A)`` (G^2 x G^2), and
  evaluates every observable each step, as the reference does. This is the
  reference's own CPU cost and is feasible only for small clusters.
* ``mode="sparse"`` replaces them by contiguous segment sums, a scatter-add
  by ``map_mem2ecm`` and a direct sparse Poisson solve; observables are
  evaluated only on sampled steps (they never feed the state).

Parity pins (see ``tests/test_oracle_*.py``): closed forms of the GHK and pump
fluxes (Fick limit, antisymmetry, equilibrium), golden vectors of the
finite-difference stencils written by ``tests/golden/make_golden_vectors.py``,
``scipy.ndimage.gaussian_filter`` itself, dense == sparse operators, exact
mass conservation of membrane/environment exchange, the electroneutral
start (Vmem == 0), and the residual of the Poisson solve.
"""

from __future__ import annotations

import math
from typing import Callable, Dict, List, Optional

import numpy as np
import scipy.ndimage
import scipy.sparse
import scipy.sparse.linalg

FLOAT_NONCE = 1.0e-25            # sim_toolbox.py:50


class OracleUnstable(RuntimeError):
    """Raised where the reference raises ``BetseSimUnstableException``."""


# --------------------------------------------------------------------------
# flux laws
# --------------------------------------------------------------------------
def electroflux(cA, cB, Dc, d, zc, vBA, T, p, rho=1):
    """GHK flux; mutates ``vBA`` in place like the reference (Q5)."""
    vBA += FLOAT_NONCE
    zc = zc + FLOAT_NONCE
    alpha = (zc * vBA * p.F) / (p.R * T)
    exp_alpha = np.exp(-alpha)
    deno = -np.expm1(-alpha)
    flux = -((Dc * alpha) / d) * ((cB - cA * exp_alpha) / deno) * rho
    return flux


def pumpNaKATP(cNai, cNao, cKi, cKo, Vm, T, p, block):
    deltaGATP_o = p.deltaGATP
    cATP, cADP, cPi = p.cATP, p.cADP, p.cPi
    Qnumo = (cADP * 1e-3) * (cPi * 1e-3) * ((cNao * 1e-3) ** 3) * ((cKi * 1e-3) ** 2)
    Qdenomo = (cATP * 1e-3) * ((cNai * 1e-3) ** 3) * ((cKo * 1e-3) ** 2)
    Qdenomo = np.where(Qdenomo == 0.0, 1.0e-15, Qdenomo)
    Q = Qnumo / Qdenomo
    Keq = np.exp(-(deltaGATP_o / (p.R * T) - ((p.F * Vm) / (p.R * T))))
    numo_E = ((cNai / p.KmNK_Na) ** 3) * ((cKo / p.KmNK_K) ** 2) * (cATP / p.KmNK_ATP)
    denomo_E = (1 + (cNai / p.KmNK_Na) ** 3) * (1 + (cKo / p.KmNK_K) ** 2) * (1 + (cATP / p.KmNK_ATP))
    fwd_co = numo_E / denomo_E
    f_Na = -3 * block * p.alpha_NaK * fwd_co * (1 - (Q / Keq))
    f_K = -(2 / 3) * f_Na
    return f_Na, f_K, -f_Na


def no_negs(data):
    """Clamp negatives to 0 in place; NaN raises (sim_toolbox.py:475-503)."""
    inds_nan = np.isnan(data).nonzero()
    inds_neg = (data < 0.0).nonzero()
    if len(inds_neg[0]) > 0:
        data[inds_neg] = 0.0
    if len(inds_nan[0]) > 0:
        raise OracleUnstable("NaN in concentrations")
    return data


def pumpCaATP(cCai, cCao, Vm, T, p, block):
    deltaGATP_o = p.deltaGATP
    no_negs(cCai)
    no_negs(cCao)
    cATP, cADP, cPi = p.cATP, p.cADP, p.cPi
    Qnumo = cADP * cPi * cCao
    Qdenomo = cATP * cCai
    Qdenomo = np.where(Qdenomo == 0.0, 1.0e-16, Qdenomo)
    Q = Qnumo / Qdenomo
    Keq = np.exp(-(deltaGATP_o / (p.R * T) - 2 * ((p.F * Vm) / (p.R * T))))
    numo_E = (cCai / p.KmCa_Ca) * (cATP / p.KmCa_ATP)
    denomo_E = (1 + (cCai / p.KmCa_Ca)) * (1 + (cATP / p.KmCa_ATP))
    frwd = numo_E / denomo_E
    f_Ca = -p.alpha_Ca * frwd * (1 - (Q / Keq))
    return f_Ca


def check_v(vm):
    if np.isnan(vm).any():
        raise OracleUnstable("NaN in Vmem")


def nernst_planck_flux(c, gcx, gcy, gvx, gvy, ux, uy, D, z, T, p, mu=0.0):
    alpha = (D * z * p.q) / (p.kb * T)
    fx = -D * gcx - alpha * gvx * c + ux * c - mu * c * gvx
    fy = -D * gcy - alpha * gvy * c + uy * c - mu * c * gvy
    return fx, fy


# --------------------------------------------------------------------------
# finite differences on the environment grid (rows = y = axis 0)
# --------------------------------------------------------------------------
def fd_gradient(F, delx, dely=None):
    if dely is None:
        dely = delx
    dFx = np.zeros(F.shape)
    dFy = np.zeros(F.shape)
    dFx[:, 1:-1] = -(F[:, :-2] - F[:, 2:]) / (2 * delx)
    dFy[1:-1, :] = -(F[:-2, :] - F[2:, :]) / (2 * dely)
    dFx[:, 0] = (F[:, 1] - F[:, 0]) / delx
    dFx[:, -1] = (F[:, -1] - F[:, -2]) / delx
    dFy[0, :] = (F[1, :] - F[0, :]) / dely
    dFy[-1, :] = (F[-1, :] - F[-2, :]) / dely
    return dFx, dFy


def fd_diff(F, delx, axis=0):
    dF = np.zeros(F.shape)
    if axis == 1:
        dF[1:-1, :] = -(F[:-2, :] - F[2:, :]) / (2 * delx)
        dF[0, :] = -(F[1, :] - F[0, :]) / delx
        dF[-1, :] = -(F[-1, :] - F[-2, :]) / delx
    elif axis == 0:
        dF[:, 1:-1] = -(F[:, :-2] - F[:, 2:]) / (2 * delx)
        dF[:, 0] = (F[:, 0] - F[:, 1]) / delx
        dF[:, -1] = (F[:, -2] - F[:, -1]) / delx
    return dF


def fd_divergence(Fx, Fy, delx, dely):
    return fd_diff(Fx, delx, axis=0) + fd_diff(Fy, dely, axis=1)


def fd_integrator(P, sharp=0.5):
    F = np.zeros(P.shape)
    sides = (1 - sharp) / 4
    F[:, :] = sharp * P
    F[0:-1, :] += sides * P[1:, :]
    F[1:, :] += sides * P[0:-1, :]
    F[:, 0:-1] += sides * P[:, 1:]
    F[:, 1:] += sides * P[:, 0:-1]
    F[:, 0] = P[:, 0]
    F[:, -1] = P[:, -1]
    F[0, :] = P[0, :]
    F[-1, :] = P[-1, :]
    return F


def gaussian_filter_const(img, sigma=1.0):
    """The call at ``ion_current.py:104``: sigma 1, mode 'constant', cval 0."""
    return scipy.ndimage.gaussian_filter(img, sigma, mode="constant", cval=0.0)


def make_laplacian_value(ny, nx, delta, sparse=False):
    """5-point Laplacian with identity ('value') boundary rows, divided by delta**2
    (``finitediff.py:252-431`` with all bounds 'value')."""
    n = ny * nx
    k = np.arange(n)
    i, j = k // nx, k % nx
    interior = (i > 0) & (j > 0) & (i < ny - 1) & (j < nx - 1)
    rows, cols, vals = [], [], []
    ki = k[interior]
    for off in (nx, -nx, 1, -1):
        rows.append(ki); cols.append(ki + off); vals.append(np.ones(len(ki)))
    rows.append(ki); cols.append(ki); vals.append(-4.0 * np.ones(len(ki)))
    kb = k[~interior]
    rows.append(kb); cols.append(kb); vals.append(np.ones(len(kb)))
    A = scipy.sparse.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))
    A = A / (delta ** 2)
    return A if sparse else A.toarray()


def sampled_steps(n_steps: int, t_resample: float):
    """Indices of sampled steps (``sim.py:2455-2462``)."""
    out = []
    i = 0
    while i < n_steps - t_resample:
        i += t_resample
        i = int(i)
        out.append(i)
    return out


# --------------------------------------------------------------------------
# voltage-gated channels
# --------------------------------------------------------------------------
def nav1p2_rates(V):
    """Nav1p2._calculate_state (vg_na.py:169-186); V in mV."""
    mAlpha = (0.182 * ((V - 10.0) - -35.0)) / (1 - (np.exp(-((V - 10.0) - -35.0) / 9)))
    mBeta = (0.124 * (-(V - 10.0) - 35.0)) / (1 - (np.exp(-(-(V - 10.0) - 35.0) / 9)))
    mInf = mAlpha / (mAlpha + mBeta)
    mTau = (1 / (mAlpha + mBeta))
    hInf = 1.0 / (1 + np.exp((V - -65.0 - 10.0) / 6.2))
    hTau = (1 / (
        (0.024 * ((V - 10.0) - -50.0)) /
        (1 - (np.exp(-((V - 10.0) - -50.0) / 5))) + (
            0.0091 * (-(V - 10.0) - 75.000123)) / (1 - (np.exp(-(-(V - 10) - 75.000123) / 5)))))
    return mInf, mTau, hInf, hTau


def nav1p2_init(V):
    """Nav1p2._init_state (vg_na.py:140-166)."""
    mAlpha = (0.182 * ((V - 10.0) - -35.0)) / (1 - (np.exp(-((V - 10.0) - -35.0) / 9)))
    mBeta = (0.124 * (-(V - 10.0) - 35.0)) / (1 - (np.exp(-(-(V - 10.0) - 35.0) / 9)))
    return mAlpha / (mAlpha + mBeta), 1.0 / (1 + np.exp((V - -65.0 - 10.0) / 6.2))


def kv1p1_rates(V):
    """Kv1p1._calculate_state (vg_k.py:167-178); V in mV."""
    mInf = 1.0000 / (1 + np.exp((V - -30.5000) / -11.3943))
    mTau = 30.0000 / (1 + np.exp((V - -76.5600) / 26.1479))
    hInf = 1.0000 / (1 + np.exp((V - -30.0000) / 27.3943))
    hTau = 15000.0000 / (1 + np.exp((V - -160.5600) / -100.0000))
    return mInf, mTau, hInf, hTau


def kv1p1_init(V):
    """Kv1p1._init_state (vg_k.py:142-163)."""
    return (1.0000 / (1 + np.exp((V - -30.5000) / -11.3943)),
            1.0000 / (1 + np.exp((V - -30.0000) / 27.3943)))


def cav1p2_rates(V):
    """Cav1p2._calculate_state (vg_ca.py:328-340); V in mV, shifted by -10 mV first."""
    V = V - 10
    mInf = 1.0000 / (1 + np.exp((V - -30.000) / -6))
    mTau = (5.0000 + 20.0000 / (1 + np.exp((V - -25.000) / 5)))
    hInf = 1.0000 / (1 + np.exp((V - -80.000) / 6.4))
    hTau = (20.0000 + 50.0000 / (1 + np.exp((V - -40.000) / 7)))
    return mInf, mTau, hInf, hTau


def cav1p2_init(V):
    """Cav1p2._init_state (vg_ca.py:302-326)."""
    V = V - 10
    return 1.0000 / (1 + np.exp((V - -30.000) / -6)), 1.0000 / (1 + np.exp((V - -80.000) / 6.4))


CHANNEL_KINDS = {
    # kind: (ion, rates, init, mpower, hpower, time_unit)
    "Nav1p2": ("Na", nav1p2_rates, nav1p2_init, 3, 1, 1.0e3),
    "Kv1p1": ("K", kv1p1_rates, kv1p1_init, 1, 2, 1.0e3),
    "Cav1p2": ("Ca", cav1p2_rates, cav1p2_init, 2, 1, 1.0e3),
}


class OracleChannel:
    """One voltage-gated channel on every membrane (targets = all)."""

    def __init__(self, kind, max_dm, modulator=1.0):
        self.kind = kind
        self.ion, self._rates, self._init, self._mpower, self._hpower, self.time_unit = CHANNEL_KINDS[kind]
        self.maxDm = max_dm
        self.modulator = modulator
        self.v_corr = 0.0
        self.ions = [self.ion]
        self.rel_perm = [1.0]

    def init(self, vm):
        V = vm * 1000 + self.v_corr
        self.m, self.h = self._init(V)

    def run(self, vm, p):
        V = vm * 1000 + self.v_corr
        self._mInf, self._mTau, self._hInf, self._hTau = self._rates(V)
        dt = p.dt * self.time_unit
        self.m = (self._mTau * self.m + dt * self._mInf) / (self._mTau + dt)
        self.h = (self._hTau * self.h + dt * self._hInf) / (self._hTau + dt)
        self.P = (self.m ** self._mpower) * (self.h ** self._hpower)


# --------------------------------------------------------------------------
class OracleSim:
    """Simulator state and step, restated from ``betse/science/sim.py``."""

    def __init__(self, cells, p, mode: str = "sparse"):
        if mode not in ("dense", "sparse"):
            raise ValueError(mode)
        self.cells = cells
        self.p = p
        self.mode = mode
        C, M = cells.n_cells, cells.n_mems
        G2 = cells.X.size
        self.shape = cells.X.shape
        if mode == "dense":
            Msum = np.zeros((C, M))
            Msum[cells.mem_to_cells, np.arange(M)] = 1.0
            self.M_sum_mems = Msum
            Mdiv = np.zeros((G2, M))
            for m, k in enumerate(cells.map_mem2ecm):
                Mdiv[k, m] += cells.mem_sa[m]
            self.M_divmap_mem2ecm = Mdiv
            A = make_laplacian_value(self.shape[0], self.shape[1], cells.delta)
            self.lapENVinv = np.linalg.pinv(A)
        else:
            self._lap = make_laplacian_value(self.shape[0], self.shape[1], cells.delta, sparse=True).tocsc()
            self._lap_solve = scipy.sparse.linalg.factorized(self._lap)
        self.storage: List[Dict[str, np.ndarray]] = []

    # ---- operators -------------------------------------------------------
    def msum(self, x):
        """``np.dot(cells.M_sum_mems, x)``."""
        if self.mode == "dense":
            return np.dot(self.M_sum_mems, x)
        return np.add.reduceat(x, self.cells.mem_ptr[:-1])

    def divmap(self, x):
        """``np.dot(cells.M_divmap_mem2ecm, x)``."""
        if self.mode == "dense":
            return np.dot(self.M_divmap_mem2ecm, x)
        c = self.cells
        return np.bincount(c.map_mem2ecm, weights=c.mem_sa * x, minlength=c.X.size)

    def lap_inv(self, b):
        """``np.dot(cells.lapENVinv, b)``."""
        if self.mode == "dense":
            return np.dot(self.lapENVinv, b)
        return self._lap_solve(b)

    # ---- setup -----------------------------------------------------------
    def init_core(self):
        p, cells = self.p, self.cells
        self.mdl = cells.n_mems
        self.cdl = cells.n_cells
        self.edl = cells.X.size
        self.extra_rho_cells = np.zeros(self.cdl)
        self.extra_rho_env = np.zeros(self.edl)
        self.extra_J_mem = np.zeros(self.mdl)
        self.extra_Jenv_x = np.zeros(self.edl)
        self.extra_Jenv_y = np.zeros(self.edl)
        # intracellular reaction network (MasterOfNetworks), substances uniform at c0
        self.network = getattr(p, "network", None)
        self.c_net = (np.tile(np.asarray(self.network.c0, dtype=np.float64)[:, None], (1, self.cdl))
                      if self.network is not None else None)
        self.vgj = np.zeros(self.mdl)
        self.gj_block = 1
        self.T = p.T
        self.vm = np.zeros(self.mdl)
        self.vm_ave = np.zeros(self.cdl)
        self.rho_cells = np.zeros(self.cdl)
        self.v_env = np.zeros(self.edl)
        self.rho_env = np.zeros(self.edl)
        self.E_env_x = np.zeros(cells.X.shape)
        self.E_env_y = np.zeros(cells.X.shape)
        self.Jn = np.zeros(self.mdl)
        self.names = p.ion_names()
        cc_cells, cc_env, zs, Dm, D_free, D_gj, D_env, cbound, Dtj = [], [], [], [], [], [], [], [], []
        for i, name in enumerate(self.names):
            setattr(self, "i" + name, i)
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
        self.fluxes_mem = np.zeros((len(self.names), self.mdl))
        self.bound_V = {"T": 0, "B": 0, "L": 0, "R": 0}
        self.initDenv()
        self.gjopen = np.ones(self.mdl) * cells.gj_default_weights
        self.fluxes_gj = np.copy(self.fluxes_mem)
        self.cc_at_mem = np.asarray([cc[cells.mem_to_cells] for cc in self.cc_cells])
        self.Phi_b = np.zeros(self.edl)
        self.init_channels()

    def init_channels(self):
        """``Channel.init_channel`` (networks.py:6550-6628) for each configured channel;
        gates start from ``init(vm)`` or, if the parameters give them, from
        explicit per-membrane arrays."""
        p, cells = self.p, self.cells
        self.channels = []
        if getattr(p, "channels", ()) and getattr(p, "channel_random_state", False) \
                and getattr(p, "channel_gates0", None) is None:
            raise ValueError("the workload's seeded channel state must be drawn before the oracle runs")
        for k, spec in enumerate(getattr(p, "channels", ())):
            mod = 1.0
            mults = getattr(p, "channel_cell_mult", None)
            if mults is not None and mults[k] is not None:
                mod = np.asarray(mults[k])[cells.mem_to_cells]
            ch = OracleChannel(spec.kind, spec.max_dm, mod)
            ch.init(self.vm)
            gates = getattr(p, "channel_gates0", None)
            if gates is not None and gates[k] is not None:
                ch.m = np.array(gates[k][0], dtype=np.float64)
                ch.h = np.array(gates[k][1], dtype=np.float64)
            self.channels.append(ch)

    def initDenv(self):
        p, cells = self.p, self.cells
        for i in range(len(self.D_env)):
            Denv_o = np.ones(self.edl) * self.D_free[i]
            Denv_o[cells.envInds_inClust] = self.D_free[i] * p.D_adh
            Denv_o[cells.all_bound_mem_inds] = self.D_free[i] * p.D_tj * self.Dtj_rel[i]
            Denv_o[cells.interior_bound_mem_inds] = self.D_free[i] * p.D_tj * self.Dtj_rel[i]
            Denv_o[cells.ecm_inds_bound_cell] = self.D_free[i] * p.D_tj * self.Dtj_rel[i]
            self.D_env[i] = Denv_o * 1.0
        self.TJ_modulator = np.ones(self.D_env.shape)

    def init_dynamics(self, kind: str, rng: Optional[np.random.RandomState] = None):
        p, cells = self.p, self.cells
        nfrac = p.smooth_cells
        nm = cells.num_mems[cells.mem_to_cells]
        self.smooth_weight_mem = (nfrac * nm - 1) / (nfrac * nm)
        self.smooth_weight_o = 1 / (nfrac * nm)
        self.initDenv()
        if p.cbnd is not None:
            for i, name in enumerate(self.names):
                self.c_env_bound[i] = p.cbnd[name]
        # tissue profiles (tishandler.py:363-443): a default tissue covering all
        # membranes, then any extra profiles in order
        for i, name in enumerate(self.names):
            self.Dm_cells[i][:] = p.mem_perms[name]
        for cell_inds, dms in getattr(p, "tissue_profiles", ()):
            mems = np.concatenate([cells.cell_to_mems[c] for c in cell_inds]).astype(np.int64)
            for name, val in dms.items():
                if name in self.names:
                    self.Dm_cells[self.names.index(name)][mems] = val
        self.D_env_base = np.copy(self.D_env)
        self.fluxes_env_x = np.zeros((len(self.zs), self.edl))
        self.fluxes_env_y = np.zeros((len(self.zs), self.edl))
        self.Dm_base = np.copy(self.Dm_cells)
        self.Dm_scheduled = np.zeros_like(self.Dm_cells)
        rng = rng if rng is not None else np.random
        self.channel_noise_factor = rng.random_sample(self.mdl)
        if "K" in self.names:
            iK = self.iK
            self.Dm_cells[iK] = (p.channel_noise_level * self.channel_noise_factor + 1) * self.Dm_cells[iK]
        self.gj_block = 1 if p.gj_block == 1.0 else p.gj_block
        self.NaKATP_block = p.NaKATP_block
        self.CaATP_block = np.ones(self.mdl)
        self.rho_pump = 1
        self.rho_channel = 1
        self.ko_env = (np.sqrt(np.dot((p.NAv * (p.q ** 2) * self.zs ** 2) / (p.er * p.eo * p.kb * p.T),
                                      self.cc_env))).mean()
        self.ko_cell = (np.sqrt(np.dot((p.NAv * (p.q ** 2) * self.zs ** 2) / (p.er * p.eo * p.kb * p.T),
                                       self.cc_cells))).mean()

    # ---- phase driver ----------------------------------------------------
    def run_sim_core(self, kind: str, n_steps: Optional[int] = None,
                     on_sample: Optional[Callable] = None, observables: Optional[bool] = None):
        """``run_sim_core`` then ``_run_sim_core_loop``; returns sampled indices."""
        p = self.p
        p.set_phase_time(kind)
        self.init_dynamics(kind)
        self.fluxes_mem.fill(0)
        self.fluxes_gj.fill(0)
        self.update_V(observables=True)
        n = n_steps if n_steps is not None else (p.init_tsteps if kind == "init" else p.sim_tsteps)
        sampled = set(sampled_steps(n, p.t_resample))
        obs_all = (self.mode == "dense") if observables is None else observables
        for it in range(n):
            is_sampled = it in sampled
            self.step(kind, observables=(obs_all or is_sampled))
            if is_sampled:
                self.write2storage(it)
                if on_sample is not None:
                    on_sample(it, self)
        return sorted(sampled)

    def write2storage(self, it):
        rec = {
            "step": it, "vm": self.vm.copy(), "vm_ave": self.vm_ave.copy(),
            "cc_cells": self.cc_cells.copy(), "cc_env": self.cc_env.copy(),
            "v_env": self.v_env.copy(), "E_env_x": self.E_env_x.copy(), "E_env_y": self.E_env_y.copy(),
        }
        for k in ("Jn", "I_mem", "J_cell_x", "J_cell_y", "J_env_x", "J_env_y", "Jtx", "Jty", "B_field",
                  "rate_NaKATP", "E_gj_x", "E_gj_y", "fluxes_mem", "fluxes_gj", "fluxes_env_x", "fluxes_env_y"):
            if hasattr(self, k):
                rec[k] = np.array(getattr(self, k), copy=True)
        self.storage.append(rec)

    # ---- one iteration of _run_sim_core_loop -----------------------------
    def step(self, kind: str = "sim", observables: bool = True):
        p, cells = self.p, self.cells
        self.fluxes_mem.fill(0)
        self.fluxes_gj.fill(0)
        self.fluxes_env_x = np.zeros((len(self.zs), self.edl))
        self.fluxes_env_y = np.zeros((len(self.zs), self.edl))
        if kind == "sim":                       # fire_events -> makeAllChanges (Q10)
            self.Dm_cells = self.Dm_scheduled + self.Dm_base
        if p.alpha_NaK == 0.0:
            self.rate_NaKATP = np.zeros(self.mdl)
        if p.alpha_NaK > 0.0:
            fNa, fK, self.rate_NaKATP = pumpNaKATP(
                self.cc_at_mem[self.iNa], self.cc_env[self.iNa][cells.map_mem2ecm],
                self.cc_at_mem[self.iK], self.cc_env[self.iK][cells.map_mem2ecm],
                self.vm, self.T, p, self.NaKATP_block)
            fNa = self.rho_pump * fNa
            fK = self.rho_pump * fK
            if p.cluster_open is False:
                fNa[cells.bflags_mems] = 0
                fK[cells.bflags_mems] = 0
            self.fluxes_mem[self.iNa] += fNa
            self.fluxes_mem[self.iK] += fK
        IdM = np.ones(self.mdl)
        for i in self.movingIons:
            f_ED = electroflux(self.cc_env[i][cells.map_mem2ecm], self.cc_at_mem[i], self.Dm_cells[i],
                               IdM * p.tm, self.zs[i] * IdM, self.vm, self.T, p, rho=self.rho_channel)
            if not p.cluster_open:
                f_ED[cells.bflags_mems] = 0
            self.fluxes_mem[i] += f_ED
            self.update_gj(i)
            self.update_ecm(i)
            self.update_intra(i)
        if "Ca" in self.names:
            self.ca_handler()
        if self.channels and (kind == "sim" or getattr(self.p, "channels_init_active", True)):
            self.extra_J_mem = np.zeros(self.mdl)       # clear_run_loop (networks.py:2790-2801)
            self.run_loop_channels()
        if self.network is not None:                    # molecules run_loop (sim.py:1286-1300)
            from oracle import network_oracle
            self.c_net = network_oracle.step(self.network, self.c_net, p.dt, p.R, self.T)
        self.update_all_concs()
        self.update_V(observables=observables)
        check_v(self.vm)

    def update_gj(self, i):
        p, cells = self.p, self.cells
        self.vgj = self.vm[cells.nn_i] - self.vm[cells.mem_i]
        self.Egj = -self.vgj / cells.gj_len
        self.E_gj_x = self.Egj * cells.mem_vects_flat[:, 2]
        self.E_gj_y = self.Egj * cells.mem_vects_flat[:, 3]
        self.gjopen = self.gj_block * np.ones(self.mdl) * cells.gj_default_weights
        conc_mem = self.cc_at_mem[i]
        fgj_X = electroflux(conc_mem[cells.mem_i], conc_mem[cells.nn_i],
                            self.D_gj[i] * p.gj_surface * self.gjopen,
                            cells.gj_len * np.ones(self.mdl), self.zs[i] * np.ones(self.mdl),
                            self.vgj, p.T, p, rho=1)
        fgj_X[cells.bflags_mems] = 0.0
        self.fluxes_gj[i] = self.fluxes_gj[i] + fgj_X

    def update_ecm(self, i):
        p, cells = self.p, self.cells
        cenv = self.cc_env[i].reshape(cells.X.shape)       # view: boundary reset writes state (Q3)
        cenv[:, 0] = self.c_env_bound[i]
        cenv[:, -1] = self.c_env_bound[i]
        cenv[0, :] = self.c_env_bound[i]
        cenv[-1, :] = self.c_env_bound[i]
        gcx, gcy = fd_gradient(cenv, cells.delta)
        ux = np.zeros(cells.X.shape)
        uy = np.zeros(cells.X.shape)
        denv = self.D_env[i].reshape(cells.X.shape) * self.TJ_modulator[i].reshape(cells.X.shape)
        fx, fy = nernst_planck_flux(cenv, gcx, gcy, -self.E_env_x, -self.E_env_y, ux, uy,
                                    denv, self.zs[i], self.T, p)
        self.fluxes_env_x[i] = fx.ravel()
        self.fluxes_env_y[i] = fy.ravel()
        div_fa = fd_divergence(-fx, -fy, cells.delta, cells.delta)
        cenv = cenv + div_fa * p.dt
        if p.sharpness < 1.0:
            cenv = fd_integrator(cenv, sharp=p.sharpness)
        self.cc_env[i] = cenv.ravel()

    def update_intra(self, i):
        self.cc_at_mem[i] = self.cc_cells[i][self.cells.mem_to_cells] * 1

    def ca_handler(self):
        p, cells = self.p, self.cells
        if p.alpha_Ca > 0.0:
            cCao = self.cc_env[self.iCa][cells.map_mem2ecm]
            f_CaATP = pumpCaATP(self.cc_at_mem[self.iCa], cCao, self.vm, self.T, p, self.CaATP_block)
            f_CaATP = self.rho_pump * f_CaATP
            self.rate_CaATP = f_CaATP
            if p.cluster_open is False:
                f_CaATP[cells.bflags_mems] = 0
            self.fluxes_mem[self.iCa] += self.rho_pump * (f_CaATP)

    def run_loop_channels(self):
        """``MasterOfNetworks.run_loop_channels`` (networks.py:3115-3215): gates
        from the current (nonce-carrying) Vmem, GHK through the channel with the
        fresh cell and post-diffusion env concentrations, then an immediate
        ``update_Co`` of cell and env."""
        p, cells = self.p, self.cells
        IdM = np.ones(self.mdl)
        for chan in self.channels:
            moddy = chan.modulator
            chan.run(self.vm, p)
            for ion, rel_perm in zip(chan.ions, chan.rel_perm):
                ioni = self.names.index(ion)
                zzz = self.zs[ioni]
                DChan = chan.P * rel_perm * chan.maxDm * moddy
                cenv = self.cc_env[ioni]
                ccell = self.cc_cells[ioni]
                cmem = self.cc_cells[ioni][cells.mem_to_cells]
                f_ED = electroflux(cenv[cells.map_mem2ecm], cmem, DChan, IdM * p.tm, zzz * IdM, self.vm,
                                   self.T, p, rho=self.rho_channel)
                delta_cells = self.msum(f_ED * cells.mem_sa) / cells.cell_vol
                cX_cell = ccell + delta_cells * p.dt
                cX_mem = cX_cell[cells.mem_to_cells]
                delta_env = self.divmap(-f_ED) / (p.cell_height * cells.delta ** 2)
                cX_env = cenv + delta_env * p.dt
                self.cc_cells[ioni], self.cc_at_mem[ioni], self.cc_env[ioni] = cX_cell, cX_mem, cX_env
                self.extra_J_mem += -f_ED * p.F * zzz
                chan.chan_flux = f_ED

    def update_all_concs(self):
        p, cells = self.p, self.cells
        for i in self.movingIons:
            flux = self.fluxes_mem[i]
            f_gj_i = self.fluxes_gj[i]
            # update_Co (sim_toolbox.py:1155-1207) with update_at_mems False
            delta_cells = self.msum(flux * cells.mem_sa) / cells.cell_vol
            cX_cell = self.cc_cells[i] + delta_cells * p.dt
            cX_mem = cX_cell[cells.mem_to_cells]
            delta_env = self.divmap(-flux) / (p.cell_height * cells.delta ** 2)
            cX_env = self.cc_env[i] + delta_env * p.dt
            self.cc_cells[i], self.cc_at_mem[i], self.cc_env[i] = cX_cell, cX_mem, cX_env
            delta_cgj = self.msum(-f_gj_i * cells.mem_sa) / cells.cell_vol
            self.cc_cells[i] += p.dt * delta_cgj
            no_negs(self.cc_cells[i])

    def update_V(self, observables: bool = True):
        p, cells = self.p, self.cells
        vmo = self.vm * 1
        self.get_current(observables=observables)
        rho_surf = self.rho_cells * cells.diviterm
        self.vm = (1 / p.cm) * rho_surf[cells.mem_to_cells] - self.Phi_b[cells.map_mem2ecm]
        self.vm_ave = self.msum(self.vm) / cells.num_mems
        self.dvm = (self.vm - vmo) / p.dt

    def get_current(self, observables: bool = True):
        p, cells = self.p, self.cells
        self.rho_cells = np.dot(self.zs * p.F, self.cc_cells) + self.extra_rho_cells
        if observables:
            self.Jmem = -np.dot(self.zs * p.F, self.fluxes_mem) + self.extra_J_mem
            self.Jgj = np.dot(self.zs * p.F, self.fluxes_gj)
            self.Jn = self.Jmem + self.Jgj
            Jn_ave = self.msum(self.Jn * cells.mem_sa) / cells.cell_sa
            self.Jn = self.smooth_weight_mem * self.Jn + Jn_ave[cells.mem_to_cells] * self.smooth_weight_o
            self.I_mem = -self.Jn * cells.mem_sa
            Jcx = self.Jn * cells.mem_vects_flat[:, 2]
            Jcy = self.Jn * cells.mem_vects_flat[:, 3]
            self.J_cell_x = self.msum(Jcx * cells.mem_sa) / cells.cell_sa
            self.J_cell_y = self.msum(Jcy * cells.mem_sa) / cells.cell_sa
            J_env_x_o = (np.dot(p.F * self.zs, self.fluxes_env_x) + self.extra_Jenv_x).reshape(cells.X.shape)
            J_env_y_o = (np.dot(p.F * self.zs, self.fluxes_env_y) + self.extra_Jenv_y).reshape(cells.X.shape)
            BB, self.J_env_x, self.J_env_y, _, Ja, Jb = self.HH_Decomp(J_env_x_o, J_env_y_o, bounds=self.bound_V)
            self.B_field = BB * p.mu
            self.Jtx = self.J_env_x + Ja
            self.Jty = self.J_env_y + Jb
        self.rho_env = np.dot(self.zs * p.F, self.cc_env) + self.extra_rho_env
        div_Jb = np.zeros(cells.X.shape)
        div_Jb[:, -1] = -self.bound_V["R"] / cells.delta ** 2
        div_Jb[:, 0] = -self.bound_V["L"] / cells.delta ** 2
        div_Jb[-1, :] = -self.bound_V["T"] / cells.delta ** 2
        div_Jb[0, :] = -self.bound_V["B"] / cells.delta ** 2
        if self.mode == "dense" or any(v != 0 for v in self.bound_V.values()):
            Phi_b = self.lap_inv(-div_Jb.ravel())
        else:
            Phi_b = np.zeros(self.edl)        # lapENVinv . 0 == 0 (Q20)
        self.rho_env_surf = np.zeros(self.edl)
        mp = cells.map_mem2ecm
        self.rho_env_surf[mp] = self.rho_env[mp] * (p.cell_height * cells.delta ** 2) / (
            cells.memSa_per_envSquare[mp].mean())
        self.v_env = self.rho_env_surf / (self.ko_env * p.eo * p.er)
        screen = (2 / (self.ko_env * cells.delta)) * (p.cell_radius / p.true_cell_size)
        self.v_env = gaussian_filter_const(self.v_env.reshape(cells.X.shape), 1).ravel() + Phi_b
        gVex, gVey = fd_gradient(screen * self.v_env.reshape(cells.X.shape), cells.delta)
        self.E_env_x = -gVex
        self.E_env_y = -gVey
        self.Phi_b = Phi_b

    def HH_Decomp(self, JJx, JJy, bounds=None, sigma=1.0):
        cells = self.cells
        if bounds is not None:
            Lb, Rb, Tb, Bb = bounds["L"], bounds["R"], bounds["T"], bounds["B"]
        else:
            Lb = Rb = Tb = Bb = 0.0
        Jxr = -JJy.reshape(cells.X.shape)
        Jyr = JJx.reshape(cells.X.shape)
        divJr = fd_divergence(Jxr, Jyr, cells.delta, cells.delta)
        divJr[:, 0] = 0.0
        divJr[:, -1] = 0.0
        divJr[0, :] = 0.0
        divJr[-1, :] = 0.0
        AA = self.lap_inv(-divJr.ravel())
        gAx, gAy = fd_gradient(AA.reshape(cells.X.shape), cells.delta)
        Fx = -gAy
        Fy = gAx
        divJd = fd_divergence(JJx.reshape(cells.X.shape) / sigma, JJy.reshape(cells.X.shape) / sigma,
                              cells.delta, cells.delta)
        divJd[:, 0] = -Lb * (1 / cells.delta ** 2)
        divJd[:, -1] = -Rb * (1 / cells.delta ** 2)
        divJd[0, :] = -Bb * (1 / cells.delta ** 2)
        divJd[-1, :] = -Tb * (1 / cells.delta ** 2)
        BB = self.lap_inv(divJd.ravel())
        Gx, Gy = fd_gradient(BB.reshape(cells.X.shape), cells.delta)
        return AA, Fx, Fy, BB, Gx, Gy
