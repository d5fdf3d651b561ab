"""Voltage-gated channels applied to every membrane (host side).

The GPU kernel evaluates the rate laws each step; the host only sets the
channels up, as ``Channel.init_channel`` does (``chemistry/networks.py:6550-6628``):

* ``init_gates(kind, vm)`` -- the channel's ``_init_state`` at the current
  Vmem (``channels/vg_na.py:140-166`` for Nav1p2, ``channels/vg_k.py:142-163``
  for Kv1p1, ``channels/vg_ca.py:302-326`` for Cav1p2);
* ``draw_channel_state(p, cells)`` -- the seeded per-cell density multipliers
  (log-uniform) and per-membrane gates (uniform in [0, 1]) of a workload that
  asks for them.

The per-cell multiplier is the channel's modulator ``moddy``: the kernel forms
``DChan = P * rel_perm * maxDm * moddy`` (``networks.py:3164``).
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Optional

import numpy as np

# kind -> (ion, m power, h power, time unit)
CHANNEL_KINDS = {
    "Nav1p2": ("Na", 3, 1, 1.0e3),
    "Kv1p1": ("K", 1, 2, 1.0e3),
    "Cav1p2": ("Ca", 2, 1, 1.0e3),
}


def init_gates(kind: str, vm: np.ndarray):
    """Gate values of ``_init_state`` at Vmem ``vm`` (V)."""
    V = vm * 1000 + 0.0
    if kind == "Nav1p2":
        mAlpha = (0.182 * ((V - 10.0) - -35.0)) / (1 - (np.exp(-((V - 10.0) - -35.0) / 9)))
        mBeta = (0.124 * (-(V - 10.0) - 35.0)) / (1 - (np.exp(-(-(V - 10.0) - 35.0) / 9)))
        return mAlpha / (mAlpha + mBeta), 1.0 / (1 + np.exp((V - -65.0 - 10.0) / 6.2))
    if kind == "Kv1p1":
        return (1.0000 / (1 + np.exp((V - -30.5000) / -11.3943)),
                1.0000 / (1 + np.exp((V - -30.0000) / 27.3943)))
    if kind == "Cav1p2":
        W = V - 10
        return (1.0000 / (1 + np.exp((W - -30.000) / -6)),
                1.0000 / (1 + np.exp((W - -80.000) / 6.4)))
    raise ValueError(f"unsupported channel kind {kind!r}")


def draw_channel_state(p, n_cells: int, n_mems: int, seed: Optional[int] = None,
                       mult_range=(0.5, 2.0)) -> None:
    """Attach seeded per-cell multipliers and per-membrane gates to ``p``.

    Multipliers are ``exp(U(log lo, log hi))`` per cell and channel; gates are
    ``U(0, 1)`` per membrane, gate and channel, drawn from
    ``RandomState(seed)`` in channel order (multiplier, then m, then h).
    """
    rng = np.random.RandomState(p.channel_seed if seed is None else seed)
    lo, hi = np.log(mult_range[0]), np.log(mult_range[1])
    mults, gates = [], []
    for _ in p.channels:
        mults.append(np.exp(rng.uniform(lo, hi, n_cells)))
        m = rng.uniform(0.0, 1.0, n_mems)
        h = rng.uniform(0.0, 1.0, n_mems)
        gates.append((m, h))
    p.channel_cell_mult = mults
    p.channel_gates0 = gates
