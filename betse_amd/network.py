"""Per-cell reaction network of intracellular substances (host side).

Describes the cell-zone part of ``MasterOfNetworks`` (``chemistry/networks.py``):

* every substance has a growth/decay rate (``write_growth_and_decay``,
  ``networks.py:1187-1305``)::

      r = mod * r_production * alpha_growth - r_decay * c/scale - decay_max * c/scale

* every reaction has a rate (``write_reactions``, ``networks.py:1312-1570``)::

      irreversible:  r = vmax * alpha * prod_reactants ((c/Km)**k / (1 + (c/Km)**k))
      reversible:    r = vmax * alpha * (1 - Q/Keq),  Keq = exp(-delta_Go/(R T)),
                     Q = prod_products(1e-3 * c**k) / prod_reactants(1e-3 * c**k)

* ``alpha`` is the product of Hill activator terms ``(c/Km)**n/(1+(c/Km)**n)``
  and inhibitor terms ``1/(1+(c/Km)**n)`` (``get_influencers``,
  ``networks.py:5084-5600``);
* a step is ``c += (reaction_matrix . rates) * dt`` then ``no_negs``
  (``run_loop``, ``networks.py:2805-2982``; ``create_reaction_matrix``,
  ``networks.py:2686-2725``).

The substances here stay inside their cells (zero membrane and gap-junction
permeability), so ``molecule_mover`` moves nothing and the update is per cell.
The device kernel evaluates the network every step (``bt_set_network``).
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import _native


@dataclass
class Hill:
    species: int
    km: float
    n: float


@dataclass
class Growth:
    """Growth/decay of one substance (``simple_growth``)."""
    species: int
    r_production: float = 0.0
    r_decay: float = 0.0
    decay_max: float = 0.0
    gad_scale: float = 1.0
    time_factor: float = 1.0            # modify_time_factor
    activators: List[Hill] = field(default_factory=list)
    inhibitors: List[Hill] = field(default_factory=list)


@dataclass
class Reaction:
    reactants: List[Tuple[int, float, float]]      # (species, coefficient, Km)
    products: List[Tuple[int, float, float]]
    vmax: float
    delta_Go: Optional[float] = None
    activators: List[Hill] = field(default_factory=list)
    inhibitors: List[Hill] = field(default_factory=list)


@dataclass
class ReactionNetwork:
    names: List[str]
    c0: np.ndarray                       # initial concentration per substance (mM)
    growth: List[Growth]                 # one per substance, in substance order
    reactions: List[Reaction]

    @property
    def n_species(self) -> int:
        return len(self.names)

    def reaction_matrix(self) -> np.ndarray:
        """``create_reaction_matrix`` (networks.py:2686-2725): one growth/decay
        column per substance, then one column per reaction."""
        S = self.n_species
        M = np.zeros((S, S + len(self.reactions)))
        for i, g in enumerate(self.growth):
            M[g.species, i] += 1
        for jo, r in enumerate(self.reactions):
            j = jo + S
            for s, k, _ in r.reactants:
                M[s, j] += -k
            for s, k, _ in r.products:
                M[s, j] += k
        return M

    def keq(self, r: Reaction, R: float, T: float) -> float:
        return float(np.exp(-r.delta_Go / (R * T)))

    def pack(self, R: float, T: float) -> "_native.BtNetwork":
        """The C description (``bt_network``)."""
        S = self.n_species
        if S > _native.BT_NET_MAX_SPECIES or S + len(self.reactions) > _native.BT_NET_MAX_RATES:
            raise ValueError("reaction network too large for the kernel")
        if len(self.growth) != S or any(g.species != i for i, g in enumerate(self.growth)):
            raise ValueError("one growth/decay entry per substance, in order")
        M = self.reaction_matrix()
        net = _native.BtNetwork()
        net.n_species = S
        net.n_rates = S + len(self.reactions)

        def hills(rt, acts, inhs):
            if len(acts) > _native.BT_NET_MAX_HILL or len(inhs) > _native.BT_NET_MAX_HILL:
                raise ValueError("too many activators/inhibitors")
            rt.n_act, rt.n_inh = len(acts), len(inhs)
            for q, h in enumerate(acts):
                rt.act[q], rt.act_km[q], rt.act_n[q] = h.species, h.km, h.n
            for q, h in enumerate(inhs):
                rt.inh[q], rt.inh_km[q], rt.inh_n[q] = h.species, h.km, h.n

        def stoich(rt, j):
            nz = np.nonzero(M[:, j])[0]
            if len(nz) > _native.BT_NET_MAX_STOICH:
                raise ValueError("reaction touches too many substances")
            rt.n_stoich = len(nz)
            for q, s in enumerate(nz):
                rt.stoich_species[q], rt.stoich_coeff[q] = int(s), float(M[s, j])

        for i, g in enumerate(self.growth):
            rt = net.rates[i]
            rt.kind = _native.BT_RATE_GROWTH
            rt.species = g.species
            rt.r_prod = g.r_production * g.time_factor
            rt.r_decay = g.r_decay * g.time_factor
            rt.decay_max = g.decay_max * g.time_factor
            rt.gad_scale = g.gad_scale
            hills(rt, g.activators, g.inhibitors)
            stoich(rt, i)
        for jo, r in enumerate(self.reactions):
            rt = net.rates[S + jo]
            rt.kind = _native.BT_RATE_REACTION
            rt.reversible = 1 if r.delta_Go is not None else 0
            if len(r.reactants) > _native.BT_NET_MAX_TERMS or len(r.products) > _native.BT_NET_MAX_TERMS:
                raise ValueError("too many reaction terms")
            rt.n_reac, rt.n_prod = len(r.reactants), len(r.products)
            for q, (s, k, km) in enumerate(r.reactants):
                rt.reac[q], rt.reac_coeff[q], rt.reac_km[q] = s, k, km
            for q, (s, k, km) in enumerate(r.products):
                rt.prod[q], rt.prod_coeff[q], rt.prod_km[q] = s, k, km
            rt.vmax = r.vmax
            rt.keq = self.keq(r, R, T) if r.delta_Go is not None else 1.0
            hills(rt, r.activators, r.inhibitors)
            stoich(rt, S + jo)
        return net


def draw_network(seed: int, n_species: int = 8, n_reactions: int = 8, rate_range=(1.0e-3, 1.0),
                 hill_ns: Sequence[int] = (1, 2, 4), km_range=(0.1, 1.0), c0_range=(0.1, 1.0),
                 p_reversible: float = 0.5) -> ReactionNetwork:
    """A random mass-action/Hill network: rate constants log-uniform over
    ``rate_range`` (s^-1, production in mM/s), Hill exponents uniform over
    ``hill_ns``, half-saturation constants and initial concentrations
    log-uniform over their ranges (mM), reversible reactions with
    delta_Go uniform in [-15, 5] kJ/mol."""
    rng = np.random.RandomState(seed)
    lo, hi = np.log10(rate_range[0]), np.log10(rate_range[1])

    def rate():
        return float(10.0 ** rng.uniform(lo, hi))

    def km():
        return float(10.0 ** rng.uniform(np.log10(km_range[0]), np.log10(km_range[1])))

    def hill(exclude=None):
        s = int(rng.randint(n_species))
        while exclude is not None and s == exclude and n_species > 1:
            s = int(rng.randint(n_species))
        return Hill(s, km(), float(rng.choice(hill_ns)))

    names = [f"S{i}" for i in range(n_species)]
    c0 = 10.0 ** rng.uniform(np.log10(c0_range[0]), np.log10(c0_range[1]), n_species)
    growth = []
    for i in range(n_species):
        g = Growth(i, r_production=rate(), r_decay=rate())
        if rng.rand() < 0.5:
            g.activators.append(hill(exclude=i))
        if rng.rand() < 0.5:
            g.inhibitors.append(hill(exclude=i))
        growth.append(g)
    reactions = []
    for _ in range(n_reactions):
        sp = rng.permutation(n_species)
        nr, npd = int(rng.randint(1, 3)), int(rng.randint(1, 3))
        reac = [(int(s), 1.0, km()) for s in sp[:nr]]
        prod = [(int(s), 1.0, km()) for s in sp[nr:nr + npd]]
        r = Reaction(reac, prod, vmax=rate())
        if rng.rand() < p_reversible:
            r.delta_Go = float(rng.uniform(-15.0e3, 5.0e3))
        if rng.rand() < 0.3:
            r.activators.append(hill())
        if rng.rand() < 0.3:
            r.inhibitors.append(hill())
        reactions.append(r)
    return ReactionNetwork(names, c0, growth, reactions)
