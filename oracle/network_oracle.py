"""FP64 NumPy restatement of the cell-zone reaction network -- TEST INFRASTRUCTURE ONLY.

Follows ``MasterOfNetworks.run_loop`` (``betse/science/chemistry/networks.py:2805-2982``)
for substances that stay inside their cells: every rate is evaluated from the
pre-step concentrations, ``delta = reaction_matrix . rates``
(``create_reaction_matrix``, ``networks.py:2686-2725``), then each substance
is updated ``c + delta*dt`` and clamped by ``no_negs``
(``sim_toolbox.py:475-503``). The reference evaluates strings it writes at
set-up; this module evaluates the same expressions in the same operation
order:

* growth/decay (``write_growth_and_decay``, ``networks.py:1233-1249``)::

      mod*r_prod*A*I - r_decay*cc - dec_max*DA*DI*cc,   cc = c/gad_scale

  with ``mod`` = ``growth_mod_function_cells`` (ones), ``r_prod`` =
  ``r_production*modify_time_factor``; the decay influencers ``DA``/``DI``
  are empty products (ones);
* reactions (``write_reactions``, ``networks.py:1346-1535``): the forward
  coefficient is the product over reactants of ``((c/Km)**k)/(1 + (c/Km)**k)``;
  ``Q = (prod(1.0e-3*c**k over products)/prod(1.0e-3*c**k over reactants))``
  -- Python's precedence makes each factor ``1.0e-3*(c**k)``; with
  ``delta_Go`` the rate is ``vmax*A*I*(1.0 - (Q/Keq))``, without it
  ``vmax*A*I*(fwd - ((0/1)*0))``;
* influencers (``get_influencers``, ``networks.py:5084-5600``): ``A`` is the
  product of activator terms ``(((c/Km)**n)/(1 + (c/Km)**n))``, ``I`` the
  product of inhibitor terms ``(1/(1 + (c/Km)**n))``; an empty product is
  ``np.ones``.

The network description is duck-typed (``names``, ``growth``, ``reactions``
with the fields of ``betse_amd.network``).

Pinned by ``tests/test_network_oracle.py`` through closed forms: production
and decay alone follow the explicit Euler recursion of ``dc/dt = p - d c``
(and reach ``p/d``); an A -> B reaction conserves A + B; a reversible
reaction at ``Q = Keq`` has zero rate; a Hill activator at ``c = Km`` is one
half; ``no_negs`` clamps and raises on NaN. The eval-string order itself is restated from the
source, not pinned by a reference run (the reference has no network fixture).
"""

from __future__ import annotations

import numpy as np


def no_negs(x):
    """``stb.no_negs`` on a substance (``networks.py:2939`` ->
    ``sim_toolbox.py:475-503``): negatives clamp to 0, a NaN raises."""
    from oracle.betse_oracle import no_negs as _no_negs
    return _no_negs(x)


def _act(c, km, n):
    return (((c / km) ** n) / (1 + (c / km) ** n))


def _inh(c, km, n):
    return (1 / (1 + (c / km) ** n))


def _alpha(C, acts, inhs, n_cells):
    a = np.ones(n_cells)
    for q, h in enumerate(acts):
        t = _act(C[h.species], h.km, h.n)
        a = t if q == 0 else a * t
    i = np.ones(n_cells)
    for q, h in enumerate(inhs):
        t = _inh(C[h.species], h.km, h.n)
        i = t if q == 0 else i * t
    return a, i


def reaction_matrix(net):
    S = len(net.names)
    M = np.zeros((S, S + len(net.reactions)))
    for i, g in enumerate(net.growth):
        M[g.species, i] += 1
    for jo, r in enumerate(net.reactions):
        for s, k, _ in r.reactants:
            M[s, S + jo] += -k
        for s, k, _ in r.products:
            M[s, S + jo] += k
    return M


def rates(net, C, R: float, T: float):
    """All rates ([n_species + n_reactions][n_cells]) at concentrations C ([S][n_cells])."""
    n = C.shape[1]
    out = []
    for g in net.growth:
        A, I = _alpha(C, g.activators, g.inhibitors, n)
        cc = C[g.species] / g.gad_scale
        mod = np.ones(n)
        r_prod = g.r_production * g.time_factor
        r_decay = g.r_decay * g.time_factor
        dec_max = g.decay_max * g.time_factor
        out.append(mod * r_prod * A * I - r_decay * cc - dec_max * np.ones(n) * np.ones(n) * cc)
    for r in net.reactions:
        A, I = _alpha(C, r.activators, r.inhibitors, n)
        if r.delta_Go is not None:
            num = None
            for s, k, _ in r.products:
                f = (1.0e-3 * C[s] ** k)
                num = f if num is None else num * f
            den = None
            for s, k, _ in r.reactants:
                f = (1.0e-3 * C[s] ** k)
                den = f if den is None else den * f
            Q = (num / den)
            Keq = np.exp(-r.delta_Go / (R * T))
            out.append(r.vmax * A * I * (1.0 - ((Q / Keq))))
        else:
            fwd = None
            for s, k, km in r.reactants:
                f = _act(C[s], km, k)
                fwd = f if fwd is None else fwd * f
            out.append(r.vmax * A * I * (fwd - ((0 / 1) * 0)))
    return np.asarray(out)


def step(net, C, dt: float, R: float, T: float):
    """One ``run_loop`` update of the substances; returns the new C."""
    delta = np.dot(reaction_matrix(net), rates(net, C, R, T))
    out = np.empty_like(C)
    for s in range(C.shape[0]):
        out[s] = C[s] + delta[s] * dt
        no_negs(out[s])
    return out
