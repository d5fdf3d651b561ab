"""The restated reaction network against closed forms (no GPU)."""

import numpy as np
import pytest

from betse_amd.network import Growth, Hill, Reaction, ReactionNetwork, draw_network
from oracle import network_oracle as no

R, T = 8.314, 310.0


def _net(growth, reactions, c0):
    return ReactionNetwork([f"S{i}" for i in range(len(c0))], np.asarray(c0, dtype=float), growth, reactions)


def test_production_decay_follows_euler_recursion_to_steady_state():
    p, d, dt = 0.3, 0.05, 0.1
    net = _net([Growth(0, r_production=p, r_decay=d)], [], [1.0])
    C = np.array([[1.0, 2.0, 10.0]])
    want = C.copy()
    for _ in range(5000):
        C = no.step(net, C, dt, R, T)
        want = want + (p - d * want) * dt          # dc/dt = p - d c, explicit Euler
    assert np.allclose(C, want, rtol=1e-12)
    assert np.allclose(C, p / d, rtol=1e-6)


def test_irreversible_conversion_conserves_total_and_follows_saturating_law():
    km, vmax, dt = 0.5, 0.2, 0.01
    net = _net([Growth(0), Growth(1)], [Reaction([(0, 1.0, km)], [(1, 1.0, km)], vmax)], [1.0, 0.0])
    C = np.array([[1.0, 0.3], [0.0, 0.2]])
    r = no.rates(net, C, R, T)
    assert np.allclose(r[2], vmax * (C[0] / km) / (1 + C[0] / km))
    tot = C.sum(axis=0)
    for _ in range(5000):
        C = no.step(net, C, dt, R, T)
    assert np.allclose(C.sum(axis=0), tot, rtol=1e-13)
    assert np.all(C[0] < 0.05)


def test_reversible_reaction_is_at_rest_when_q_equals_keq():
    dG = -2000.0
    keq = np.exp(-dG / (R * T))
    # A <-> B, Q = (1e-3 cB)/(1e-3 cA) = cB/cA
    cA = np.array([0.4, 1.0])
    C = np.vstack([cA, keq * cA])
    net = _net([Growth(0), Growth(1)], [Reaction([(0, 1.0, 1.0)], [(1, 1.0, 1.0)], 0.7, delta_Go=dG)], [1, 1])
    r = no.rates(net, C, R, T)
    assert np.allclose(r[2], 0.0, atol=1e-15)
    # away from equilibrium the rate is vmax (1 - Q/Keq)
    C2 = np.vstack([cA, 0.5 * keq * cA])
    assert np.allclose(no.rates(net, C2, R, T)[2], 0.7 * 0.5)


def test_hill_terms_and_empty_products():
    net = _net([Growth(0, r_production=1.0, activators=[Hill(1, 0.3, 4.0)], inhibitors=[Hill(2, 0.2, 2.0)]),
                Growth(1), Growth(2)], [], [1, 1, 1])
    C = np.array([[1.0], [0.3], [0.2]])            # activator and inhibitor at their Km
    assert np.allclose(no.rates(net, C, R, T)[0], 0.5 * 0.5)
    C = np.array([[1.0], [0.0], [0.0]])            # no activator -> 0; inhibitor absent -> 1
    assert np.allclose(no.rates(net, C, R, T)[0], 0.0)
    net2 = _net([Growth(0, r_production=0.8), Growth(1)], [], [1, 1])
    assert np.allclose(no.rates(net2, np.ones((2, 3)), R, T)[0], 0.8)


def test_reaction_matrix_and_no_negs():
    net = _net([Growth(0), Growth(1), Growth(2)],
               [Reaction([(0, 1.0, 1.0), (1, 2.0, 1.0)], [(2, 1.0, 1.0)], 5.0)], [1, 1, 1])
    M = no.reaction_matrix(net)
    assert np.array_equal(M, net.reaction_matrix())
    assert np.array_equal(M[:, 3], [-1.0, -2.0, 1.0])
    C = np.array([[1e-3], [1.0], [0.0]])
    out = no.step(net, C, 1.0, R, T)              # overshoots below zero -> clamped
    assert out[0, 0] == 0.0 and out[2, 0] > 0


def test_nan_substance_raises_like_no_negs():
    """A reversible reaction with reactant and product both at zero has
    Q = 0/0: the substance turns NaN and ``no_negs`` raises
    (``sim_toolbox.py:475-503``, called on every substance by run_loop)."""
    import warnings
    from oracle.betse_oracle import OracleUnstable
    net = _net([Growth(0), Growth(1)], [Reaction([(0, 1.0, 1.0)], [(1, 1.0, 1.0)], 1.0, delta_Go=-1.0e3)], [1, 1])
    C = np.array([[0.0, 1.0], [0.0, 1.0]])          # cell 0: 0/0; cell 1 finite
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        with pytest.raises(OracleUnstable):
            no.step(net, C, 1e-3, R, T)
        out = no.step(net, C[:, 1:], 1e-3, R, T)   # the finite cell alone steps
    assert np.all(np.isfinite(out))


def test_drawn_network_is_well_formed_and_packs():
    net = draw_network(seed=5)
    assert net.n_species == 8 and len(net.reactions) == 8
    M = net.reaction_matrix()
    assert M.shape == (8, 16)
    packed = net.pack(R, T)
    assert packed.n_species == 8 and packed.n_rates == 16
    for j in range(16):
        r = packed.rates[j]
        cols = {r.stoich_species[q]: r.stoich_coeff[q] for q in range(r.n_stoich)}
        assert cols == {int(s): float(M[s, j]) for s in np.nonzero(M[:, j])[0]}
    rng = np.random.RandomState(0)
    C = np.tile(net.c0[:, None], (1, 4)) * rng.uniform(0.5, 2.0, (8, 4))
    for _ in range(100):
        C = no.step(net, C, 1e-4, R, T)
    assert np.all(np.isfinite(C)) and np.all(C >= 0)
