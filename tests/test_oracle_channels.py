"""The oracle's voltage-gated channels: rate laws against reduced closed forms
and anchor points, the semi-implicit gate update, and the channel step."""

import numpy as np
import pytest

from betse_amd.cells import Cells
from betse_amd.channels import draw_channel_state, init_gates
from betse_amd.parameters import Parameters
from oracle import betse_oracle as O


def test_nav1p2_minf_reduced_form():
    """alpha/beta of Hammil's m gate share the factor x/(1-e^(-x/9)), x = V+25 mV:
    m_inf = 1/(1 + (0.124/0.182) e^(-x/9)), tau_m = (1-e^(-x/9))/(x (0.182 + 0.124 e^(-x/9)))."""
    V = np.linspace(-90, 40, 131) + 0.123
    mInf, mTau, hInf, hTau = O.nav1p2_rates(V)
    x = V + 25.0
    np.testing.assert_allclose(mInf, 1 / (1 + (0.124 / 0.182) * np.exp(-x / 9)), rtol=1e-12)
    np.testing.assert_allclose(mTau, (1 - np.exp(-x / 9)) / (x * (0.182 + 0.124 * np.exp(-x / 9))), rtol=1e-11)
    # h_inf is a Boltzmann with half point -55 mV and slope 6.2 mV
    np.testing.assert_allclose(O.nav1p2_rates(np.array([-55.0]))[2], 0.5, rtol=1e-15)
    assert np.all(np.diff(hInf) < 0) and np.all(np.diff(mInf) > 0)
    assert np.all(hTau > 0)


def test_kv1p1_anchor_points():
    mInf, mTau, hInf, hTau = O.kv1p1_rates(np.array([-30.5, -76.56, -30.0, -160.56]))
    assert mInf[0] == 0.5 and hInf[2] == 0.5
    assert mTau[1] == 15.0 and hTau[3] == 7500.0


def test_cav1p2_anchor_points():
    """Carlin's L-type model after the -10 mV shift: Boltzmann half points at
    -20 mV (m) and -70 mV (h); tau midpoints at -15 mV (m: 5 + 20/2) and
    -30 mV (h: 20 + 50/2)."""
    mInf, mTau, hInf, hTau = O.cav1p2_rates(np.array([-20.0, -15.0, -70.0, -30.0]))
    assert mInf[0] == 0.5 and hInf[2] == 0.5
    assert mTau[1] == 15.0 and hTau[3] == 45.0
    V = np.linspace(-100, 40, 141)
    mInf, mTau, hInf, hTau = O.cav1p2_rates(V)
    assert np.all(np.diff(mInf) > 0) and np.all(np.diff(hInf) < 0)
    assert np.all((mTau > 5) & (mTau < 25)) and np.all((hTau > 20) & (hTau < 70))


def test_init_gates_match_steady_state():
    V = np.array([-0.07, -0.03, 0.01])
    for kind, rates in (("Nav1p2", O.nav1p2_rates), ("Kv1p1", O.kv1p1_rates), ("Cav1p2", O.cav1p2_rates)):
        m0, h0 = init_gates(kind, V)
        mInf, _, hInf, _ = rates(V * 1000)
        np.testing.assert_allclose(m0, mInf, rtol=1e-14)
        np.testing.assert_allclose(h0, hInf, rtol=1e-14)
        ch = O.OracleChannel(kind, 1e-17)
        ch.init(V)
        np.testing.assert_array_equal(ch.m, m0)


def test_gate_update_fixed_point_and_limit():
    class P:
        dt = 1e-4
    ch = O.OracleChannel("Kv1p1", 1e-17)
    V = np.array([-0.05, 0.02])
    mInf, _, hInf, _ = O.kv1p1_rates(V * 1000)
    ch.m, ch.h = mInf.copy(), hInf.copy()
    ch.run(V, P)
    np.testing.assert_allclose(ch.m, mInf, rtol=1e-14)
    P.dt = 1e6
    ch.m, ch.h = np.zeros(2), np.ones(2)
    ch.run(V, P)
    np.testing.assert_allclose(ch.m, mInf, rtol=1e-6)
    np.testing.assert_allclose(ch.P, ch.m * ch.h ** 2)


def test_channel_flux_conserves_mass_and_stays_stable():
    p = Parameters.preset("tiny_channels")
    cells = Cells.make_world(p)
    draw_channel_state(p, cells.n_cells, cells.n_mems)
    o = O.OracleSim(cells, p.copy(), "sparse")
    o.init_core()
    o.run_sim_core("init", n_steps=50)
    assert np.all(np.isfinite(o.vm))
    f = o.channels[0].chan_flux
    tot = np.sum(cells.cell_vol * o.msum(f * cells.mem_sa) / cells.cell_vol) \
        + np.sum(p.cell_height * cells.delta ** 2 * o.divmap(-f) / (p.cell_height * cells.delta ** 2))
    assert abs(tot) < 1e-12 * np.sum(np.abs(f * cells.mem_sa))


def test_seeded_channel_state_required():
    p = Parameters.preset("tiny_channels")
    cells = Cells.make_world(p)
    with pytest.raises(ValueError):
        O.OracleSim(cells, p.copy(), "sparse").init_core()


# ---------------------------------------------------------------------------
# against the reference's own methods (tests/golden/make_golden_vectors.py:
# Nav1p2/Kv1p1/Cav1p2._calculate_state and ChannelsABC.update_mh run on seeded
# inputs, including the 0/0 points of the rate laws where the reference gives NaN)
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def chan_gold():
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_channels.npz")
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.mark.parametrize("tag,fn", [("nav", O.nav1p2_rates), ("kv", O.kv1p1_rates), ("cav", O.cav1p2_rates)])
def test_rate_laws_match_reference(chan_gold, tag, fn):
    import warnings
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        got = fn(chan_gold["V"].copy())
    for k, g in zip(("_mInf", "_mTau", "_hInf", "_hTau"), got):
        ref = chan_gold[tag + k]
        assert np.array_equal(np.isnan(g), np.isnan(ref)), (tag, k)
        ok = ~np.isnan(ref)
        np.testing.assert_array_equal(np.asarray(g)[ok], ref[ok], err_msg=f"{tag}{k}")


def test_gate_update_matches_reference(chan_gold):
    g = chan_gold
    ch = O.OracleChannel("Kv1p1", 1e-17)
    rates = (g["upd_in_mInf"], g["upd_in_mTau"], g["upd_in_hInf"], g["upd_in_hTau"])
    ch._rates = lambda V: rates
    ch.m, ch.h = g["upd_inm"].copy(), g["upd_inh"].copy()

    class P:
        dt = float(g["upd_dt"])
    ch.run(np.zeros(len(ch.m)), P)
    np.testing.assert_array_equal(ch.m, g["upd_m"])
    np.testing.assert_array_equal(ch.h, g["upd_h"])
