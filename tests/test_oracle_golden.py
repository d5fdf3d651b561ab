"""The oracle against vectors written by the reference's own functions
(``tests/golden/make_golden_vectors.py``)."""

from types import SimpleNamespace

import numpy as np
import pytest

from oracle import betse_oracle as O

P = SimpleNamespace(F=96485, R=8.314, q=1.602e-19, kb=1.3806e-23, deltaGATP=-37000, cATP=1.5, cADP=0.1,
                    cPi=0.1, KmNK_Na=12.0, KmNK_K=0.2, KmNK_ATP=0.5, KmCa_Ca=1.0e-3, KmCa_ATP=0.5,
                    alpha_NaK=1.0e-7, alpha_Ca=5.0e-8, dt=1.0e-3, cell_height=10.0e-6)


def same(a, b):
    """Bit-identical, NaN-aware."""
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape
    assert np.array_equal(a, b, equal_nan=True), float(np.nanmax(np.abs(a - b)))


@pytest.mark.parametrize("n", range(4))
def test_finite_differences(golden, n):
    F, G, d = golden[f"fd{n}_F"], golden[f"fd{n}_G"], float(golden[f"fd{n}_delta"])
    gx, gy = O.fd_gradient(F, d)
    same(gx, golden[f"fd{n}_grad_x"])
    same(gy, golden[f"fd{n}_grad_y"])
    same(O.fd_diff(F, d, axis=0), golden[f"fd{n}_diff0"])
    same(O.fd_diff(F, d, axis=1), golden[f"fd{n}_diff1"])
    same(O.fd_divergence(F, G, d, d), golden[f"fd{n}_div"])
    same(O.fd_integrator(F, 0.5), golden[f"fd{n}_integ"])


def test_electroflux(golden):
    g = {k[4:]: golden[k] for k in golden if k.startswith("ghk_")}
    v = g["v"].copy()
    f = O.electroflux(g["cA"], g["cB"], g["Dc"], g["d"], g["z"].copy(), v, 310.0, P)
    same(f, g["flux"])
    same(v, g["v_after"])          # in-place +1e-25 on the voltage argument (Q5)


def test_pumps(golden):
    g = {k[4:]: golden[k] for k in golden if k.startswith("nak_")}
    fNa, fK, _ = O.pumpNaKATP(g["cNai"], g["cNao"], g["cKi"], g["cKo"], g["Vm"], 310.0, P, 1)
    same(fNa, g["fNa"])
    same(fK, g["fK"])
    ci, co = golden["ca_cCai"].copy(), golden["ca_cCao"].copy()
    fCa = O.pumpCaATP(ci, co, golden["ca_Vm"], 310.0, P, 1)
    same(fCa, golden["ca_fCa"])
    same(ci, golden["ca_cCai_after"])
    same(co, golden["ca_cCao_after"])


def test_nernst_planck(golden):
    g = {k[3:]: golden[k] for k in golden if k.startswith("np_")}
    z = np.zeros(g["c"].shape)
    fx, fy = O.nernst_planck_flux(g["c"], g["gcx"], g["gcy"], g["gvx"], g["gvy"], z, z, g["D"],
                                  int(g["z"]), 310.0, P)
    same(fx, g["fx"])
    same(fy, g["fy"])


def test_no_negs(golden):
    x = golden["nn_in"].copy()
    same(O.no_negs(x), golden["nn_out"])
    with pytest.raises(O.OracleUnstable):
        O.no_negs(np.array([1.0, np.nan]))


def _uc_cells(golden):
    m2c = golden["uc_m2c"]
    C = len(golden["uc_cell"])
    G = int(golden["uc_G"])
    ptr = np.zeros(C + 1, dtype=np.int64)
    ptr[1:] = np.cumsum(np.bincount(m2c, minlength=C))
    return SimpleNamespace(mem_to_cells=m2c, mem_sa=golden["uc_mem_sa"], cell_vol=golden["uc_cell_vol"],
                           map_mem2ecm=golden["uc_map"], mem_ptr=ptr, delta=float(golden["uc_delta"]),
                           X=np.zeros((G, G)), n_cells=C, n_mems=len(m2c))


@pytest.mark.parametrize("mode", ["dense", "sparse"])
def test_update_co(golden, mode):
    """update_Co + div_env as the oracle's update_all_concs composes them."""
    cells = _uc_cells(golden)
    o = O.OracleSim.__new__(O.OracleSim)
    o.cells, o.mode = cells, mode
    if mode == "dense":
        C, M, G2 = cells.n_cells, cells.n_mems, cells.X.size
        o.M_sum_mems = np.zeros((C, M))
        o.M_sum_mems[cells.mem_to_cells, np.arange(M)] = 1.0
        o.M_divmap_mem2ecm = np.zeros((G2, M))
        for m, k in enumerate(cells.map_mem2ecm):
            o.M_divmap_mem2ecm[k, m] += cells.mem_sa[m]
    flux = golden["uc_flux"]
    delta_cells = o.msum(flux * cells.mem_sa) / cells.cell_vol
    cell = golden["uc_cell"] + delta_cells * P.dt
    env = golden["uc_env"] + (o.divmap(-flux) / (P.cell_height * cells.delta ** 2)) * P.dt
    np.testing.assert_allclose(cell, golden["uc_cell_out"], rtol=1e-15, atol=0)
    np.testing.assert_allclose(cell[cells.mem_to_cells], golden["uc_mem_out"], rtol=1e-15, atol=0)
    np.testing.assert_allclose(env, golden["uc_env_out"], rtol=1e-15, atol=0)


@pytest.mark.parametrize("mode", ["dense", "sparse"])
def test_hh_decomp(golden, mode):
    JJx, JJy, dl = golden["hh_JJx"], golden["hh_JJy"], float(golden["hh_delta"])
    ny, nx = JJx.shape
    o = O.OracleSim.__new__(O.OracleSim)
    o.cells = SimpleNamespace(X=np.zeros((ny, nx)), delta=dl)
    o.mode = mode
    A = O.make_laplacian_value(ny, nx, dl)
    if mode == "dense":
        o.lapENVinv = np.linalg.pinv(A)
    else:
        import scipy.sparse.linalg as sla
        o._lap_solve = sla.factorized(O.make_laplacian_value(ny, nx, dl, sparse=True).tocsc())
    res = o.HH_Decomp(JJx, JJy, bounds={"L": 0.0, "R": 0.0, "T": 0.0, "B": 0.0})
    for k, r in zip(("AA", "Fx", "Fy", "BB", "Gx", "Gy"), res):
        ref = golden["hh_" + k]
        scale = np.abs(ref).max()
        np.testing.assert_allclose(r, ref, rtol=0, atol=1e-10 * scale)
