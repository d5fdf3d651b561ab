"""CSV exports in the reference's layout (header of column names, %.18e, commas)."""

import numpy as np

from betse_amd import export
from betse_amd.cells import Cells
from betse_amd.parameters import Parameters
from betse_amd.sim import HostState, SimPhase, SimPhaseKind


def _stored_state():
    p = Parameters.preset("tiny")
    cells = Cells.make_world(p)
    st = HostState()
    ph = SimPhase(SimPhaseKind.INIT, p, cells)
    st.init_core(ph)
    st.init_dynamics(ph)
    st.clear_storage()
    rng = np.random.RandomState(0)
    for t in range(3):
        st.time.append(0.01 * t)
        st.vm_time.append(rng.uniform(-0.07, -0.05, cells.n_mems))
        st.rate_NaKATP_time.append(rng.rand(cells.n_mems))
        st.cc_time.append(st.cc_cells * (1 + 0.01 * t))
        st.dd_time.append(st.Dm_cells.copy())
        st.I_mem_time.append(rng.randn(cells.n_mems))
    return st, cells


def test_cell_series_columns_and_values(tmp_path):
    st, cells = _stored_state()
    f = tmp_path / "ExportedData.csv"
    export.export_cell_series(st, 2, str(f))
    header = f.read_text().splitlines()[0].split(",")
    assert header[:4] == ["time_s", "Vmem_mV", "Goldman_Vmem_mV", "NaK-ATPase_Rate_mol/m2s"]
    assert header[4:4 + len(st.names)] == [f"cell_{n}_mmol/L" for n in st.names]
    assert header[-4:] == ["I_A/m2", "HydroP_Pa", "OsmoP_Pa", "Displacement_um"]
    data = np.loadtxt(f, delimiter=",", skiprows=1)
    assert data.shape == (3, len(header))
    m0 = cells.mem_ptr[2]
    vm_cell = [1e3 * st.vm_time[t][cells.mem_ptr[2]:cells.mem_ptr[3]].mean() for t in range(3)]
    assert np.allclose(data[:, 1], vm_cell, rtol=1e-14)
    assert np.array_equal(data[:, 3], [r[m0] for r in st.rate_NaKATP_time])
    assert np.array_equal(data[:, 4], [cc[0][2] for cc in st.cc_time])
    # %.18e round-trips doubles exactly
    assert np.array_equal(data[:, 0], st.time)


def test_cells_vmem_files(tmp_path):
    st, cells = _stored_state()
    export.export_cells_vmem(st, str(tmp_path / "Vmem2D_TextExport"))
    files = sorted((tmp_path / "Vmem2D_TextExport").iterdir())
    assert [p.name for p in files] == ["Vmem2D_0.csv", "Vmem2D_1.csv", "Vmem2D_2.csv"]
    assert files[0].read_text().splitlines()[0] == "x [um],y [um],Vmem [mV]"
    d = np.loadtxt(files[1], delimiter=",", skiprows=1)
    assert d.shape == (cells.n_cells, 3)
    assert np.allclose(d[:, 0], cells.cell_centres[:, 0] * 1e6)


def test_comma_names_are_quoted(tmp_path):
    f = tmp_path / "a.csv"
    export.write_csv(str(f), {"a,b": [1.0, 2.0], "c": [3.0, 4.0]})
    assert f.read_text().splitlines()[0] == '"a,b",c'
