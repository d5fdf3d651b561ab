"""CSV exports of a finished phase, in the reference's file format.

The reference writes its CSV exports with ``npcsv.write_csv``
(``betse/lib/numpy/npcsv.py``): one header line of comma-joined column names
(names holding a comma are double-quoted), then the columns stacked as rows,
every value as ``%.18e``, comma-delimited, no comment prefix. Two exporters of
``SimPipesExportCSV`` (``betse/science/pipe/export/pipeexpcsv.py``) are
mirrored here for the stored time series of a ``Simulator``:

* ``export_cell_series`` -- the single-cell series ``ExportedData.csv``
  (``pipeexpcsv.py:60-190``): time, Vmem (mV, the cell average of the
  membrane voltages), the Goldman Vmem column, the NaK-ATPase rate, cell
  concentrations, membrane diffusivities and the membrane current at the
  cell's first membrane; physics this path does not model (Goldman
  potential, hydrostatic and osmotic pressure, displacement) are zero
  columns, as the reference writes them when disabled;
* ``export_cells_vmem`` -- one ``Vmem2D_<n>.csv`` per sampled step with the
  cell centres in um and Vmem in mV (``pipeexpcsv.py:251-390``).
"""

from __future__ import annotations

import os
from typing import Dict, Sequence

import numpy as np


def write_csv(filename: str, columns: Dict[str, Sequence[float]]) -> None:
    """Columns of equal length -> CSV in the reference's layout."""
    names, data, n = [], [], None
    for name, values in columns.items():
        if '"' in name or "\n" in name:
            raise ValueError(f"column name {name!r} holds a reserved character")
        values = np.asarray(values, dtype=np.float64)
        if values.ndim != 1:
            raise ValueError(f"column {name!r} is not one-dimensional")
        if n is not None and len(values) != n:
            raise ValueError(f"column {name!r} has length {len(values)}, the others {n}")
        n = len(values)
        names.append(f'"{name}"' if "," in name else name)
        data.append(values)
    parent = os.path.dirname(os.path.abspath(filename))
    os.makedirs(parent, exist_ok=True)
    np.savetxt(filename, np.column_stack(data), fmt="%.18e", header=",".join(names), delimiter=",", comments="")


def _cell_average(cells, x_mem: np.ndarray) -> np.ndarray:
    return np.add.reduceat(x_mem, cells.mem_ptr[:-1]) / cells.num_mems


def export_cell_series(sim, cell_index: int, filename: str) -> None:
    """``ExportedData.csv`` of one cell from the stored time series."""
    cells = sim.cells
    n = len(sim.time)
    zeros = np.zeros(n)
    m0 = int(cells.mem_ptr[cell_index])            # cell_to_mems[cell_index][0]
    cols = {"time_s": sim.time,
            "Vmem_mV": [1e3 * _cell_average(cells, vm)[cell_index] for vm in sim.vm_time],
            "Goldman_Vmem_mV": zeros,
            "NaK-ATPase_Rate_mol/m2s": [r[m0] for r in sim.rate_NaKATP_time]}
    for i, name in enumerate(sim.names):
        cols[f"cell_{name}_mmol/L"] = [cc[i][cell_index] for cc in sim.cc_time]
    for i, name in enumerate(sim.names):
        cols[f"Dm_{name}_m2/s"] = [dd[i][m0] for dd in sim.dd_time]
    cols["I_A/m2"] = [im[m0] for im in sim.I_mem_time]
    cols["HydroP_Pa"] = zeros
    cols["OsmoP_Pa"] = zeros
    cols["Displacement_um"] = zeros
    write_csv(filename, cols)


def export_cells_vmem(sim, dirname: str, prefix: str = "Vmem2D_") -> None:
    """One CSV per sampled step: cell centres (um) and the cell-average Vmem (mV)."""
    cells = sim.cells
    x = cells.cell_centres[:, 0] * 1e6
    y = cells.cell_centres[:, 1] * 1e6
    for t, vm in enumerate(sim.vm_time):
        write_csv(os.path.join(dirname, f"{prefix}{t}.csv"),
                  {"x [um]": x, "y [um]": y, "Vmem [mV]": 1e3 * _cell_average(cells, vm)})
