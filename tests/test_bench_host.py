"""Host-side checks of bench.py (no GPU): the one-GPU refusal and the
algorithmic-byte model of the step kernels."""

import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402
from betse_amd.cells import Cells  # noqa: E402
from betse_amd.parameters import Parameters  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in r.stdout + r.stderr


def test_algorithmic_bytes_counts_each_array_once():
    """The model of the tiny world, term by term: every array element the step
    kernels must touch, at its storage width, reads and writes separately."""
    p = Parameters.preset("tiny")
    cells = Cells.make_world(p)
    C, M, G2 = cells.n_cells, cells.n_mems, cells.X.size
    n_in = len(cells.envInds_inClust)
    ni = 5
    b = bench.algorithmic_bytes(cells, ni, ca=False, flux_mode="fp64")
    # membranes: 5 int32 index arrays, mem_sa, gjopen, Dm per ion; flux writes
    mem = M * (5 * 4 + 8 + 8 + 8 * ni) + M * 8 * ni
    # cells: c_cell and c_lag read and written per ion, Vmem read and written,
    # cell_vol, diviterm, mem_ptr; env gathers of the squares under the cluster
    mem += C * (8 * ni * 4 + 16 + 16 + 4) + n_in * 8 * ni
    assert b["mem_cell"] == mem
    assert b["env_field"] == G2 * (8 + 24 * ni)
    n_live = int(np.count_nonzero(np.diff(cells.sq_ptr)))        # squares with membranes
    assert 0 < n_live < G2
    assert b["env_exchange"] == n_live * (16 * ni + 21) + M * 8 * ni
    assert b["step"] == b["mem_cell"] + b["env_field"] + b["env_exchange"]
    # channels add the gates (read and write), the channel fluxes, the per-cell
    # multipliers and rates, and the rates the env kernel writes from Vmem
    bc = bench.algorithmic_bytes(cells, ni, ca=False, flux_mode="fp64", n_ch=2)
    assert bc["mem_cell"] - b["mem_cell"] == M * 2 * 40 + C * 2 * 40 + n_in * 8 * 2
    assert bc["env_field"] - b["env_field"] == C * (8 + 32 * 2)
    # a D_env that is one value per ion is not loaded by the env kernel
    bd = bench.algorithmic_bytes(cells, ni, ca=False, flux_mode="fp64", denv_loaded=0)
    assert b["env_field"] - bd["env_field"] == G2 * 8 * ni
    # per-cluster uniform Dm and gjopen are not loaded per membrane
    bu = bench.algorithmic_bytes(cells, ni, ca=False, flux_mode="fp64", dm_loaded=0, gj_loaded=False)
    assert b["mem_cell"] - bu["mem_cell"] == M * 8 * (ni + 1)
    assert np.isfinite(float(b["step"]))
