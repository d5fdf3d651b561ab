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


def test_step_bytes_match_the_scope_tables_figures():
    """Per-unit bytes of the step against the figures the hot-path scope table
    of the survey lists for its configurations (C1/C3: 64/196/176 bytes per
    membrane/cell/square; C2 with Nav/Kv m,h gates and two per-cell
    multipliers: 136/244/200; C5 with three channels and eight substances:
    168/508/200) and its B_step totals (C1 0.643 MB, C2 37.8 MB)."""
    c1 = bench.b_step(350, 2100, 2500, 5)
    assert (c1["b_M"], c1["b_C"], c1["b_G"]) == (64, 196, 176)
    assert abs(c1["bytes"] / 1e6 - 0.643) < 0.001
    c2 = bench.b_step(23000, 140000, 65536, 6, n_gates=4, n_mult=2)
    assert (c2["b_M"], c2["b_C"], c2["b_G"]) == (136, 244, 200)
    assert abs(c2["bytes"] / 1e6 - 37.8) < 0.05
    c5 = bench.b_step(100000, 600000, 262144, 6, n_gates=6, n_mult=3, n_species=8)
    assert (c5["b_M"], c5["b_C"], c5["b_G"]) == (168, 508, 200)
    assert abs(c5["bytes"] / 1e6 - 204) < 1


def test_workload_bytes_follow_the_physics():
    """The workload's B_step takes its ions, channels and substances from the
    parameters: the large preset has Nav/Kv (four gates, two multipliers)."""
    p = Parameters.preset("tiny_channels")
    cells = Cells.make_world(p)
    b = bench.workload_b_step([(p, cells)])
    assert (b["b_M"], b["b_C"], b["b_G"]) == (136, 244, 200)
    assert b["bytes"] == cells.n_mems * 136 + cells.n_cells * 244 + cells.X.size * 200
    assert np.isfinite(b["per_cell"])


def test_measured_traffic_reads_the_profile_summary():
    t = bench.measured_traffic("large", "fp64")
    assert t is not None and set(t["kernels"]) >= {"mem_cell", "env_field", "env_exchange"}
    assert t["bytes"] == sum(t["kernels"].values())
    assert bench.measured_traffic("no_such_workload", "fp64") is None
