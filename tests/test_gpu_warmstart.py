"""The learned warm start on the GPU: init-phase labels from the plain
method, the network trained by the library's kernels, and what it gains over
the default start on held-out clusters (same stepper, same residual)."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def test_warm_start_gains_over_default_start():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from betse_amd import warmstart as ws
    kw = dict(cells_range=(60, 200), grid_range=(16, 24))
    say("case: warm start, 96 training clusters of 60-200 cells, labels from 20000 + 10000 steps")
    data, stats, base = ws.make_dataset(96, seed=77, batch=96, horizon=20_000, window=100, log=say, **kw)
    say(f"labels: {stats}")
    assert sum(d["ok"] for d in data) >= 48
    starter, hist, _ = ws.train(data, epochs=120, clusters_per_batch=16, log=say)
    assert hist[-1] < 0.1 * hist[0]
    say("case: warm start evaluation, 32 held-out clusters")
    ev = ws.evaluate(starter, 32, seed=4242, max_steps=40_000, window=100, **kw)
    say(f"evaluation: {ev}")
    # every prediction is usable and the run stays stable
    assert ev["warm_injected"] >= 28 and ev["restarts"] == 0
    # the plain stepper from the prediction reaches the residual for at least as
    # many clusters, in fewer steps overall
    assert ev["warm_converged"] >= ev["cold_converged"]
    assert ev["warm_total_steps"] < 0.8 * ev["cold_total_steps"]
    assert ev["warm_seconds"] < ev["cold_seconds"]


def test_warm_start_gains_at_spec_cluster_sizes():
    """The learned warm start with the workload's cluster sizes (500-2000
    cells, grids 48-88, Dm and NaK rate over two decades): 64 training clusters
    and 16 held-out ones (the full run uses 4096 and 256)."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from betse_amd import warmstart as ws
    say("case: warm start at spec sizes, 64 training clusters of 500-2000 cells")
    data, stats, base = ws.make_dataset(64, seed=1000, batch=64, horizon=20_000, flux_mode="fp32", log=say)
    say(f"labels: {stats}")
    assert sum(d["ok"] for d in data) >= 32
    starter, hist, _ = ws.train(data, epochs=60, clusters_per_batch=16, log=say)
    assert hist[-1] < 0.2 * hist[0]
    say("case: warm start evaluation at spec sizes, 16 held-out clusters")
    ev = ws.evaluate(starter, 16, seed=999_000, max_steps=60_000, window=100, flux_mode="fp32")
    say(f"evaluation: {ev}")
    assert ev["warm_injected"] >= 12 and ev["restarts"] == 0
    assert ev["warm_converged"] >= ev["cold_converged"]
    assert ev["warm_total_steps"] < 0.8 * ev["cold_total_steps"]
