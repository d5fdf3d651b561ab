"""Phase time vector and its sampled steps (``sim.py:2455-2462``)."""

import numpy as np
import pytest

from betse_amd.sim import phase_time_steps


def test_sampled_steps_every_t_resample():
    # 2,000 steps sampled every 100: indices 100, 200, ..., 1900 (the loop stops
    # once fewer than t_resample steps remain), 19 samples
    ts, sampled = phase_time_steps(2000, 1e-3, 100)
    assert len(ts) == 2000
    idx = sorted(i for i, t in enumerate(ts) if t in sampled)
    assert idx == list(range(100, 2000, 100))


def test_short_phase_has_no_samples():
    ts, sampled = phase_time_steps(20, 1e-3, 100)
    assert len(ts) == 20 and not sampled


def test_time_vector_is_linspace_over_the_phase():
    ts, _ = phase_time_steps(50, 1e-3, 10)
    assert ts[0] == 0.0 and ts[-1] == pytest.approx(0.05)
    assert np.all(np.diff(ts) > 0)


@pytest.mark.parametrize("tr", [0, 0.5, -3])
def test_sub_step_resampling_is_refused(tr):
    # below one step the reference's loop would never advance
    with pytest.raises(ValueError):
        phase_time_steps(100, 1e-3, tr)
