"""Ensemble tolerance of the device step against the oracle: both free-run
several seeds from the same preload (the device's Poisson solution differs
from SuperLU's by about 1e-11, so trajectories are not kept identical), and
the sweep observables -- beamlet current, theta95 divergence, saddle-point
potential, accel-grid impingement (Python/physics_engine.py:1482-1497,
:1655-1676) -- agree within max(3 standard errors, stated tolerance)
(tests/ensemble.py)."""
import pytest

from ensemble import compare, run_device, run_oracle, window_means

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def test_device_ensemble_matches_oracle():
    seeds = [1, 2, 3, 4]
    dev, ora = [], []
    for s in seeds:
        print(f"[gpu] ensemble seed {s}", flush=True)
        dev.append(window_means(run_device(s, 60), 10))
        ora.append(window_means(run_oracle(s, 60, "ref32"), 10))
    for k, (md, mo, d, bound, ok) in compare(dev, ora).items():
        print(f"  {k}: device {md:.6g} oracle {mo:.6g} diff {d:.3g} bound {bound:.3g}", flush=True)
        assert ok, k
