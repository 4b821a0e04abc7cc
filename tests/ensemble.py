"""Ensemble statistics of the observables a perveance sweep reads
(test infrastructure): beamlet current, theta95 divergence, saddle-point
potential and accel-grid impingement, as the reference computes them per step
(Python/physics_engine.py:1482-1497, :1655-1676).

A run is a free-running simulation of one seed; its statistic per observable
is the mean over a window of steps (theta95 over the steps that define it,
more than 5 crossings).  Two ensembles (several seeds each) agree when, for
every observable, the mean over seeds of the paired per-seed differences is
within max(3 standard errors of that mean, the observable's stated tolerance
relative to its magnitude).
"""
import math

import numpy as np

from conftest import c1_params

OBSERVABLES = ("current_A", "theta95_deg", "saddle_V", "impingement")
# stated tolerances, relative to the ensemble mean's magnitude
TOLERANCE = {"current_A": 0.02, "theta95_deg": 0.02, "saddle_V": 1e-3, "impingement": 0.02}


def case_params(seed):
    """Two-grid case at 128 x 48 with a dense preload, so that more than five
    ions cross the exit plane per step and theta95 is defined."""
    return c1_params(nx=128, ny=48, ions=50000, electrons=5000, seed=seed)


def record(sim, result, transmitted, lost):
    """One step's observables: transmitted ions (the step's beamlet current,
    transmitted * q * w / dt), theta95, saddle potential, ions lost to a grid."""
    return {"current_A": transmitted * sim.q * sim.macro_weight / sim.dt,
            "theta95_deg": result[2], "saddle_V": result[1], "impingement": lost}


def window_means(rows, first):
    out = {}
    for k in OBSERVABLES:
        v = np.array([r[k] for r in rows[first:]], np.float64)
        out[k] = float(np.nanmean(v)) if np.any(~np.isnan(v)) else float("nan")
    return out


def compare(a, b):
    """a, b: lists (one per seed, same seeds in the same order) of window
    means.  Returns {observable: (mean_a, mean_b, diff, bound, ok)}."""
    out = {}
    for k in OBSERVABLES:
        xa = np.array([s[k] for s in a])
        xb = np.array([s[k] for s in b])
        d = xa - xb
        n = len(d)
        se = float(np.std(d, ddof=1) / math.sqrt(n)) if n > 1 else 0.0
        mag = max(abs(float(np.mean(xa))), abs(float(np.mean(xb))))
        bound = max(3.0 * se, TOLERANCE[k] * mag)
        diff = float(np.mean(d))
        out[k] = (float(np.mean(xa)), float(np.mean(xb)), diff, bound, abs(diff) <= bound)
    return out


def run_oracle(seed, steps, mode="ref32"):
    from oracle.sim import OracleSimulator
    p = case_params(seed)
    o = OracleSimulator(mode=mode)
    o.m_e = o.m_ion / 200.0
    o.build_domain(p)
    rows = []
    for _ in range(steps):
        res = o.step(p)
        rows.append(record(o, res, o.transmitted_ions_step, o.lost_to_grid_step))
    return rows


def run_device(seed, steps):
    from thruster_cx_code_amd import DigitalTwinSimulator
    p = case_params(seed)
    g = DigitalTwinSimulator()
    g.m_e = g.m_ion / 200.0
    g.build_domain(p)
    rows = []
    for _ in range(steps):
        res = g.step(p)
        rows.append(record(g, res, g.transmitted_ions_step, g.lost_to_grid_step))
    return rows
