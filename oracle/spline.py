"""User cross-section tables (test infrastructure, oracle side): the cubic
smoothing spline of log10(sigma) in log10(E) that the reference fits with
SciPy's ``UnivariateSpline`` (Python/run_simulation_from_config.py:124-133,
Python/main.py:270-281) and evaluates in ``lookup_user_cs``
(Python/physics_engine.py:475-484).

``bspline_eval`` restates FITPACK's evaluation (``splev``: the knot interval
l with t[l] <= x < t[l+1], capped at the last interval, and ``fpbspl``'s
Cox-de Boor recursion, in FITPACK's operation order), which the device
evaluator repeats; it is pinned against ``scipy.interpolate.splev``
(tests/test_oracle_spline.py).  ``table_value`` applies the reference's
clipping of log10(E) to the table's energy range; 10**y is evaluated as
det_exp(y ln 10) (oracle.dmath), which is deterministic on both sides and
within an ulp or two of NumPy's power.
"""
import numpy as np

from .dmath import det_exp, det_log, INV_LN10

LN10 = 2.302585092994045684


def bspline_eval(t, c, k, x):
    """Value of the B-spline (t, c, k) at x (array), FITPACK splev/fpbspl."""
    t = np.asarray(t, np.float64)
    c = np.asarray(c, np.float64)
    x = np.atleast_1d(np.asarray(x, np.float64))
    n = len(t)
    lo, hi = k, n - k - 2
    l = np.searchsorted(t, x, side="right") - 1     # largest l with t[l] <= x
    l = np.clip(l, lo, hi)
    h = np.zeros((k + 1, len(x)))
    h[0] = 1.0
    for j in range(1, k + 1):
        hh = h[:j].copy()
        h[0] = 0.0
        for i in range(1, j + 1):
            li = l + i
            lj = li - j
            tl, tj = t[li], t[lj]
            same = tl == tj
            d = np.where(same, 1.0, tl - tj)
            f = hh[i - 1] / d
            h[i - 1] = np.where(same, h[i - 1], h[i - 1] + f * (tl - x))
            h[i] = np.where(same, 0.0, f * (x - tj))
    s = np.zeros(len(x))
    for j in range(k + 1):
        s = s + c[l - k + j] * h[j]
    return s


def spline_table(spline, energy):
    """(t, c, k, log10 E_min, log10 E_max) of a fitted table."""
    t, c, k = spline._eval_args
    e = np.asarray(energy, np.float64)
    return (np.asarray(t, np.float64), np.asarray(c, np.float64), int(k),
            float(np.log10(max(e[0], 1e-30))), float(np.log10(e[-1])))


def table_value(table, E):
    """10 ** spline(clip(log10(max(E, 1e-30)), lo, hi)) with deterministic
    log and exp (oracle.dmath)."""
    t, c, k, lo, hi = table
    E = np.maximum(np.asarray(E, np.float64), 1e-30)
    le = det_log(E) * INV_LN10
    le = np.minimum(np.maximum(le, lo), hi)
    return det_exp(bspline_eval(t, c, k, le) * LN10)


def lookup(user_cs, prefix):
    """The table the reference would use for a prefix ('CX', 'SEE'): the first
    entry whose label starts with it and has a fitted spline."""
    for label, ds in user_cs.items():
        if label.startswith(prefix) and ds.get('spline') is not None:
            return spline_table(ds['spline'], ds['energy'])
    return None
