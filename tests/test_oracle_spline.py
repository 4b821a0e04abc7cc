"""Pins of the oracle's user cross-section evaluation (CPU): FITPACK's
splev through SciPy, and the reference's clipping of the energy range."""
import os

import numpy as np
import pytest
from scipy.interpolate import UnivariateSpline, splev

from oracle import spline as SP

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "config")


def _table_fit(s=0.0):
    raw = np.loadtxt(os.path.join(GOLDEN, "sample_cs.txt"), comments="#")
    e, cs = raw[:, 0], raw[:, 1]
    order = np.argsort(e)
    e, cs = e[order], cs[order]
    sp = UnivariateSpline(np.log10(np.maximum(e, 1e-30)), np.log10(np.maximum(cs, 1e-50)), s=s, k=3)
    return e, cs, sp


@pytest.mark.parametrize("s", [0.0, 0.05])
def test_bspline_eval_matches_scipy_splev(s):
    e, cs, sp = _table_fit(s)
    t, c, k = sp._eval_args
    x = np.concatenate([np.linspace(np.log10(e[0]), np.log10(e[-1]), 5001), np.log10(e), t[k:len(t) - k]])
    got = SP.bspline_eval(t, c, k, x)
    want = splev(x, (t, c, k))
    assert np.max(np.abs(got - want)) <= 4e-15 * max(1.0, np.max(np.abs(want)))


def test_bspline_eval_random_splines():
    r = np.random.default_rng(3)
    for trial in range(20):
        xs = np.sort(r.uniform(0, 10, 12 + trial))
        ys = r.normal(size=len(xs))
        sp = UnivariateSpline(xs, ys, s=float(r.uniform(0, 2)), k=3)
        t, c, k = sp._eval_args
        x = r.uniform(xs[0], xs[-1], 400)
        assert np.allclose(SP.bspline_eval(t, c, k, x), splev(x, (t, c, k)), rtol=0, atol=1e-13)


def test_interpolating_table_reproduces_points_and_clips():
    e, cs, sp = _table_fit(0.0)
    table = SP.spline_table(sp, e)
    got = SP.table_value(table, e)
    assert np.max(np.abs(got / cs - 1.0)) < 1e-12           # s = 0 interpolates
    below = SP.table_value(table, np.array([1e-3, 0.5]))
    above = SP.table_value(table, np.array([1e4, 1e7]))
    assert np.allclose(below, cs[0], rtol=1e-12) and np.allclose(above, cs[-1], rtol=1e-12)
    # the reference's own expression on the same energies
    ref = 10.0 ** sp(np.clip(np.log10(np.maximum(e * 1.37, 1e-30)), np.log10(e[0]), np.log10(e[-1])))
    assert np.allclose(SP.table_value(table, e * 1.37), ref, rtol=1e-13)


def test_lookup_prefers_first_fitted_label():
    e, cs, sp = _table_fit(0.0)
    store = {"Custom": {"energy": e, "cs": cs, "spline": sp},
             "CX_a": {"energy": e, "cs": cs, "spline": None},
             "CX_b": {"energy": e, "cs": cs * 2, "spline": sp}}
    assert SP.lookup(store, "SEE") is None
    t = SP.lookup(store, "CX")
    assert t is not None and t[2] == 3
