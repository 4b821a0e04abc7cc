"""Pins of the oracle's electrode-surface physics (CPU): closed forms, a
literal per-node loop, conservation, fixed points."""
import math

import numpy as np
import pytest

from oracle import surface as S

F32 = np.float32
Q = 1.602e-19
M_XE = 131.293 * 1.6605e-27
MO = dict(k=138.0, rho=10280.0, cp=250.0, emissivity=0.80, alpha=4.8e-6, Y=1.05e-4, E_th=30.0)


def _c_cell(dx=0.02, dy=0.02):
    return MO['rho'] * (dx * 1e-3) * (dy * 1e-3) * 1e-3 * MO['cp']


def _hits(n=2000, seed=0):
    r = np.random.default_rng(seed)
    v = r.normal(0.0, 3e4, (3, n)).astype(F32)
    return v[0], v[1], v[2]


def test_impact_heat_float32_order_and_closed_form():
    vx, vy, vz = _hits()
    w, C, acc = 1e3, _c_cell(), 1e7
    dT = S.impact_heat(vx, vy, vz, M_XE, w, C, acc)
    assert dT.dtype == F32
    # explicit float32 scalar arithmetic in the reference's order (:1361-1367)
    for k in range(0, len(vx), 97):
        v2 = F32(F32(F32(vx[k] * vx[k]) + F32(vy[k] * vy[k])) + F32(vz[k] * vz[k]))
        ej = F32(F32(F32(0.5 * M_XE) * v2) * F32(w))
        want = F32(F32(ej / F32(C)) * F32(acc))
        assert dT[k] == want
    # closed form E = m v^2 / 2 in float64
    v2 = vx.astype(np.float64) ** 2 + vy.astype(np.float64) ** 2 + vz.astype(np.float64) ** 2
    exact = 0.5 * M_XE * v2 * w / C * acc
    assert np.max(np.abs(dT / exact - 1.0)) < 1e-6


def test_sputter_damage_threshold_and_closed_form():
    vx, vy, vz = _hits(4000, 1)
    w = 3e5
    valid, dmg = S.sputter_damage(vx, vy, vz, M_XE, Q, MO['E_th'], MO['Y'], w)
    v2 = vx.astype(np.float64) ** 2 + vy.astype(np.float64) ** 2 + vz.astype(np.float64) ** 2
    E = 0.5 * M_XE * v2 / Q
    far = np.abs(E - MO['E_th']) > 1e-3
    assert np.array_equal(valid[far], (E > MO['E_th'])[far])
    exact = MO['Y'] * (E[valid] - MO['E_th']) * w
    rel = np.abs(dmg / exact - 1.0)
    assert np.max(rel[np.abs(E[valid] - MO['E_th']) > 1.0]) < 1e-5
    # exactly at the threshold: not damaging (strict >)
    vx0 = F32(math.sqrt(2 * MO['E_th'] * Q / M_XE))
    ok, _ = S.sputter_damage(np.array([vx0]), np.zeros(1, F32), np.zeros(1, F32), M_XE, Q,
                             MO['E_th'], MO['Y'], w)
    E0 = F32(F32(F32(0.5 * M_XE) * F32(vx0 * vx0)) / F32(Q))
    assert bool(ok[0]) == bool(E0 > F32(MO['E_th']))


def test_radiative_cooling_closed_form_and_fixed_point():
    r = np.random.default_rng(2)
    T = r.uniform(250, 3000, (12, 20)).astype(F32)
    B = r.random((12, 20)) < 0.5
    cf = S.cooling_factor(MO['emissivity'], 5.67e-8, 2 * 2e-5 * 1e-3, 5e-10, 1e7, _c_cell())
    T0 = T.copy()
    S.radiative_cooling(T, B, cf)
    assert np.array_equal(T[~B], T0[~B])
    exact = T0[B].astype(np.float64) - cf * (T0[B].astype(np.float64) ** 4 - 300.0 ** 4)
    assert np.max(np.abs(T[B] - exact) / np.abs(exact)) < 2e-6
    T3 = np.full((4, 4), 300.0, F32)
    S.radiative_cooling(T3, np.ones((4, 4), bool), cf)
    assert np.all(T3 == F32(300.0))


def test_fourier_numbers_limit():
    fx, fy = S.fourier_numbers(MO['k'], MO['rho'], MO['cp'], 5e-10, 1e7, 0.02, 0.025)
    a = MO['k'] / (MO['rho'] * MO['cp']) * 5e-3
    rx, ry = a / (0.02e-3) ** 2, a / (0.025e-3) ** 2
    s = 0.2 / max(rx, ry)
    assert fx == pytest.approx(rx * s, rel=1e-15) and fy == pytest.approx(ry * s, rel=1e-15)
    assert max(fx, fy) == pytest.approx(0.2, rel=1e-15)
    # small Fo: unchanged
    gx, gy = S.fourier_numbers(MO['k'], MO['rho'], MO['cp'], 5e-10, 1.0, 0.02, 0.02)
    assert gx == pytest.approx(MO['k'] / (MO['rho'] * MO['cp']) * 5e-10 / (0.02e-3) ** 2, rel=1e-15)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_conduction_vectorised_equals_literal_loop(seed):
    r = np.random.default_rng(seed)
    ny, nx = 9, 13
    T = r.uniform(100, 2500, (ny, nx)).astype(F32)    # values below 300 exercise the floor
    B = r.random((ny, nx)) < 0.6
    if seed == 2:
        T[4, 6] = np.nan
    a = S.conduction(T, B, 0.2, 0.13, sweeps=10)
    b = S.conduction_loop(T, B, 0.2, 0.13, sweeps=10)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_conduction_conserves_heat_in_a_closed_electrode():
    ny, nx = 16, 20
    B = np.zeros((ny, nx), bool)
    B[4:12, 5:15] = True                         # insulated block away from the edges
    r = np.random.default_rng(3)
    T = np.full((ny, nx), 300.0, F32)
    T[B] = r.uniform(400, 4000, B.sum()).astype(F32)
    out = S.conduction(T, B, 0.2, 0.2, sweeps=10)
    assert abs(out[B].astype(np.float64).sum() / T[B].astype(np.float64).sum() - 1.0) < 1e-6
    assert np.array_equal(out[~B], T[~B])
    assert out[B].std() < T[B].std()             # diffusion smooths
    U = np.full((ny, nx), 812.5, F32)
    assert np.array_equal(S.conduction(U, B, 0.2, 0.1, 10), U)


def test_clamp_and_grid_means():
    T = np.array([[250.0, 400.0], [np.nan, 299.0]], F32)
    B = np.array([[True, True], [True, False]])
    S.clamp_bound(T, B)
    assert T[0, 0] == 300.0 and T[0, 1] == 400.0 and np.isnan(T[1, 0]) and T[1, 1] == F32(299.0)
    m0 = np.array([[True, True], [False, False]])
    m1 = np.zeros((2, 2), bool)
    g = S.grid_means(T, [m0, m1], [1.0, 777.0])
    assert g[0] == np.mean(T[m0]) and g[1] == 777.0


@pytest.mark.parametrize("geometry", ["half_hole", "one_hole", "two_holes"])
def test_bow_deflection_closed_form(geometry):
    grid = {"t": 0.5, "r": 0.55}
    Ly, pitch = 2.3, 6.1447
    Tg = F32(812.25)
    d = S.bow_deflection(Tg, grid, geometry, Ly, pitch, MO['alpha'])
    L = {"half_hole": Ly - 0.55, "one_hole": 0.5 * (Ly - 1.1), "two_holes": pitch - 1.1}[geometry]
    L = max(0.2, L) * 1e-3
    exact = MO['alpha'] * (812.25 - 300.0) * L ** 2 / (2 * 0.5e-3) * 1e3
    assert abs(d - exact) / exact < 1e-6
    assert isinstance(d, np.floating) and d.dtype == F32     # float32, as the reference computes it
    assert S.bow_deflection(F32(300.0), grid, geometry, Ly, pitch, MO['alpha']) == 0.0


def test_rf_potential_closed_form():
    f = 13.56
    assert S.rf_potential(100.0, f, 1.0 / (4 * f * 1e6)) == pytest.approx(100.0, rel=1e-12)
    assert abs(S.rf_potential(100.0, f, 1.0 / (2 * f * 1e6))) < 1e-9


def test_fixed_point_exact_above_8():
    v = np.array([8.0, 1234.5678, 3.3e7, 16777215.0], F32)
    q = S.fixed_point(v, S.HEAT_FRACTION_BITS)
    assert np.array_equal(q.astype(np.float64) / 2 ** S.HEAT_FRACTION_BITS, v.astype(np.float64))


@pytest.mark.parametrize("face", ["upstream", "downstream", "any"])
def test_groove_profile_equals_row_scan(face):
    r = np.random.default_rng(7)
    ny, nx, dy, thresh = 11, 17, 0.02, 1e4
    mask = np.zeros((ny, nx), bool)
    mask[2:9, 5:9] = True
    mask[4, 5] = False                          # a notch: the first node of row 4 moves right
    depth = r.random((ny, nx)) * 0.05
    damage = r.random((ny, nx)) * 3e3
    y, d = S.groove_profile(mask, depth, damage, dy, thresh, True, face)
    want = np.zeros(ny)
    for iy in range(ny):
        xs = [ix for ix in range(nx) if mask[iy, ix]]
        if not xs:
            continue
        tot = [depth[iy, ix] + damage[iy, ix] / thresh * dy for ix in range(nx)]
        if face == 'upstream':
            want[iy] = tot[xs[0]]
        elif face == 'downstream':
            want[iy] = tot[xs[-1]]
        else:
            cols = [ix for ix in range(nx) if mask[:, ix].any()]
            want[iy] = max(tot[ix] for ix in cols)
    assert np.allclose(d, want * 1000.0, rtol=1e-15, atol=0.0)
    assert np.array_equal(y, np.arange(ny) * dy)
    _, d0 = S.groove_profile(mask, depth, damage, dy, None, True, face)
    assert np.allclose(d0[mask.any(axis=1)], 1000.0 * np.array(
        [depth[iy, [ix for ix in range(nx) if mask[iy, ix]][0 if face == 'upstream' else -1]]
         if face != 'any' else depth[iy, 5:9].max() for iy in range(ny) if mask[iy].any()]), rtol=1e-15)
    assert S.groove_profile(np.zeros((3, 3), bool), depth[:3, :3], damage[:3, :3], dy)[0].size == 0
    with pytest.raises(ValueError):
        S.groove_profile(mask, depth, damage, dy, face='side')
