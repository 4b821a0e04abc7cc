"""Pins of the oracle's particle kernels (CPU): closed forms, brute force,
library routines."""
import math

import numpy as np
import pytest

from oracle import kernels as K

Q = 1.602e-19
M_XE = 131.293 * 1.6605e-27


def _uniform_fields(nx, ny, ex=0.0, ey=0.0, bx=0.0, by=0.0, bz=0.0):
    f = lambda v: np.full((ny, nx), v, np.float32)
    return f(ex), f(ey), f(bx), f(by), f(bz)


def test_boris_gyration_closed_form():
    """Discrete Boris orbit in uniform Bz (Python/tests/particle_push_test.py:6-27
    parameters): speed is conserved exactly, the velocity turns by
    2*atan(w dt/2) per step and positions lie on a circle of radius
    r*sqrt(1+(w dt/2)^2), r = m v0/(qB)."""
    Bz, dt, v0, steps = 0.5, 1e-9, 1e4, 2000
    qm = Q / M_XE
    w = qm * Bz
    nx = ny = 2
    Ex, Ey, Bx, By, Bzf = _uniform_fields(nx, ny, bz=Bz)
    x = np.array([0.0]); y = np.array([0.0])
    vx = np.array([v0]); vy = np.array([0.0]); vz = np.array([0.0])
    xs, ys = [], []
    for _ in range(steps):
        K.push_boris(x, y, vx, vy, vz, Ex, Ey, Bx, By, Bzf, 1, 1e6, 1e6, nx, ny, dt, qm, mode="fp64")
        xs.append(x[0]); ys.append(y[0])
    xs, ys = np.array(xs), np.array(ys)
    speed = math.hypot(vx[0], vy[0])
    assert abs(speed - v0) / v0 < 1e-12
    theta = 2 * math.atan(w * dt / 2)
    ang = math.atan2(vy[0], vx[0])
    want = -(steps * theta)
    assert abs(math.remainder(ang - want, 2 * math.pi)) < 1e-9
    R = (v0 / w) * math.sqrt(1 + (w * dt / 2) ** 2) * 1e3  # mm
    # centre from the first three points; all points on the circle
    A = np.array([[2 * (xs[1] - xs[0]), 2 * (ys[1] - ys[0])], [2 * (xs[2] - xs[0]), 2 * (ys[2] - ys[0])]])
    rhs = np.array([xs[1] ** 2 - xs[0] ** 2 + ys[1] ** 2 - ys[0] ** 2,
                    xs[2] ** 2 - xs[0] ** 2 + ys[2] ** 2 - ys[0] ** 2])
    cx, cy = np.linalg.solve(A, rhs)
    r = np.hypot(xs - cx, ys - cy)
    assert np.max(np.abs(r - R)) / R < 1e-9


def test_boris_constant_E_closed_form():
    """Uniform E, B = 0: v_k = v0 + k a dt, x_k = x0 + 1000 dt (k v0 + a dt k(k+1)/2)."""
    E0, dt, v0, steps = 2.0e4, 5e-10, 1500.0, 300
    qm = Q / M_XE
    a = qm * E0
    Ex, Ey, Bx, By, Bz = _uniform_fields(3, 3, ex=E0)
    x = np.array([0.5]); y = np.array([0.5])
    vx = np.array([v0]); vy = np.array([0.0]); vz = np.array([0.0])
    for _ in range(steps):
        K.push_boris(x, y, vx, vy, vz, Ex, Ey, Bx, By, Bz, 1, 10.0, 10.0, 3, 3, dt, qm, mode="fp64")
    k = steps
    assert abs(vx[0] - (v0 + k * a * dt)) / (v0 + k * a * dt) < 1e-12
    xw = 0.5 + 1000 * dt * (k * v0 + a * dt * k * (k + 1) / 2)
    assert abs(x[0] - xw) / xw < 1e-12


def test_push_ref32_matches_loop():
    """Vectorised ref32 push equals a per-particle restatement with the same
    promotions (f32 state, f64 weights and kicks, f32 drift)."""
    rng = np.random.default_rng(0)
    nx, ny = 17, 9
    dx, dy = 0.05, 0.04
    Ex = rng.normal(0, 1e5, (ny, nx)).astype(np.float32)
    Ey = rng.normal(0, 1e5, (ny, nx)).astype(np.float32)
    Z = np.zeros((ny, nx), np.float32)
    n = 300
    x = rng.uniform(-0.1, (nx - 1) * dx + 0.1, n).astype(np.float32)
    y = rng.uniform(-0.1, (ny - 1) * dy + 0.1, n).astype(np.float32)
    vx, vy, vz = [rng.normal(0, 3e4, n).astype(np.float32) for _ in range(3)]
    X, Y, VX, VY, VZ = [a.copy() for a in (x, y, vx, vy, vz)]
    dt, qm = 5e-10, -1.47e8
    K.push_boris(X, Y, VX, VY, VZ, Ex, Ey, Z, Z, Z, n, dx, dy, nx, ny, dt, qm)
    for i in range(n):
        gx = np.float32(x[i] / np.float32(dx))
        gy = np.float32(y[i] / np.float32(dy))
        i0 = min(max(int(math.floor(gx)), 0), nx - 2)
        j0 = min(max(int(math.floor(gy)), 0), ny - 2)
        fx, fy = float(gx) - i0, float(gy) - j0
        w = [(1 - fx) * (1 - fy), fx * (1 - fy), (1 - fx) * fy, fx * fy]
        ex = float(Ex[j0, i0]) * w[0] + float(Ex[j0, i0 + 1]) * w[1] + float(Ex[j0 + 1, i0]) * w[2] + float(Ex[j0 + 1, i0 + 1]) * w[3]
        ey = float(Ey[j0, i0]) * w[0] + float(Ey[j0, i0 + 1]) * w[1] + float(Ey[j0 + 1, i0]) * w[2] + float(Ey[j0 + 1, i0 + 1]) * w[3]
        h = 0.5 * qm * dt
        nvx = np.float32((float(vx[i]) + h * ex) + h * ex)
        nvy = np.float32((float(vy[i]) + h * ey) + h * ey)
        nx_ = np.float32(x[i] + np.float32(np.float32(nvx * np.float32(dt)) * np.float32(1000.0)))
        assert VX[i] == nvx and VY[i] == nvy and X[i] == nx_ and VZ[i] == vz[i]


def test_deposit_counts_brute_force():
    rng = np.random.default_rng(1)
    nx, ny, dx, dy = 20, 11, 0.1, 0.07
    x = rng.uniform(-0.3, 2.3, 5000).astype(np.float32)
    y = rng.uniform(-0.3, 1.0, 5000).astype(np.float32)
    x[:5] = [0.05, 0.15, 0.25, np.nan, 1e30]  # ties at half-cells, NaN, huge
    got = K.deposit_counts(x, y, dx, dy, nx, ny).reshape(ny, nx)
    want = np.zeros((ny, nx), np.int64)
    for xi, yi in zip(x, y):
        gx, gy = np.float32(xi / np.float32(dx)), np.float32(yi / np.float32(dy))
        def idx(g, n):
            r = np.round(g)
            i = int(r) if np.isfinite(r) and abs(r) < 2 ** 31 else -2 ** 31
            return min(max(i, 1), n - 2)
        want[idx(gy, ny), idx(gx, nx)] += 1
    assert np.array_equal(got, want)


def test_rho_from_counts_equals_add_at():
    rng = np.random.default_rng(2)
    nx, ny, dx, dy = 30, 12, 0.1, 0.1
    xi = rng.uniform(0, 3, 20000).astype(np.float32); yi = rng.uniform(0, 1.1, 20000).astype(np.float32)
    xe = rng.uniform(0, 3, 7000).astype(np.float32); ye = rng.uniform(0, 1.1, 7000).astype(np.float32)
    cd = 1.602e-19 * 1000.0 / (1e-4 * 1e-4 * 1e-3)
    rho = np.zeros((ny, nx), np.float32)
    K.accumulate_rho(xi, yi, rho, len(xi), dx, dy, nx, ny, cd)
    K.accumulate_rho(xe, ye, rho, len(xe), dx, dy, nx, ny, -cd)
    ci = K.deposit_counts(xi, yi, dx, dy, nx, ny)
    ce = K.deposit_counts(xe, ye, dx, dy, nx, ny)
    got = K.rho_from_counts(ci, cd, ce, -cd).reshape(ny, nx)
    assert np.array_equal(got.view(np.uint32), rho.view(np.uint32))
    # total charge invariant
    tot = float(np.sum(got, dtype=np.float64))
    assert abs(tot - cd * (len(xi) - len(xe))) < 1e-5 * cd * len(xi)


def test_substeps_brute_force():
    rng = np.random.default_rng(3)
    nx, ny, dx, dy = 25, 14, 0.05, 0.03
    Ex = rng.normal(0, 3e5, (ny, nx)).astype(np.float32)
    Ey = rng.normal(0, 3e5, (ny, nx)).astype(np.float32)
    n = 400
    x = rng.uniform(0, 1.2, n).astype(np.float32); y = rng.uniform(0, 0.39, n).astype(np.float32)
    vx, vy, vz = [rng.normal(0, 5e4, n).astype(np.float32) for _ in range(3)]
    qm, dt = -1.47e8, 5e-10
    n_sub, ds_max, ds_lim = K.substeps(x, y, vx, vy, vz, Ex, Ey, dx, dy, nx, ny, qm, dt)
    best = 0.0
    for i in range(n):
        gx, gy = np.float32(x[i] / np.float32(dx)), np.float32(y[i] / np.float32(dy))
        i0 = min(max(int(np.floor(gx)), 0), nx - 2); j0 = min(max(int(np.floor(gy)), 0), ny - 2)
        fx, fy = float(gx) - i0, float(gy) - j0
        def g(F):
            return (float(F[j0, i0]) * (1 - fx) * (1 - fy) + float(F[j0, i0 + 1]) * fx * (1 - fy)
                    + float(F[j0 + 1, i0]) * (1 - fx) * fy + float(F[j0 + 1, i0 + 1]) * fx * fy)
        a = math.sqrt((qm * g(Ex)) ** 2 + (qm * g(Ey)) ** 2)
        v = np.sqrt(np.float32(np.float32(vx[i] * vx[i] + vy[i] * vy[i]) + vz[i] * vz[i]))
        best = max(best, float(np.float32(v * np.float32(dt))) + 0.5 * a * dt ** 2)
    assert ds_max == best
    assert n_sub == max(1, math.ceil(best / (0.25 * min(dx, dy) * 1e-3)))


def test_segment_hits_brute_force():
    rng = np.random.default_rng(4)
    nx, ny, dx, dy = 30, 20, 0.1, 0.1
    isB = rng.random((ny, nx)) < 0.1
    n = 2000
    x0 = rng.uniform(0, 2.9, n).astype(np.float32); y0 = rng.uniform(0, 1.9, n).astype(np.float32)
    x1 = (x0 + rng.normal(0, 0.3, n)).astype(np.float32); y1 = (y0 + rng.normal(0, 0.3, n)).astype(np.float32)
    hit, hix, hiy = K.segment_hits(isB, x0, y0, x1, y1, dx, dy, nx, ny)
    for i in range(n):
        h, fx_, fy_ = False, None, None
        for k in range(1, 9):
            f = np.float32(k / 8.0)
            xi = np.float32(x0[i] + np.float32(np.float32(x1[i] - x0[i]) * f))
            yi = np.float32(y0[i] + np.float32(np.float32(y1[i] - y0[i]) * f))
            cx = min(max(int(np.round(np.float32(xi / np.float32(dx)))), 0), nx - 1)
            cy = min(max(int(np.round(np.float32(yi / np.float32(dy)))), 0), ny - 1)
            if isB[cy, cx] and not h:
                h, fx_, fy_ = True, cx, cy
        assert hit[i] == h
        if h:
            assert (hix[i], hiy[i]) == (fx_, fy_)


def test_ppc_map_trunc():
    x = np.array([0.0, 0.099, 0.1, 0.25, 2.99], np.float32)
    y = np.array([0.0, 0.05, 0.15, 0.2999, 0.0], np.float32)
    m = K.ppc_map(x, y, 0.1, 0.1, 30, 4)
    assert m.sum() == 5 and m[0, 0] == 2 and m[1, 1] == 1 and m[2, 2] == 1 and m[0, 29] == 1


def test_mcc_det64_matches_ref32_formula():
    """The float64 deterministic CEX probability restates the reference's
    float32 formula (Roy plume neutrals, built-in cross-section)."""
    rng = np.random.default_rng(5)
    n = 20000
    x = rng.uniform(1.0, 8.0, n).astype(np.float32); y = rng.uniform(0, 1.5, n).astype(np.float32)
    vx = rng.normal(3e4, 5e3, n).astype(np.float32); vy, vz = [rng.normal(0, 1e3, n).astype(np.float32) for _ in range(2)]
    p64 = K.mcc_prob_det64(x, y, vx, vy, vz, 1.5, 1e20, 5e-10, M_XE, Q)
    p32 = K.mcc_prob_ref32(x, y, vx, vy, vz, 1.5, 1e20, 5e-10, M_XE, Q)
    assert np.max(np.abs(p64 - p32) / p64) < 5e-5
    # closed form at the axis origin: theta = 0, R = r_T -> n = n0*a*(1-1/sqrt2) = n0
    p = K.mcc_prob_det64(np.array([1.0], np.float32), np.array([0.0], np.float32),
                         np.array([1e4], np.float32), np.zeros(1, np.float32), np.zeros(1, np.float32),
                         1.5, 1e20, 5e-10, M_XE, Q)
    sig = (-0.8821 * math.log(1e4) + 15.1262) ** 2 * 1e-20
    assert abs(p[0] - (1 - math.exp(-1e20 * sig * 1e4 * 5e-10))) < 1e-15


def test_mcc_loglinear_uniform():
    p = K.mcc_prob_det64(np.array([2.0], np.float32), np.array([0.3], np.float32),
                         np.array([2e4], np.float32), np.zeros(1, np.float32), np.zeros(1, np.float32),
                         1.5, 1e18, 5e-10, M_XE, Q, neutral_model="uniform", cx_model="loglinear")
    E = 0.5 * M_XE * 2e4 ** 2 / Q
    sig = (87.3 - 13.6 * math.log10(E)) * 1e-20
    assert abs(p[0] - (1 - math.exp(-1e18 * sig * 2e4 * 5e-10))) < 1e-15
