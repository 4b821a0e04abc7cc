"""Pins of the oracle's bilinear (cloud-in-cell) deposit
(Python/docs/PY-BEMCS_Manual.tex:215-230): a literal per-particle loop in
Python floats, charge and first-moment conservation, particles on nodes, and
the float64 formula."""
import math

import numpy as np
import pytest

from oracle import kernels as K

F32 = np.float32


def _particles(n, nx, ny, dx, dy, seed):
    r = np.random.default_rng(seed)
    x = (r.random(n) * (nx - 1) * dx).astype(F32)
    y = (r.random(n) * (ny - 1) * dy).astype(F32)
    # edges: both ends of each axis, a node, just past the last node
    x[:6] = np.array([0.0, (nx - 1) * dx, 3 * dx, 2.5 * dx, (nx - 1) * dx * 1.0000001, 0.0], F32)
    y[:6] = np.array([0.0, (ny - 1) * dy, 5 * dy, 0.5 * dy, 0.0, (ny - 1) * dy * 1.0000001], F32)
    return x, y


def _literal(x, y, dx, dy, nx, ny):
    """Per-particle loop in Python floats: g = x / f32(dx) in float32,
    i = clip(floor(g), 0, n-2), h = g - i, weights rounded half to even."""
    W = {}
    for px, py in zip(x, y):
        gx = float(F32(px) / F32(dx))
        gy = float(F32(py) / F32(dy))
        if not (-1.0 <= gx <= nx and -1.0 <= gy <= ny):
            continue
        ix = min(max(math.floor(gx), 0), nx - 2)
        iy = min(max(math.floor(gy), 0), ny - 2)
        hx, hy = gx - ix, gy - iy
        for (cx, cy), wt in (((0, 0), (1 - hx) * (1 - hy)), ((1, 0), hx * (1 - hy)),
                             ((0, 1), (1 - hx) * hy), ((1, 1), hx * hy)):
            k = (iy + cy) * nx + ix + cx
            W[k] = W.get(k, 0) + round(wt * 2.0 ** 32)
    out = np.zeros(nx * ny, np.int64)
    for k, v in W.items():
        out[k] = v
    return out


@pytest.mark.parametrize("nx,ny,seed", [(13, 9, 0), (40, 17, 1), (7, 5, 2)])
def test_deposit_equals_literal_loop(nx, ny, seed):
    dx, dy = 0.02, 0.025
    x, y = _particles(700, nx, ny, dx, dy, seed)
    x[10], y[11] = np.nan, np.inf                       # non-finite: no deposit
    x[12] = -0.5 * dx * 3                               # g = -1.5: outside [-1, n], no deposit
    x[13] = -0.5 * dx                                   # g = -0.5: extrapolated, kept
    assert np.array_equal(K.deposit_cic(x, y, dx, dy, nx, ny), _literal(x, y, dx, dy, nx, ny))


def test_charge_and_first_moment_conservation():
    nx, ny, dx, dy = 64, 24, 0.02, 0.02
    x, y = _particles(20000, nx, ny, dx, dy, 3)
    W = K.deposit_cic(x, y, dx, dy, nx, ny)
    n = len(x)
    # each particle's four rounded weights sum to 2^32 within 2 units
    assert abs(int(W.sum()) - n * 2 ** 32) <= 2 * n
    Wn = W.reshape(ny, nx).astype(np.float64) / 2.0 ** 32
    gx = (x / F32(dx)).astype(np.float64)
    gy = (y / F32(dy)).astype(np.float64)
    # bilinear weights reproduce the position exactly: sum_nodes W i = sum_p g
    mx = (Wn * np.arange(nx)[None, :]).sum()
    my = (Wn * np.arange(ny)[:, None]).sum()
    assert abs(mx - gx.sum()) <= 1e-8 * gx.sum()
    assert abs(my - gy.sum()) <= 1e-8 * gy.sum()


def test_particle_on_node_and_rho():
    nx, ny, dx, dy = 10, 6, 0.03125, 0.03125     # spacing 2^-5: node positions exact
    x = np.array([3 * dx, 9 * dx], F32)      # interior node; last column (clamped cell, h = 1)
    y = np.array([2 * dy, 5 * dy], F32)
    W = K.deposit_cic(x, y, dx, dy, nx, ny).reshape(ny, nx)
    assert W[2, 3] == 2 ** 32 and W[5, 9] == 2 ** 32 and int(W.sum()) == 2 * 2 ** 32
    # rho against the manual's float64 formula
    r = np.random.default_rng(4)
    xs = (r.random(5000) * (nx - 1) * dx).astype(F32)
    ys = (r.random(5000) * (ny - 1) * dy).astype(F32)
    cd = 1.602e-19 * 1e3 / (dx * dy * 1e-9)
    rho = K.rho_from_weights(K.deposit_cic(xs, ys, dx, dy, nx, ny), cd,
                             K.deposit_cic(ys * 0 + xs[::-1], ys, dx, dy, nx, ny), -cd)
    ref = np.zeros(nx * ny)
    for sx, sign in ((xs, 1.0), (xs[::-1], -1.0)):
        gx = (sx / F32(dx)).astype(np.float64)
        gy = (ys / F32(dy)).astype(np.float64)
        ix = np.clip(np.floor(gx), 0, nx - 2).astype(int)
        iy = np.clip(np.floor(gy), 0, ny - 2).astype(int)
        hx, hy = gx - ix, gy - iy
        for ox, oy, w in ((0, 0, (1 - hx) * (1 - hy)), (1, 0, hx * (1 - hy)),
                          (0, 1, (1 - hx) * hy), (1, 1, hx * hy)):
            np.add.at(ref, (iy + oy) * nx + ix + ox, sign * cd * w)
    scale = np.abs(ref).max()
    assert np.max(np.abs(rho.astype(np.float64) - ref)) <= 1e-6 * scale


def test_rho_from_weights_order():
    W1 = np.array([2 ** 32, 3 * 2 ** 31, 0, -5], np.int64)
    W2 = np.array([2 ** 30, 0, 7, 2 ** 32], np.int64)
    got = K.rho_from_weights(W1, 1.5, W2, -0.25)
    want = np.zeros(4, F32)
    for W, cd in ((W1, 1.5), (W2, -0.25)):
        want = (want.astype(np.float64) + cd * (W.astype(np.float64) / 2.0 ** 32)).astype(F32)
    assert np.array_equal(got, want)
