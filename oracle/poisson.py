"""Poisson operator, Picard loop and E field (test infrastructure, oracle side).

Restates ``build_sparse_matrix`` (Python/physics_engine.py:525-608),
``_recalc_poisson_cpu`` (:886-902) and the field part of ``recalc_poisson``
(:850-884) with SciPy SuperLU (``splu``) as the linear solver, exactly the
library the reference reaches through ``factorized``.

Pinned by: the manufactured-solution O(dx^2) convergence test of
Python/tests/poisson_solver_test.py (slope 2.00), the residual of every solve,
a hand-assembled operator on a 5x4 mesh, and ``np.gradient``.
"""
import numpy as np
import scipy.sparse as sp
from scipy.sparse.linalg import splu

# node kinds (the same codes are used by the HIP operator set-up)
K_BOUND = 0      # Dirichlet: V = V_fixed
K_UNKNOWN = 1    # interior 5-point row (also periodic bottom/top rows)
K_RIGHT = 2      # V[iy, nx-1] = V[iy, nx-2]
K_TOP = 3        # V[ny-1, ix] = V[ny-2, ix]
K_BOTTOM = 4     # V[0, ix]    = V[1, ix]


def node_kinds(isBound, periodic_y):
    ny, nx = isBound.shape
    kind = np.full((ny, nx), K_UNKNOWN, dtype=np.uint8)
    b = isBound.astype(bool)
    xs = np.arange(nx)[None, :]
    ys = np.arange(ny)[:, None]
    right = (xs == nx - 1) & ~b
    if not periodic_y:
        top = (ys == ny - 1) & ~b & ~right
        bottom = (ys == 0) & ~b & ~right & ~top
        kind[top] = K_TOP
        kind[bottom] = K_BOTTOM
    kind[right] = K_RIGHT
    kind[b] = K_BOUND
    return kind


def build_operator(isBound, periodic_y):
    """Assemble the reference's N x N matrix (CSC) and its row masks."""
    ny, nx = isBound.shape
    N = nx * ny
    kind = node_kinds(isBound, periodic_y).ravel()
    idx = np.arange(N)
    yy = idx // nx
    rows, cols, vals = [], [], []

    def add(r, c, v):
        rows.append(r)
        cols.append(c)
        vals.append(np.full(r.shape, v, dtype=np.float64))

    ib = idx[kind == K_BOUND]
    add(ib, ib, 1.0)
    ir = idx[kind == K_RIGHT]
    add(ir, ir, 1.0)
    add(ir, ir - 1, -1.0)
    it = idx[kind == K_TOP]
    add(it, it, 1.0)
    add(it, it - nx, -1.0)
    ibo = idx[kind == K_BOTTOM]
    add(ibo, ibo, 1.0)
    add(ibo, ibo + nx, -1.0)
    iu = idx[kind == K_UNKNOWN]
    if periodic_y:
        south = np.where(yy[iu] == 0, iu + (ny - 1) * nx, iu - nx)
        north = np.where(yy[iu] == ny - 1, iu - (ny - 1) * nx, iu + nx)
    else:
        south, north = iu - nx, iu + nx
    add(iu, iu, -4.0)
    add(iu, iu - 1, 1.0)
    add(iu, iu + 1, 1.0)
    add(iu, south, 1.0)
    add(iu, north, 1.0)
    A = sp.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(N, N)).tocsc()
    return A, kind.reshape(ny, nx)


class SuperLUPoisson:
    """The reference Poisson backend: SuperLU factorisation + Picard loop."""

    def __init__(self, isBound, periodic_y):
        self.A, self.kind = build_operator(isBound, periodic_y)
        self.lu = splu(self.A)
        self.bound = (self.kind == K_BOUND).ravel()
        self.rhs = (self.kind == K_UNKNOWN).ravel()

    def solve(self, b):
        return self.lu.solve(np.asarray(b, np.float64))

    def rhs_vector(self, rho_total_flat, V_fixed_flat, coeff):
        b = np.zeros(self.A.shape[0], dtype=np.float64)
        b[self.bound] = V_fixed_flat[self.bound]
        b[self.rhs] = -coeff * rho_total_flat[self.rhs]
        return b

    def picard(self, V, rho, V_fixed, coeff, V_plasma, Te_up, n0, q, iterations,
               omega=0.2, exp=np.exp):
        """`iterations` Picard/under-relaxation sweeps; returns the new V."""
        ny, nx = V.shape
        Vf = V_fixed.ravel()
        for _ in range(iterations):
            rho_e = -q * n0 * exp((np.minimum(V, V_plasma) - V_plasma) / Te_up)
            rho_total = rho + rho_e
            b = self.rhs_vector(rho_total.ravel(), Vf, coeff)
            V_new = self.solve(b).reshape(ny, nx)
            V = ((1 - omega) * V + omega * V_new).astype(np.float64)
        return V


def efield(V, dx, dy, periodic_y):
    """Ex, Ey (float32) = -grad V with np.gradient and the periodic Ey fix."""
    Ey, Ex = np.gradient(-V, dy * 1e-3, dx * 1e-3)
    if periodic_y:
        ny = V.shape[0]
        two_dy_m = 2.0 * dy * 1e-3
        Ey[0, :] = (V[ny - 1, :] - V[1, :]) / two_dy_m
        Ey[ny - 1, :] = (V[ny - 2, :] - V[0, :]) / two_dy_m
    return Ex.astype(np.float32), Ey.astype(np.float32)
