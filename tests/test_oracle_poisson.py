"""Pins of the oracle's Poisson operator (CPU)."""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle.poisson import build_operator, node_kinds, SuperLUPoisson, efield, K_BOUND, K_UNKNOWN, \
    K_RIGHT, K_TOP, K_BOTTOM


def hand_operator(isB, periodic):
    """Per-row loops following Python/physics_engine.py:525-608."""
    ny, nx = isB.shape
    N = nx * ny
    A = np.zeros((N, N))
    for iy in range(ny):
        for ix in range(nx):
            i = iy * nx + ix
            if isB[iy, ix]:
                A[i, i] = 1
            elif ix == nx - 1:
                A[i, i], A[i, i - 1] = 1, -1
            elif not periodic and iy == ny - 1:
                A[i, i], A[i, i - nx] = 1, -1
            elif not periodic and iy == 0:
                A[i, i], A[i, i + nx] = 1, -1
            else:
                A[i, i] = -4
                A[i, i - 1] += 1
                A[i, i + 1] += 1
                A[i, ((iy - 1) % ny) * nx + ix] += 1
                A[i, ((iy + 1) % ny) * nx + ix] += 1
    return A


@pytest.mark.parametrize("periodic", [False, True])
def test_operator_matches_hand_assembly(periodic):
    r = np.random.default_rng(0)
    isB = r.random((6, 7)) < 0.2
    isB[:, 0] = True
    A, kind = build_operator(isB, periodic)
    assert np.array_equal(A.toarray(), hand_operator(isB, periodic))


def test_node_kinds_corners():
    isB = np.zeros((5, 6), bool)
    isB[:, 0] = True
    k = node_kinds(isB, False)
    assert k[0, 5] == K_RIGHT and k[4, 5] == K_RIGHT  # right wins over top/bottom
    assert k[0, 3] == K_BOTTOM and k[4, 3] == K_TOP and k[2, 3] == K_UNKNOWN
    assert np.all(k[:, 0] == K_BOUND)


def test_manufactured_solution_second_order():
    """Python/tests/poisson_solver_test.py: V = sin(pi x/L) sin(pi y/L) on an
    all-Dirichlet box; RMS error must fall as dx^2 (published slope 2.00)."""
    L, eps0 = 20.0, 8.854e-12
    errs, dxs = [], [0.5, 0.25, 0.125]
    for dx in dxs:
        n = int(L / dx) + 1
        xs = np.linspace(0, L, n)
        X, Y = np.meshgrid(xs, xs)
        k = np.pi / L
        Vx = np.sin(k * X) * np.sin(k * Y)
        rho = eps0 * 2 * (k * 1000) ** 2 * Vx
        isB = np.zeros((n, n), bool)
        isB[0, :] = isB[-1, :] = isB[:, 0] = isB[:, -1] = True
        P = SuperLUPoisson(isB, False)
        b = P.rhs_vector(rho.ravel(), np.zeros(n * n), (dx * 1e-3) ** 2 / eps0)
        V = P.solve(b).reshape(n, n)
        errs.append(np.sqrt(np.mean((V - Vx) ** 2)))
    slope = np.polyfit(np.log(dxs), np.log(errs), 1)[0]
    assert abs(slope - 2.0) < 0.05
    assert errs[1] < 1e-4  # published figure: ~6.5e-5 at dx = 0.25 mm


def test_reduced_system_is_spd_and_equivalent():
    """Eliminating Dirichlet and Neumann rows gives an SPD system whose
    solution equals the full system's (the form the MG-PCG solves)."""
    r = np.random.default_rng(3)
    ny, nx = 9, 12
    isB = r.random((ny, nx)) < 0.15
    isB[:, 0] = True
    A, kind = build_operator(isB, False)
    b = r.normal(0, 1, nx * ny)
    b[(kind == K_RIGHT).ravel() | (kind == K_TOP).ravel() | (kind == K_BOTTOM).ravel()] = 0.0
    V = sp.linalg.spsolve(A.tocsc(), b)
    u = np.where(kind.ravel() == K_UNKNOWN)[0]
    pos = {j: k for k, j in enumerate(u)}
    R = np.zeros((len(u), len(u)))
    br = np.zeros(len(u))
    kf = kind.ravel()
    for k, i in enumerate(u):
        iy, ix = divmod(i, nx)
        d = 4
        br[k] = -b[i]
        for j in (i - 1, i + 1, i - nx, i + nx):
            if kf[j] == K_UNKNOWN:
                R[k, pos[j]] -= 1
            elif kf[j] == K_BOUND:
                br[k] += b[j]
            else:
                d -= 1
        R[k, k] = d
    assert np.allclose(R, R.T)
    assert np.min(np.linalg.eigvalsh(R)) > 0
    assert np.allclose(np.linalg.solve(R, br), V[u], rtol=1e-10, atol=1e-12)


def test_efield_matches_np_gradient():
    r = np.random.default_rng(1)
    V = r.normal(0, 100, (8, 11))
    Ex, Ey = efield(V, 0.05, 0.04, False)
    gy, gx = np.gradient(-V, 0.04e-3, 0.05e-3)
    assert np.allclose(Ex, gx.astype(np.float32)) and np.allclose(Ey, gy.astype(np.float32))
    Ex, Ey = efield(V, 0.05, 0.04, True)
    assert np.isclose(Ey[0, 3], np.float32((V[7, 3] - V[1, 3]) / (2 * 0.04e-3)))
