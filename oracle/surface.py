"""Electrode-surface physics of the reference time step, restated in NumPy
(test infrastructure, oracle side): impact heating, sputter damage and cell
failure, radiative cooling, heat conduction, grid temperatures, thermal bow
and the RF modulation of a grid potential.

dtype semantics follow the reference's NumPy path: the temperature map is
float32 and a Python float combined with float32 data is cast to float32
(NumPy 2 weak scalars); the damage map is float64 and receives float32
increments.

Functions and what pins them (tests/test_oracle_surface.py)
-----------------------------------------------------------
impact_heat        Python/physics_engine.py:1359-1367 (dT per hitting ion).
                   Pinned against the same value evaluated with explicit
                   np.float32 scalar arithmetic in the stated order, and by
                   the closed form E = m v^2 / 2 in float64 to float32 rounding.
sputter_damage     :1419-1433.  Pinned likewise; threshold E_th exactly.
radiative_cooling  :1566-1570.  Pinned by the closed form
                   T - c (T^4 - 300^4) (float64, float32 rounding) and T = 300
                   being a fixed point.
fourier_numbers    :1572-1585 (Fo limited to 0.2).  Closed form.
conduction         :1598-1613 (the NumPy loop: interior bound nodes only,
                   insulated towards non-electrode neighbours, floor 300 K).
                   Pinned bit-for-bit by ``conduction_loop``, a literal
                   per-node loop, on small maps; by heat conservation on a
                   closed electrode region away from the edges; and by
                   uniform maps being fixed points.
grid_means         :1615-1617 (np.mean over each mask, float32).
bow_deflection     :1619-1643.  Closed form of the cantilever formula
                   delta = alpha dT L^2 / (2 t) per geometry, float32 as the
                   reference evaluates it.
rf_potential       :984-991.  Closed form A sin(2 pi f t).
groove_profile     :1903-1932 (per-row depth of a grid face).  Pinned by a
                   literal per-row scan on small maps for every face.
"""
import numpy as np

F32 = np.float32

HEAT_FRACTION_BITS = 20    # fixed-point scale of the device heat accumulator
DAMAGE_FRACTION_BITS = 24  # fixed-point scale of the device damage accumulator


def impact_heat(vx, vy, vz, m_ion, macro_weight, C_cell, thermal_accel):
    """Temperature rise dT (float32) deposited by each hitting ion (:1361-1367)."""
    vx, vy, vz = (np.asarray(a, F32) for a in (vx, vy, vz))
    v_mag_sq = vx ** 2 + vy ** 2 + vz ** 2
    E_joules = 0.5 * m_ion * v_mag_sq * macro_weight
    return (E_joules / C_cell) * thermal_accel


def sputter_damage(vx, vy, vz, m_ion, q, E_th, Y_coeff, macro_weight):
    """(valid mask, float32 damage of the valid ions) (:1419-1433)."""
    vx, vy, vz = (np.asarray(a, F32) for a in (vx, vy, vz))
    E_eV = (0.5 * m_ion * (vx ** 2 + vy ** 2 + vz ** 2)) / q
    valid = E_eV > E_th
    damage = (Y_coeff * (E_eV[valid] - E_th)) * macro_weight
    return valid, damage


def cooling_factor(emissivity, sb_sigma, A_cell, dt, thermal_accel, C_cell):
    return (emissivity * sb_sigma * A_cell * dt * thermal_accel) / C_cell


def radiative_cooling(T, isBound, factor):
    """In place: T[bound] -= factor * (T[bound]^4 - 300^4), float32 (:1566-1570)."""
    T_bound = T[isBound]
    T[isBound] -= factor * (T_bound ** 4 - 300.0 ** 4)
    return T


def fourier_numbers(k, rho, cp, dt, thermal_accel, dx, dy, max_Fo=0.2):
    """(Fo_x, Fo_y) of the explicit conduction step, both scaled down so the
    larger is at most max_Fo (:1572-1585); dx, dy in mm."""
    alpha = k / (rho * cp)
    dt_th = dt * thermal_accel
    Fo_x = alpha * dt_th / (dx * 1e-3) ** 2
    Fo_y = alpha * dt_th / (dy * 1e-3) ** 2
    if Fo_x > max_Fo or Fo_y > max_Fo:
        s = max_Fo / max(Fo_x, Fo_y)
        Fo_x *= s
        Fo_y *= s
    return Fo_x, Fo_y


def conduction_loop(T, isBound, Fo_x, Fo_y, sweeps=10):
    """Literal per-node restatement of the reference loop (:1598-1613); small
    maps only.  Returns the new map (float32)."""
    T = np.array(T, F32)
    ny, nx = T.shape
    T_new = T.copy()
    for _ in range(sweeps):
        T_new[:] = T[:]
        for iy in range(1, ny - 1):
            for ix in range(1, nx - 1):
                if not isBound[iy, ix]:
                    continue
                c = T[iy, ix]
                l = T[iy, ix - 1] if isBound[iy, ix - 1] else c
                r = T[iy, ix + 1] if isBound[iy, ix + 1] else c
                d = T[iy - 1, ix] if isBound[iy - 1, ix] else c
                u = T[iy + 1, ix] if isBound[iy + 1, ix] else c
                dT = Fo_x * (l - 2.0 * c + r) + Fo_y * (d - 2.0 * c + u)
                T_new[iy, ix] = max(c + dT, 300.0)
        T, T_new = T_new, T
    return T


def conduction(T, isBound, Fo_x, Fo_y, sweeps=10):
    """Vectorised equivalent of conduction_loop (same float32 operations)."""
    T = np.array(T, F32)
    B = np.asarray(isBound, bool)
    inner = B[1:-1, 1:-1]
    fx, fy = F32(Fo_x), F32(Fo_y)
    for _ in range(sweeps):
        c = T[1:-1, 1:-1]
        l = np.where(B[1:-1, :-2], T[1:-1, :-2], c)
        r = np.where(B[1:-1, 2:], T[1:-1, 2:], c)
        d = np.where(B[:-2, 1:-1], T[:-2, 1:-1], c)
        u = np.where(B[2:, 1:-1], T[2:, 1:-1], c)
        dT = fx * ((l - F32(2.0) * c) + r) + fy * ((d - F32(2.0) * c) + u)
        s = c + dT
        s = np.where(F32(300.0) > s, F32(300.0), s)   # Python max(s, 300.0): NaN stays
        out = T.copy()
        out[1:-1, 1:-1] = np.where(inner, s, c)
        T = out
    return T


def clamp_bound(T, isBound):
    """T[bound] = max(T[bound], 300) (np.maximum, NaN propagates) (:1615)."""
    T[isBound] = np.maximum(T[isBound], 300.0)
    return T


def grid_means(T, masks, previous):
    """Mean temperature of each grid's nodes; grids with an empty mask keep
    their previous value (:1617-1619)."""
    out = list(previous)
    for i, m in enumerate(masks):
        if np.any(m):
            out[i] = np.mean(T[m])
    return out


def cantilever_length_mm(grid, geometry, Ly, pitch_mm):
    """Free length of the web that bows (:1624-1631)."""
    if geometry == 'two_holes':
        return max(0.2, pitch_mm - 2.0 * grid['r'])
    if geometry == 'one_hole':
        return max(0.2, 0.5 * (Ly - 2.0 * grid['r']))
    return max(0.2, Ly - grid['r'])


def bow_deflection(T_grid, grid, geometry, Ly, pitch_mm, alpha_thermal):
    """Thermal bow in mm (:1621-1638): delta = alpha dT L^2 / (2 t), with the
    reference's dtype (float32 when T_grid is a float32 mean)."""
    dT = T_grid - 300.0
    L_m = cantilever_length_mm(grid, geometry, Ly, pitch_mm) * 1e-3
    t_m = grid['t'] * 1e-3
    if t_m > 0 and L_m > 0:
        return alpha_thermal * dT * L_m ** 2 / (2.0 * t_m) * 1e3
    return 0.0


def rf_potential(amp, freq_mhz, t):
    """Sinusoidal RF component added to the RF grid (:986-989)."""
    return amp * np.sin(2.0 * np.pi * (freq_mhz * 1e6) * t)


def fixed_point(values, fraction_bits):
    """Round non-negative float32 values to the device accumulators' fixed
    point (value * 2^fraction_bits, round half to even); returns uint64."""
    s = np.asarray(values, np.float64) * float(2 ** fraction_bits)
    return np.rint(s).astype(np.uint64)


def groove_profile(mask, eroded_depth, damage_map, dy, thresh=None, accumulate_subcell=True,
                   face='upstream'):
    """(y_mm, depth_um) of one grid's face per mesh row (:1903-1932)."""
    mask = np.asarray(mask, bool)
    if not np.any(mask):
        return np.array([]), np.array([])
    ny, nx = mask.shape
    depth = np.asarray(eroded_depth, np.float64)
    if accumulate_subcell and thresh and thresh > 0:
        depth = depth + (np.asarray(damage_map, np.float64) / thresh) * dy
    per_y = np.zeros(ny, np.float64)
    rows = mask.any(axis=1)
    if face == 'any':
        cols = mask.any(axis=0)
        per_y[rows] = depth[rows][:, cols].max(axis=1)
    elif face == 'upstream':
        per_y[rows] = depth[np.nonzero(rows)[0], np.argmax(mask, axis=1)[rows]]
    elif face == 'downstream':
        last = nx - 1 - np.argmax(mask[:, ::-1], axis=1)
        per_y[rows] = depth[np.nonzero(rows)[0], last[rows]]
    else:
        raise ValueError(f"face must be 'upstream','downstream', or 'any' (got {face!r})")
    return np.arange(ny) * dy, per_y * 1000.0
