"""Per-particle kernels of the reference time step, restated in NumPy
(test infrastructure, oracle side).

dtype semantics follow the reference's NumPy path exactly ("ref32"): particle
state is float32, a Python float combined with a float32 array is cast to
float32 (NumPy 2 weak scalars, also in comparisons), and float32 - int32
promotes to float64.  ``mode="fp64"`` evaluates the same formulas with float64
state, which is the FP64 yardstick for stating float32 error.

Functions and what pins them
----------------------------
push_boris          Python/physics_engine.py:188-254.  Pinned by the discrete
                    Boris-orbit closed form (circle radius r*sqrt(1+(w dt/2)^2),
                    rotation 2*atan(w dt/2) per step, exact speed
                    conservation), the constant-E closed form and a per-particle
                    loop on tiny inputs (tests/test_oracle_kernels.py).
nearest_index       the ``np.round(.../d).astype(int)`` + clip idiom used at
                    :183-184, :1326-1327, :1541-1542, :1883-1891.
deposit_counts      :178-185 (NGP).  Pinned by brute-force loops and by
rho_from_counts     bit equality with ``np.add.at`` (:185).
deposit_cic         the bilinear deposit of the model manual
rho_from_weights    (Python/docs/PY-BEMCS_Manual.tex:215-230) with the field
                    gather's index and fractions, weights in 2^-32 fixed point.
                    Pinned by a literal per-particle loop in Python floats, by
                    charge conservation and exact first-moment conservation
                    (sum_nodes W x_node = sum_p x_p), by particles on nodes, and
                    by the float64 formula to 1e-6 of max |rho|.
substeps            :928-966.  Pinned by a per-particle loop.
segment_hits        :1880-1901.  Pinned by a per-particle loop.
ppc_map             :1749-1758.  Pinned by ``np.bincount`` on brute-force indices.
mcc_prob_ref32      :1690-1719 evaluated with the reference's float32 math.
mcc_prob_det64      the same quantities in float64 with oracle.dmath; agrees
                    with mcc_prob_ref32 to float32 rounding (tested); user
                    cross-section tables through oracle.spline.
"""
import numpy as np

from .dmath import det_exp, det_log, INV_LN10

F32 = np.float32


def _floor_i32(v):
    """np.floor(v).astype(np.int32) with x86 out-of-range/NaN -> INT32_MIN."""
    f = np.floor(v)
    ok = (f >= -2147483648.0) & (f < 2147483648.0)
    out = np.full(f.shape, np.int32(-2147483648), dtype=np.int32)
    out[ok] = f[ok].astype(np.int32)
    return out


def _round_i64(v):
    """np.round(v).astype(int) with NaN/out-of-range -> INT64_MIN."""
    f = np.round(v)
    ok = (f >= -9.2e18) & (f < 9.2e18)
    out = np.full(f.shape, np.iinfo(np.int64).min, dtype=np.int64)
    out[ok] = f[ok].astype(np.int64)
    return out


def nearest_index(v, d, n, lo=0):
    """clip(round(v / d), lo, n-1) for float32 v and Python-float d."""
    q = np.asarray(v) / d            # float32 array / weak float -> float32
    return np.clip(_round_i64(q), lo, n - 1 - lo)


def push_boris(x, y, vx, vy, vz, Ex, Ey, Bx, By, Bz, num_p, dx, dy, nx, ny, dt,
               q_m, mode="ref32"):
    """One gather + Boris + drift pass, in place on x..vz[:num_p]."""
    n = int(num_p)
    if n == 0:
        return
    if mode == "fp64":
        dxs, dys = float(dx), float(dy)
    else:
        dxs, dys = dx, dy
    gx = x[:n] / dxs
    gy = y[:n] / dys
    i0 = np.clip(_floor_i32(gx), 0, nx - 2)
    j0 = np.clip(_floor_i32(gy), 0, ny - 2)
    fx = gx - i0                      # float32 - int32 -> float64
    fy = gy - j0
    i1 = np.minimum(i0 + 1, nx - 1)
    j1 = np.minimum(j0 + 1, ny - 1)
    a00 = (1.0 - fx) * (1.0 - fy)
    a10 = fx * (1.0 - fy)
    a01 = (1.0 - fx) * fy
    a11 = fx * fy

    def gather(F):
        return F[j0, i0] * a00 + F[j0, i1] * a10 + F[j1, i0] * a01 + F[j1, i1] * a11

    ex, ey = gather(Ex), gather(Ey)
    bx, by, bz = gather(Bx), gather(By), gather(Bz)
    h = 0.5 * q_m * dt
    ux = vx[:n] + h * ex
    uy = vy[:n] + h * ey
    uz = vz[:n]
    tx, ty, tz = h * bx, h * by, h * bz
    tt = tx * tx + ty * ty + tz * tz
    sx = 2.0 * tx / (1.0 + tt)
    sy = 2.0 * ty / (1.0 + tt)
    sz = 2.0 * tz / (1.0 + tt)
    wx = ux + (uy * tz - uz * ty)
    wy = uy + (uz * tx - ux * tz)
    wz = uz + (ux * ty - uy * tx)
    px = ux + (wy * sz - wz * sy)
    py = uy + (wz * sx - wx * sz)
    pz = uz + (wx * sy - wy * sx)
    vx[:n] = px + h * ex
    vy[:n] = py + h * ey
    vz[:n] = pz
    x[:n] += vx[:n] * dt * 1000.0
    y[:n] += vy[:n] * dt * 1000.0


def deposit_counts(x, y, dx, dy, nx, ny):
    """NGP cell counts (uint32[ny*nx]) with the deposit's clip to [1, n-2]."""
    gx = np.asarray(x) / dx
    gy = np.asarray(y) / dy
    ix = np.clip(_round_i32(gx), 1, nx - 2)
    iy = np.clip(_round_i32(gy), 1, ny - 2)
    return np.bincount(iy.astype(np.int64) * nx + ix, minlength=nx * ny).astype(np.uint32)


def _round_i32(v):
    f = np.round(v)
    ok = (f >= -2147483648.0) & (f < 2147483648.0)
    out = np.full(f.shape, np.int32(-2147483648), dtype=np.int32)
    out[ok] = f[ok].astype(np.int32)
    return out


def accumulate_rho(x, y, rho, num_p, dx, dy, nx, ny, charge_density):
    """The reference deposit itself (np.add.at into float32 rho)."""
    n = int(num_p)
    if n == 0:
        return
    ix = np.clip(_round_i32(x[:n] / dx), 1, nx - 2)
    iy = np.clip(_round_i32(y[:n] / dy), 1, ny - 2)
    np.add.at(rho, (iy, ix), charge_density)


def rho_from_counts(counts_first, cd_first, counts_second=None, cd_second=None,
                    rho0=None):
    """float32 rho equal bit-for-bit to sequential np.add.at deposits: all
    particles of the first species, then all of the second.  np.add.at with a
    Python-float value evaluates rho = f32(f64(rho) + cd) for every particle
    (tested against np.add.at), and within one deposit every added value is the
    same, so the result depends only on the per-cell count."""
    c1 = np.asarray(counts_first, np.int64)
    rho = np.zeros(c1.shape, F32) if rho0 is None else np.array(rho0, F32, copy=True)
    steps = [(c1, float(cd_first))]
    if counts_second is not None:
        steps.append((np.asarray(counts_second, np.int64), float(cd_second)))
    for cnt, v in steps:
        todo = cnt.copy()
        while True:
            m = todo > 0
            if not m.any():
                break
            rho[m] = (rho[m].astype(np.float64) + v).astype(F32)
            todo[m] -= 1
    return rho


CIC_SCALE = 2.0 ** 32   # fixed-point units per unit weight of the bilinear deposit


def cic_weights(x, y, dx, dy, nx, ny):
    """(base node int64[n] or -1, fixed-point corner weights int64[n, 4]) of
    the bilinear deposit: g = x/f32(dx), i = clip(floor(g), 0, nx-2),
    h = f64(g) - i; corners [i,j], [i+1,j], [i,j+1], [i+1,j+1]; particles
    with g outside [-1, n] (or NaN) deposit nothing."""
    gx = np.asarray(x, F32) / F32(dx)
    gy = np.asarray(y, F32) / F32(dy)
    ok = (gx >= -1.0) & (gx <= nx) & (gy >= -1.0) & (gy <= ny)
    ix = np.clip(_floor_i32(gx), 0, nx - 2).astype(np.int64)
    iy = np.clip(_floor_i32(gy), 0, ny - 2).astype(np.int64)
    hx = gx.astype(np.float64) - ix
    hy = gy.astype(np.float64) - iy
    ax, ay = 1.0 - hx, 1.0 - hy
    w = np.stack([np.rint((ax * ay) * CIC_SCALE), np.rint((hx * ay) * CIC_SCALE),
                  np.rint((ax * hy) * CIC_SCALE), np.rint((hx * hy) * CIC_SCALE)], axis=1)
    w = np.where(ok[:, None], w, 0.0).astype(np.int64)
    return np.where(ok, iy * nx + ix, -1), w


def deposit_cic(x, y, dx, dy, nx, ny):
    """Per-node sums (int64[ny*nx]) of the fixed-point bilinear weights."""
    base, w = cic_weights(x, y, dx, dy, nx, ny)
    ok = base >= 0
    b, w = base[ok], w[ok]
    W = np.zeros(nx * ny, np.int64)
    for c, off in enumerate((0, 1, nx, nx + 1)):
        np.add.at(W, b + off, w[:, c])
    return W


def rho_from_weights(w_first, cd_first, w_second=None, cd_second=None, rho0=None):
    """rho = f32(f64(rho) + cd * W 2^-32), first species then second."""
    rho = (np.zeros(np.shape(w_first), F32) if rho0 is None else np.array(rho0, F32, copy=True))
    for W, cd in ((w_first, cd_first), (w_second, cd_second)):
        if W is None:
            continue
        s = np.asarray(W, np.int64).astype(np.float64) * (1.0 / CIC_SCALE)
        rho = (rho.astype(np.float64) + float(cd) * s).astype(F32)
    return rho


def substeps(x, y, vx, vy, vz, Ex, Ey, dx, dy, nx, ny, qm, dt, frac=0.25):
    """(n_sub, ds_max, ds_lim) of compute_particle_substeps."""
    if len(x) == 0:
        return 1, 0.0, frac * min(dx, dy) * 1e-3
    gx = x / dx
    gy = y / dy
    i0 = np.clip(_floor_i64(gx), 0, nx - 2)
    j0 = np.clip(_floor_i64(gy), 0, ny - 2)
    fx = gx - i0
    fy = gy - j0
    i1 = i0 + 1
    j1 = j0 + 1

    def gather(F):
        return (F[j0, i0] * (1 - fx) * (1 - fy) + F[j0, i1] * fx * (1 - fy)
                + F[j1, i0] * (1 - fx) * fy + F[j1, i1] * fx * fy)

    ex, ey = gather(Ex), gather(Ey)
    a_mag = np.sqrt((qm * ex) ** 2 + (qm * ey) ** 2)
    v_mag = np.sqrt(vx * vx + vy * vy + vz * vz)
    ds = v_mag * dt + 0.5 * a_mag * dt ** 2
    ds_max = float(np.max(ds))
    ds_lim = frac * min(dx, dy) * 1e-3
    n_sub = max(1, int(np.ceil(ds_max / max(ds_lim, 1e-30))))
    return n_sub, ds_max, ds_lim


def _floor_i64(v):
    f = np.floor(v)
    ok = (f >= -9.2e18) & (f < 9.2e18)
    out = np.full(f.shape, np.iinfo(np.int64).min, dtype=np.int64)
    out[ok] = f[ok].astype(np.int64)
    return out


def segment_hits(isBound, x0, y0, x1, y1, dx, dy, nx, ny, samples=8):
    """First bound node met by `samples` points along old->new (float32)."""
    hit = np.zeros(len(x0), dtype=bool)
    hix = nearest_index(x1, dx, nx)
    hiy = nearest_index(y1, dy, ny)
    for i in range(1, samples + 1):
        f = i / float(samples)
        xi = x0 + (x1 - x0) * f
        yi = y0 + (y1 - y0) * f
        cx = nearest_index(xi, dx, nx)
        cy = nearest_index(yi, dy, ny)
        step_hit = isBound[cy, cx]
        new = step_hit & ~hit
        hix[new] = cx[new]
        hiy[new] = cy[new]
        hit |= step_hit
    return hit, hix, hiy


def ppc_map(x, y, dx, dy, nx, ny):
    """Particles per node with truncating indices (int64[ny, nx])."""
    ix = np.clip(_trunc_i64(np.asarray(x) / dx), 0, nx - 1)
    iy = np.clip(_trunc_i64(np.asarray(y) / dy), 0, ny - 1)
    return np.bincount(iy * nx + ix, minlength=nx * ny).reshape(ny, nx)


def _trunc_i64(v):
    ok = (v >= -9.2e18) & (v < 9.2e18)
    out = np.full(v.shape, np.iinfo(np.int64).min, dtype=np.int64)
    out[ok] = v[ok].astype(np.int64)
    return out


def _neutral_density(x64, y64, Ly, n0, model):
    if model == "uniform":
        return np.full(x64.shape, float(n0))
    r_T = Ly * 1e-3
    a_corr = 1.0 / (1.0 - 1.0 / np.sqrt(2.0))
    z = np.maximum((x64 - 1.0) * 1e-3, 0.0)
    r = y64 * 1e-3
    zr = z + r_T
    R = np.sqrt(r * r + zr * zr)
    cos_t = zr / R
    q = r_T / np.maximum(R, 1e-12)
    n = (n0 * a_corr) * (1.0 - 1.0 / np.sqrt(1.0 + q * q)) * cos_t
    return np.maximum(n, 0.0)


def _sigma_det(g, m_ion, q, model):
    if model == "loglinear":
        E = (0.5 * m_ion * (g * g)) / q
        s = (87.3 - 13.6 * (det_log(E) * INV_LN10)) * 1e-20
        return np.maximum(s, 0.0)
    t = -0.8821 * det_log(g) + 15.1262
    return (t * t) * 1e-20


def mcc_prob_det64(x, y, vx, vy, vz, Ly, n0, dt, m_ion, q,
                   neutral_model="roy", cx_model="reference", cx_table=None):
    """CEX probability in float64 from float32 state (bit-exact twin of the
    HIP kernel).  cx_table: a user cross-section table (oracle.spline), which
    takes precedence over the built-in formulas as in the reference
    (Python/physics_engine.py:1714-1716)."""
    from .spline import table_value
    x64, y64 = np.asarray(x, np.float64), np.asarray(y, np.float64)
    a = np.asarray(vx, np.float64)
    b = np.asarray(vy, np.float64)
    c = np.asarray(vz, np.float64)
    n = _neutral_density(x64, y64, Ly, n0, neutral_model)
    g = np.maximum(np.sqrt(a * a + b * b + c * c), 1.0)
    if cx_table is not None:
        sigma = table_value(cx_table, (0.5 * m_ion * (g * g)) / q)
    else:
        sigma = _sigma_det(g, m_ion, q, cx_model)
    return 1.0 - det_exp(-(((n * sigma) * g) * dt))


def mcc_prob_ref32(x, y, vx, vy, vz, Ly, n0, dt, m_ion, q):
    """CEX probability with the reference's own float32 arithmetic
    (Roy neutral model, built-in cross-section)."""
    r_T = Ly * 1e-3
    z_m = np.maximum((x - 1.0) * 1e-3, 0.0)
    r_m = y * 1e-3
    a_corr = 1.0 / (1.0 - 1.0 / np.sqrt(2.0))
    R = np.sqrt(r_m ** 2 + (z_m + r_T) ** 2)
    theta = np.arctan2(r_m, z_m + r_T)
    n_local = n0 * a_corr * (1.0 - 1.0 / np.sqrt(1.0 + (r_T / np.maximum(R, 1e-12)) ** 2)) * np.cos(theta)
    n_local = np.maximum(n_local, 0.0)
    v_mag = np.sqrt(vx ** 2 + vy ** 2 + vz ** 2)
    g = np.maximum(v_mag, 1.0)
    sigma = ((-0.8821 * np.log(g) + 15.1262) ** 2) * 1e-20
    return 1.0 - np.exp(-n_local * sigma * g * dt)
