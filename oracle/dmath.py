"""Deterministic f64 transcendentals (test infrastructure, oracle side).

The reference evaluates the CEX collision probability, the Boltzmann-electron
term and its random draws with NumPy's libm (``Python/physics_engine.py:892``,
``:1717-1718``, ``:1057-1059``).  libm and the GPU's device library round the
last bit differently, so a decision such as ``rand < P`` could flip between
CPU and GPU.  Both sides therefore evaluate these functions with the same
sequence of IEEE-754 +, -, *, / and sqrt (all correctly rounded on x86 and on
CDNA4), which makes the results bit-identical.  The HIP twin is
``thruster_cx_code_amd/csrc/bemcs_dmath.h``; the operation order below must
stay identical to it.

Methods
-------
exp   Cody-Waite reduction x = k ln2 + r with a 32-bit-significand ln2_hi (so
      k*ln2_hi is exact), Taylor polynomial of degree 13 on |r| <= ln2/2
      (truncation < 5e-18), scaling by 2^k in two exact power-of-two steps.
log   x = m 2^e with m in [sqrt(1/2), sqrt(2)); log m = 2 atanh(s),
      s = (m-1)/(m+1), odd series to s^21 (truncation < 1e-19).
sincos(2 pi u), u in [0,1)
      exact quadrant reduction of 4u, Taylor sin/cos on [0, pi/4].

Pinned by: comparison with ``np.exp``/``np.log``/``np.sin``/``np.cos`` to a
few ulp over wide ranges (tests/test_oracle_rng.py), and bit equality with
the host build of the HIP header (tests/test_abi_host.py).
"""
import numpy as np

INV_LN2 = 1.4426950408889634
LN2_HI = 0.6931471803691238      # 0x1.62e42fee00000p-1 (32 significant bits)
LN2_LO = 1.9082149292705877e-10  # ln 2 - LN2_HI
SQRT_HALF = 0.7071067811865476
HALF_PI = 1.5707963267948966
INV_LN10 = 0.43429448190325176

# 1/k!, k = 0..13
EXP_C = (1.0, 1.0, 0.5, 0.16666666666666666, 0.041666666666666664,
         0.008333333333333333, 0.001388888888888889, 0.0001984126984126984,
         2.48015873015873e-05, 2.7557319223985893e-06, 2.755731922398589e-07,
         2.505210838544172e-08, 2.08767569878681e-09, 1.6059043836821613e-10)
# 1/(2k+1), k = 1..10
LOG_C = (0.3333333333333333, 0.2, 0.14285714285714285, 0.1111111111111111,
         0.09090909090909091, 0.07692307692307693, 0.06666666666666667,
         0.058823529411764705, 0.05263157894736842, 0.047619047619047616)
# (-1)^k/(2k+1)!, k = 1..8
SIN_C = (-0.16666666666666666, 0.008333333333333333, -0.0001984126984126984,
         2.7557319223985893e-06, -2.505210838544172e-08, 1.6059043836821613e-10,
         -7.647163731819816e-13, 2.8114572543455206e-15)
# (-1)^k/(2k)!, k = 1..8
COS_C = (-0.5, 0.041666666666666664, -0.001388888888888889,
         2.48015873015873e-05, -2.755731922398589e-07, 2.08767569878681e-09,
         -1.1470745597729725e-11, 4.779477332387385e-14)


def det_exp(x):
    x = np.asarray(x, dtype=np.float64)
    nan = np.isnan(x)
    xc = np.clip(np.where(nan, 0.0, x), -1100.0, 1100.0)
    k = np.rint(xc * INV_LN2)
    r = (xc - k * LN2_HI) - k * LN2_LO
    p = np.full_like(r, EXP_C[13])
    for c in EXP_C[12::-1]:
        p = p * r + c
    ki = k.astype(np.int64)
    a = ki >> 1
    b = ki - a
    with np.errstate(over="ignore", under="ignore"):
        out = (p * np.ldexp(1.0, a)) * np.ldexp(1.0, b)
    return np.where(nan, np.nan, out)


def det_log(x):
    x = np.asarray(x, dtype=np.float64)
    m, e = np.frexp(x)
    small = m < SQRT_HALF
    m = np.where(small, m * 2.0, m)
    e = e - small.astype(e.dtype)
    f = m - 1.0
    with np.errstate(divide="ignore", invalid="ignore"):
        s = f / (2.0 + f)
    z = s * s
    R = np.full_like(z, LOG_C[9])
    for c in LOG_C[8::-1]:
        R = R * z + c
    R = R * z
    two_s = 2.0 * s
    log_m = two_s + two_s * R
    ef = e.astype(np.float64)
    out = ef * LN2_HI + (ef * LN2_LO + log_m)
    out = np.where(x == 0.0, -np.inf, out)
    out = np.where(np.isinf(x) & (x > 0), np.inf, out)
    out = np.where((x < 0) | np.isnan(x), np.nan, out)
    return out


def det_sincos_2pi(u):
    """(sin(2 pi u), cos(2 pi u)) for u in [0, 1)."""
    u = np.asarray(u, dtype=np.float64)
    t = 4.0 * u
    q = np.floor(t)
    f = t - q
    flip = f > 0.5
    g = np.where(flip, 1.0 - f, f)
    a = g * HALF_PI
    a2 = a * a
    P = np.full_like(a2, SIN_C[7])
    for c in SIN_C[6::-1]:
        P = P * a2 + c
    s = a + (a * a2) * P
    Q = np.full_like(a2, COS_C[7])
    for c in COS_C[6::-1]:
        Q = Q * a2 + c
    co = 1.0 + a2 * Q
    s, co = np.where(flip, co, s), np.where(flip, s, co)
    qi = q.astype(np.int64) & 3
    sin_out = np.select([qi == 0, qi == 1, qi == 2], [s, co, -s], -co)
    cos_out = np.select([qi == 0, qi == 1, qi == 2], [co, -s, -co], s)
    return sin_out, cos_out


def box_muller(u0, u1):
    """Two standard normals from two uniforms in [0, 1)."""
    r = np.sqrt(-2.0 * det_log(1.0 - np.asarray(u0, np.float64)))
    s, c = det_sincos_2pi(u1)
    return r * c, r * s
