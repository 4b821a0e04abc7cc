"""Scalar per-particle restatement of the step's bookkeeping sections, used to
pin the oracle's vectorised composition (test infrastructure).

Every function here walks particles (or nodes) one at a time with NumPy
float32 scalars and Python ints, following the reference text directly:

* ions after the push: entered / hit / OOB / crossed / SEE candidates /
  removal  (Python/physics_engine.py:1316-1529, segment samples :1880-1901);
* electron purge (:1534-1561);
* CEX primary selection and collision decision (:1682-1739);
* particles per cell with truncating indices (:1742-1758);
* preload placement (the additive synthetic-init option: the first of 32
  keyed candidates whose nearest node is not an electrode);
* electrode rasterisation of build_domain (:763-843).

Nothing here is shared with ``oracle/``: a shared misreading of a mask or an
index would have to be made twice, in two different formulations.
"""
import math

import numpy as np

F32 = np.float32


def nearest(v, d, n):
    """clip(round(v / d), 0, n - 1) for one float32 value: NumPy divides the
    float32 value by the Python float d rounded to float32, rounds half to
    even (Python's round does the same) and casts NaN / inf to INT64_MIN."""
    q = F32(v) / F32(d)
    if not math.isfinite(float(q)):
        return 0
    return min(max(int(round(float(q))), 0), n - 1)


def trunc_index(v, d, n):
    q = F32(v) / F32(d)
    if not math.isfinite(float(q)):
        return 0
    return min(max(int(float(q)), 0), n - 1)


def ion_section(sim, cap, grids):
    """Sections D of the step for every ion of `cap` (post-push state):
    returns per-ion outcome dicts and the step counters."""
    if grids:
        x_entry = sim.grid_x_starts[0]
        x_exit = sim.grid_x_ends[-1]
        x_plume = sim.grid_x_ends[-1] + grids[-1]["gap"]
    else:
        x_entry, x_exit, x_plume = 0.5, 3.0, 3.0
    xe, xx, xp = F32(x_entry), F32(x_exit), F32(x_plume)
    Ly32 = F32(sim.Ly)
    out = []
    cnt = dict(entered=0, crossed=0, lost=0, exited=0)
    for k in range(len(cap["x"])):
        xo, yo = F32(cap["x_old"][k]), F32(cap["y_old"][k])
        x, y = F32(cap["x"][k]), F32(cap["y"][k])
        vx = F32(cap["vx"][k])
        if xo < xe and x >= xe and vx > 0:
            cnt["entered"] += 1
        ix, iy = nearest(x, sim.dx, sim.nx), nearest(y, sim.dy, sim.ny)
        hit = bool(sim.isBound[iy, ix])
        imp = (iy, ix)
        path_hit = False
        for i in range(1, 9):
            f = i / 8.0
            xi = xo + (x - xo) * F32(f)
            yi = yo + (y - yo) * F32(f)
            cx, cy = nearest(xi, sim.dx, sim.nx), nearest(yi, sim.dy, sim.ny)
            if sim.isBound[cy, cx] and not path_hit:
                path_hit = True
                imp = (cy, cx)
        hit = hit or path_hit
        oob = bool(x < 0 or x > xp or math.isnan(float(x)))
        if not sim.periodic_y:
            oob = oob or bool(y < 0 or y > Ly32)
        crossed = bool(xo <= xx and x > xx and vx > 0)
        if x > xx and vx > 0:
            cnt["exited"] = 1
        cnt["crossed"] += int(crossed)
        if hit and not oob:
            cnt["lost"] += 1
        alive = not (hit or oob)
        keep = alive and (x <= xp or vx < 0)
        out.append(dict(pid=int(cap["pid"][k]), hit=hit, oob=oob, crossed=crossed, keep=keep,
                        see_candidate=bool(hit and x > F32(0.25)), impact=imp))
    return out, cnt


def electron_purge(sim, ex, ey):
    """Keep flags of the electron purge (:1534-1561)."""
    Lx32, Ly32 = F32(sim.Lx), F32(sim.Ly)
    keep = []
    for k in range(len(ex)):
        x, y = F32(ex[k]), F32(ey[k])
        ix, iy = nearest(x, sim.dx, sim.nx), nearest(y, sim.dy, sim.ny)
        hit = bool(sim.isBound[iy, ix])
        outb = bool(x < 0 or x > Lx32 or y < 0 or y > Ly32 or math.isnan(float(x)))
        keep.append(not (hit or outb))
    return np.array(keep, bool)


def mcc_primary(sim, x, cex):
    """Primary-ion selection of the CEX block (:1690): ~isCEX, 1 <= x <= Lx."""
    Lx32 = F32(sim.Lx)
    return [bool((not c) and F32(v) >= F32(1.0) and F32(v) <= Lx32) for v, c in zip(x, cex)]


def ppc_counts(sim, x, y):
    """{(iy, ix): count} with truncating indices (:1749-1758)."""
    out = {}
    for a, b in zip(x, y):
        key = (trunc_index(b, sim.dy, sim.ny), trunc_index(a, sim.dx, sim.nx))
        out[key] = out.get(key, 0) + 1
    return out


def electrode_mask(sim, grids, geometry, pitch):
    """Electrode map of build_domain (:763-843), one node at a time in Python
    floats: node coordinates i * L / (n - 1) as np.linspace gives them."""
    nx, ny = sim.nx, sim.ny
    Lx, Ly = sim.Lx, sim.Ly
    xs = np.linspace(0, Lx, nx)
    ys = np.linspace(0, Ly, ny)
    sr = grids[0]["r"]
    if geometry == "two_holes":
        centres = [1.5 * sr, 1.5 * sr + pitch]
    elif geometry == "one_hole":
        centres = [1.5 * sr]
    else:
        centres = [0.0]
    bound = [[False] * nx for _ in range(ny)]
    cur = sim.upstream_gap_mm
    for gi, g in enumerate(grids):
        gs, ge = cur, cur + g["t"]
        d = sim.grid_deflections[gi]
        tanc = math.tan(math.radians(g["cham"]))
        for iy in range(ny):
            Y = float(ys[iy])
            bow = 0.0
            if abs(d) > 1e-6:
                if geometry == "two_holes":
                    eta = 1.0 - abs(Y - (1.5 * sr + 0.5 * pitch)) / max(pitch * 0.5, 1e-6)
                elif geometry == "one_hole":
                    eta = 1.0 - abs(Y - 1.5 * sr) / max(Ly * 0.5, 1e-6)
                else:
                    Lc = Ly - g["r"]
                    eta = (Ly - Y) / Lc if Lc > 0 else None
                if eta is not None:
                    eta = min(max(eta, 0.0), 1.0)
                    bow = d * eta ** 2
            for ix in range(nx):
                X = float(xs[ix])
                if not (X >= gs + bow and X <= ge + bow):
                    continue
                in_hole = False
                for yc in centres:
                    r = max(g["r"] - max(0.0, X - gs - bow) * tanc, 0.0)
                    if abs(Y - yc) <= r:
                        in_hole = True
                if not in_hole:
                    bound[iy][ix] = True
        cur = ge + g["gap"]
    for iy in range(ny):
        bound[iy][0] = True
    return np.array(bound, bool)
