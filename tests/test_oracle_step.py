"""Pins of the oracle's step composition (OracleSimulator.step, the restatement
of Python/physics_engine.py:969-1784).

Three kinds of checks, each of which a plausible mistake in the composition
(a wrong mask, a swapped index, a missing term, a wrong order of the
sections) fails:

* closed-form single-particle scenarios in a field-free domain: the outcome of
  each (survives, entered, crossed, lost to a grid, out of bounds, purged,
  collided) follows from the geometry and the kinematics x1 = x0 + v dt;
* a scalar per-particle restatement (tests/literal_step.py) of sections D, E,
  H and the PPC map, run on the oracle's own intermediate states over many
  random steps, plus the particle bookkeeping identities
  num_p' = num_p + injected - removed and
  num_e' = num_e + injected electrons + SEE electrons - purged;
* the preload placement, the injection bands, the saddle-point index and the
  electrode rasterisation against literal loops and planted values.
"""
import math

import numpy as np
import pytest

from conftest import c1_params
from literal_step import (electrode_mask, electron_purge, ion_section, mcc_primary, nearest,
                          ppc_counts)
from oracle import rng
from oracle.kernels import mcc_prob_det64
from oracle.sim import OracleSimulator

F32 = np.float32


class Capture(OracleSimulator):
    """OracleSimulator that records its intermediate states."""

    def _inject(self, params, t_current):
        n0 = self.num_p
        ne0 = self.num_e
        super()._inject(params, t_current)
        self.cap_injected = self.num_p - n0
        self.cap_injected_e_rf = self.num_e - ne0

    def _neutralizer(self, params):
        ne0 = self.num_e
        super()._neutralizer(params)
        self.cap_neut = self.num_e - ne0

    def _ion_diagnostics(self, params, grids, x_old, y_old, sim_mode='Trajectories'):
        self.cap = dict(x_old=x_old.copy(), y_old=y_old.copy(), x=self.p_x.copy(), y=self.p_y.copy(),
                        vx=self.p_vx.copy(), vy=self.p_vy.copy(), vz=self.p_vz.copy(),
                        pid=self.p_pid.copy(), cex=self.p_isCEX.copy())
        self.cap_ne_before_see = self.num_e
        super()._ion_diagnostics(params, grids, x_old, y_old, sim_mode)
        self.cap_after_diag_pid = self.p_pid.copy()

    def _electron_purge(self):
        self.cap_e = (self.e_x.copy(), self.e_y.copy(), self.e_vx.copy(), self.e_vy.copy(),
                      self.e_vz.copy())
        super()._electron_purge()

    def _mcc(self, params):
        self.cap_mcc = dict(x=self.p_x.copy(), y=self.p_y.copy(), vx=self.p_vx.copy(),
                            vy=self.p_vy.copy(), vz=self.p_vz.copy(), pid=self.p_pid.copy(),
                            cex=self.p_isCEX.copy())
        super()._mcc(params)


def _sim(params, cls=OracleSimulator, m_e_ratio=200.0):
    o = cls()
    o.m_e = o.m_ion / m_e_ratio
    o.build_domain(params)
    return o


def _quiet(geometry="half_hole", nx=128, ny=48, **kw):
    """Field-free domain with no injection: the first step (odd) runs no
    Poisson solve, so particles move on straight lines."""
    p = c1_params(nx=nx, ny=ny, ions=0, electrons=0, neut_rate=0, geometry=geometry, **kw)
    o = _sim(p, Capture)
    o.injection_enabled = False
    o.Ex = np.zeros_like(o.Ex)
    o.Ey = np.zeros_like(o.Ey)
    return p, o


def _ions(o, rows, cex=None):
    """rows: (x, y, vx, vy, vz) per ion; pids 0..n-1."""
    a = np.array(rows, np.float64)
    n = len(a)
    o._add_ions(a[:, 0].astype(F32), a[:, 1].astype(F32), a[:, 2].astype(F32), a[:, 3].astype(F32),
                a[:, 4].astype(F32), np.zeros(n, bool) if cex is None else np.asarray(cex, bool),
                np.arange(n, dtype=np.uint32))


# ---------------------------------------------------------------- geometry
def test_quiet_geometry_constants():
    """The scenario positions below assume this two-grid layout."""
    p, o = _quiet()
    assert abs(o.grid_x_starts[0] - 0.7125) < 1e-12          # 0.75 * r_screen upstream gap
    assert abs(o.grid_x_ends[0] - 1.0925) < 1e-12
    assert abs(o.grid_x_starts[1] - 1.6925) < 1e-12
    assert abs(o.grid_x_ends[1] - 2.1925) < 1e-12            # x_exit
    assert abs(o.Lx - 6.1925) < 1e-12
    assert abs(o.Ly - (0.5 * 0.95 + 0.3 * 6.1447)) < 1e-12
    assert not o.periodic_y


# ------------------------------------------------------ closed-form scenarios
def test_field_free_ballistic_ion():
    p, o = _quiet()
    _ions(o, [(0.3, 0.1, 2e4, 1e3, 0.0)])
    o.step(p)
    assert o.n_sub_ion == 1                                   # |v| dt = 1e-5 m < 0.25 min(dx, dy)
    assert o.num_p == 1
    assert abs(float(o.p_x[0]) - (0.3 + 2e4 * 5e-10 * 1e3)) < 2e-7
    assert abs(float(o.p_y[0]) - (0.1 + 1e3 * 5e-10 * 1e3)) < 2e-7
    assert o.entered_optics_step == 0 and o.exit_count_step == 0 and o.lost_to_grid_step == 0


def test_ion_entering_optics_counts_once():
    p, o = _quiet()
    # crosses x_entry = 0.7125 inside the screen aperture (|y| < 0.95)
    _ions(o, [(0.705, 0.1, 2e4, 0, 0), (0.69, 0.2, 2e4, 0, 0), (0.72, 0.1, 2e4, 0, 0),
              (0.72, 0.1, -2e4, 0, 0)])
    o.step(p)
    assert o.entered_optics_step == 1
    assert o.num_p == 4 and o.lost_to_grid_step == 0


def test_ion_landing_exactly_on_entry_plane_counts():
    """x1 == f32(x_entry) exactly counts as entered (p_x >= x_entry)."""
    p, o = _quiet()
    xe = F32(0.7125)
    x0 = xe - F32(2.0 ** -9)
    dt32 = F32(o.dt)
    start = F32((float(xe) - float(x0)) / 1000.0 / float(dt32))
    found = None
    lo = hi = start
    for _ in range(20000):
        for v in (lo, hi):
            if x0 + (v * dt32) * F32(1000.0) == xe:
                found = v
                break
        if found is not None:
            break
        lo, hi = np.nextafter(lo, F32(0)), np.nextafter(hi, F32(np.inf))
    assert found is not None
    _ions(o, [(float(x0), 0.1, float(found), 0, 0)])
    o.step(p)
    assert o.p_x[0] == xe
    assert o.entered_optics_step == 1


def test_see_only_beyond_quarter_millimetre():
    """SEE candidates are hits with x > 0.25 mm: a fast ion (yield 1) ending
    next to the Dirichlet column at x = 0.02 is lost without SEE; one hitting a
    screen face placed at 0.3 mm (upstream_gap_mm) emits."""
    p, o = _quiet()
    _ions(o, [(0.1, 0.5, -1.6e5, 0, 0)])
    o.step(p)
    assert o.lost_to_grid_step == 1 and o.see_count == 0
    p = c1_params(nx=128, ny=48, ions=0, electrons=0, neut_rate=0)
    p["upstream_gap_mm"] = 0.3
    o = _sim(p, Capture)
    o.injection_enabled = False
    o.Ex = np.zeros_like(o.Ex)
    o.Ey = np.zeros_like(o.Ey)
    _ions(o, [(0.2, 1.5, 2.0e5, 0, 0)])                       # x1 = 0.3 on the screen face
    o.step(p)
    assert o.lost_to_grid_step == 1 and o.see_count == 1


def test_ion_crossing_exit_plane():
    """Crossing x_exit counts once in `crossed` and as transmitted; the ion
    stays in the domain."""
    p, o = _quiet()
    _ions(o, [(2.188, 0.1, 2e4, 0, 0), (2.30, 0.1, 2e4, 0, 0), (2.188, 0.1, -2e4, 0, 0)])
    o.step(p)
    assert o.exit_count_step == 1
    assert o.transmitted_ions_step == 1.0
    assert o.num_p == 3
    assert math.isnan(o.current_div)                          # 5 crossings or fewer


def test_theta95_of_known_exit_angles():
    """20 ions leave with angles 1..20 degrees: np.percentile's linear rule
    gives 19 + 0.05 = 19.05 degrees."""
    p, o = _quiet()
    rows = []
    for k in range(1, 21):
        th = math.radians(k)
        rows.append((2.188, 0.1, 2e4 * math.cos(th), 2e4 * math.sin(th), 0.0))
    _ions(o, rows)
    o.step(p)
    assert o.exit_count_step == 20
    assert abs(o.current_div - 19.05) < 1e-4


def test_ion_leaving_through_plume_boundary_is_not_a_grid_loss():
    p, o = _quiet()
    _ions(o, [(3.19, 0.5, 2e4, 0, 0)])                       # x_plume = 3.1925
    o.step(p)
    assert o.num_p == 0 and o.lost_to_grid_step == 0


def test_ion_hitting_accel_grid():
    """Lands on an accel-grid node (outside the 0.55 mm aperture): removed,
    counted once in lost_to_grid; its SEE electron, if drawn, starts 2/3 of a
    step back along the path."""
    p, o = _quiet()
    _ions(o, [(1.685, 0.8, 2e4, 0, 0)])
    o.step(p)
    assert o.num_p == 0 and o.lost_to_grid_step == 1
    E = 0.5 * o.m_ion * (2e4) ** 2 / o.q
    gamma = min(max(0.05 + 1e-4 * E, 0.0), 1.0)
    u, _ = rng.uniforms(o.seed, rng.P_SEE, np.uint64(0), o.iteration, 0)
    assert o.see_count == int(float(u) < gamma)


def test_see_electron_position_and_purge():
    """Dense SEE: a fast ion with yield 1 (E > 9500 eV) hits the accel grid
    face; its SEE electron starts at x1 - v dt 1000 / 1.5, which lies in the
    gap, so it survives the electron purge."""
    p, o = _quiet()
    v = 2.0e5                                                 # E = 0.5 m v^2 / q ~ 27 keV
    _ions(o, [(1.60, 0.8, v, 0, 0)])
    o.step(p)
    assert o.lost_to_grid_step == 1 and o.see_count == 1
    assert o.num_e == 1
    x1 = 1.60 + v * 5e-10 * 1e3
    assert abs(float(o.e_x[0]) - (x1 - v * 5e-10 * 1e3 / 1.5)) < 1e-5
    assert abs(float(o.e_y[0]) - 0.8) < 1e-6


def test_see_electron_inside_electrode_is_purged():
    """A slow ion deep in the accel slab: its SEE electron also starts inside
    the electrode and the purge (which runs after SEE spawning) removes it."""
    p, o = _quiet()
    v = 2.0e5
    _ions(o, [(1.85, 0.8, v, 0, 0)])                          # x1 = 1.95, x_see = 1.883
    o.step(p)
    assert o.lost_to_grid_step == 1 and o.see_count == 1
    assert o.num_e == 0


def test_ion_passing_through_thin_grid_is_caught_by_segment_samples():
    """1e6 m/s carries the ion from 0.69 mm across the 0.38 mm screen slab
    (|y| = 1.5 > r) into the gap in one step: the final node is free, a
    segment sample is not."""
    p, o = _quiet()
    _ions(o, [(0.69, 1.5, 1e6, 0, 0)])
    o.step(p)
    assert o.n_sub_ion == math.ceil(1e6 * 5e-10 / (0.25 * min(o.dx, o.dy) * 1e-3))
    assert o.num_p == 0 and o.lost_to_grid_step == 1


def test_out_of_bounds_at_upstream_wall_is_not_a_grid_loss():
    """x < 0 is out of bounds; its nearest node is the Dirichlet column 0
    (isBound), so hit & ~oob keeps lost_to_grid at 0."""
    p, o = _quiet()
    _ions(o, [(0.004, 0.5, -2e4, 0, 0)])
    o.step(p)
    assert o.num_p == 0 and o.lost_to_grid_step == 0


def test_y_walls_remove_in_half_hole_and_wrap_in_one_hole():
    p, o = _quiet()
    _ions(o, [(0.3, float(o.Ly) - 0.003, 0, 2e4, 0), (0.3, 0.003, 0, -2e4, 0)])
    o.step(p)
    assert o.num_p == 0 and o.lost_to_grid_step == 0
    p, o = _quiet(geometry="one_hole", nx=130, ny=60)
    Ly = float(o.Ly)
    _ions(o, [(0.3, Ly - 0.003, 0, 2e4, 0), (0.3, 0.003, 0, -2e4, 0)])
    o.step(p)
    assert o.num_p == 2
    ys = sorted(float(v) for v in o.p_y)
    assert abs(ys[0] - 0.007) < 1e-5 and abs(ys[1] - (Ly - 0.007)) < 1e-5


def test_electron_purge_cases():
    p, o = _quiet()
    Lx = float(o.Lx)
    rows = [(Lx - 0.003, 1.0, 2e4, 0, 0),     # leaves through x = Lx
            (4.0, 1.0, 0, 0, 0),              # stays
            (2.0, 1.5, 0, 0, 0),              # accel electrode node
            (4.0, 0.003, 0, -2e4, 0),         # leaves through y = 0
            (0.01, 1.0, 0, 0, 0)]             # nearest node is the Dirichlet column
    a = np.array(rows)
    o._add_electrons(*(a[:, k].astype(F32) for k in range(5)))
    o.step(p)
    assert o.num_e == 1
    assert float(o.e_x[0]) == F32(4.0)


def test_mcc_primary_selection_with_certain_collisions():
    """P = 1 (uniform neutrals, n0 = 1e30): every primary ion with
    1 <= x <= Lx collides, ions upstream of 1 mm and CEX ions never do; new
    velocities are 2 vth (R1+R2+R3-1.5), inside [-3 vth, 3 vth]."""
    p, o = _quiet()
    p["n0"] = 1e30
    _ions(o, [(2.5, 0.3, 2e4, 0, 0), (0.5, 0.3, 2e4, 0, 0), (2.6, 0.3, 2e4, 0, 0),
              (3.0, 0.3, 2e4, 0, 0)], cex=[False, False, True, False])
    o.step(p)
    order = np.argsort(o.p_pid)
    cex = o.p_isCEX[order]
    assert list(cex) == [True, False, True, True]
    vth = math.sqrt(2.0 * o.kB * 300.0 / o.m_ion)
    for k in (0, 3):
        v = np.array([o.p_vx[order][k], o.p_vy[order][k], o.p_vz[order][k]], np.float64)
        assert np.all(np.abs(v) <= 3 * vth * (1 + 1e-6))
    assert float(o.p_vx[order][1]) == F32(2e4)              # not a primary in range
    assert float(o.p_vx[order][2]) == F32(2e4)              # already CEX
    p["n0"] = 0.0
    p2, o2 = _quiet()
    p2["n0"] = 0.0
    _ions(o2, [(2.5, 0.3, 2e4, 0, 0)])
    o2.step(p2)
    assert not o2.p_isCEX[0]


def test_ppc_map_of_placed_ions():
    p, o = _quiet()
    _ions(o, [(0.3, 0.1, 0, 0, 0), (0.3, 0.1, 0, 0, 0), (3.0, 2.0, 0, 0, 0)])
    o.step(p)
    ix0, iy0 = int(float(F32(0.3) / F32(o.dx))), int(float(F32(0.1) / F32(o.dy)))
    ix1, iy1 = int(float(F32(3.0) / F32(o.dx))), int(float(F32(2.0) / F32(o.dy)))
    m = o.current_ppc_map
    assert m[iy0, ix0] == 2 and m[iy1, ix1] == 1 and m.sum() == 3
    assert o.total_active_cells == 2 and o.low_ppc_cells == 2


@pytest.mark.parametrize("geometry", ["half_hole", "two_holes"])
def test_saddle_point_index_by_planted_potential(geometry):
    """min_pot is V at the hole centre(s) on the column nearest the middle of
    the second grid; plant values there and next to it."""
    kw = {"pitch_mm": 3.0} if geometry == "two_holes" else {}
    nx, ny = (128, 128) if geometry == "two_holes" else (128, 48)
    p = c1_params(nx=nx, ny=ny, ions=0, electrons=0, neut_rate=0, geometry=geometry)
    p.update(kw)
    o = _sim(p)
    o.injection_enabled = False
    V = np.zeros_like(o.V)
    xi = int(round(0.5 * (1.6925 + 2.1925) / o.dx))
    sr = 0.95
    rows = [0] if geometry == "half_hole" else [int(round(1.5 * sr / o.dy)),
                                                int(round((1.5 * sr + 3.0) / o.dy))]
    vals = [-7.0] if geometry == "half_hole" else [-3.0, -5.0]
    for r, v in zip(rows, vals):
        V[r, xi] = v
        V[r, xi + 1] = -100.0
        V[min(r + 1, o.ny - 1), xi] = -100.0
    o.V = V
    res = o.step(p)                                           # odd step: no solve
    assert res[1] == min(vals)


# -------------------------------------------- literal restatement over steps
def _check_step_against_literal(o, p, grids):
    before_p, before_e = o.num_p, o.num_e
    o.step(p)
    cap = o.cap
    outs, cnt = ion_section(o, cap, grids)
    assert cnt["entered"] == o.entered_optics_step
    assert cnt["crossed"] == o.exit_count_step
    assert cnt["lost"] == o.lost_to_grid_step
    assert (cnt["crossed"] if cnt["exited"] else 0) == o.transmitted_ions_step
    kept = sorted(d["pid"] for d in outs if d["keep"])
    assert kept == sorted(int(v) for v in o.cap_after_diag_pid)
    removed = sum(1 for d in outs if not d["keep"])
    assert o.num_p == before_p + o.cap_injected - removed
    # SEE candidates and spawn decisions (yield 0.05 + 1e-4 E, keyed uniform)
    spawned = 0
    for k, d in enumerate(outs):
        if not d["see_candidate"]:
            continue
        v2 = sum(float(cap[c][k]) ** 2 for c in ("vx", "vy", "vz"))
        E = (0.5 * o.m_ion * v2) / o.q
        gamma = min(max(0.05 + 1e-4 * E, 0.0), 1.0)
        u, _ = rng.uniforms(o.seed, rng.P_SEE, np.uint64(d["pid"]), o.iteration, 0)
        spawned += int(float(u) < gamma)
    assert spawned == o.see_count
    # electron purge on the electrons present after SEE spawning
    ex, ey = o.cap_e[0], o.cap_e[1]
    assert len(ex) == before_e + o.cap_injected_e_rf + o.cap_neut + o.see_count
    keep_e = electron_purge(o, ex, ey)
    assert o.num_e == int(keep_e.sum())
    a = np.sort(np.stack([ex[keep_e], ey[keep_e]]), axis=1)
    b = np.sort(np.stack([o.e_x, o.e_y]), axis=1)
    assert np.array_equal(a, b)
    # CEX: primary selection and the collision decision per ion
    m = o.cap_mcc
    prim = mcc_primary(o, m["x"], m["cex"])
    nm, cm = p.get("neutral_model", "roy"), p.get("cx_model", "reference")
    n_coll = 0
    for k in range(len(m["x"])):
        if not prim[k]:
            assert not o.mcc_mask[k]
            continue
        P = mcc_prob_det64(m["x"][k:k + 1], m["y"][k:k + 1], m["vx"][k:k + 1], m["vy"][k:k + 1],
                           m["vz"][k:k + 1], o.Ly, p.get("n0", 1e20), o.dt, o.m_ion, o.q, nm, cm)[0]
        u, _ = rng.uniforms(o.seed, rng.P_MCC, np.uint64(m["pid"][k]), o.iteration, 0)
        c = bool(float(u) < P)
        assert c == bool(o.mcc_mask[k])
        n_coll += int(c)
        if c:
            assert o.p_isCEX[k]
    # PPC map with truncating indices
    lit = ppc_counts(o, o.p_x, o.p_y)
    m = o.current_ppc_map
    assert int(m.sum()) == o.num_p
    for (iy, ix), c in lit.items():
        assert m[iy, ix] == c
    assert len(lit) == o.total_active_cells
    return n_coll


@pytest.mark.parametrize("geometry,nx,ny,n0,models", [
    ("half_hole", 128, 48, 1e18, ("uniform", "loglinear")),
    ("half_hole", 128, 48, 3e22, ("roy", "reference")),
    ("one_hole", 128, 60, 3e22, ("uniform", "loglinear")),
])
def test_step_sections_equal_literal_restatement(geometry, nx, ny, n0, models):
    """Six steps (odd and even, Poisson included) with injection, the
    neutraliser and preloaded particles: every section of every step equals
    the scalar restatement, and the bookkeeping identities hold."""
    p = c1_params(nx=nx, ny=ny, ions=1500, electrons=1500, geometry=geometry)
    p["n0"] = n0
    p["neutral_model"], p["cx_model"] = models
    o = _sim(p, Capture)
    n_coll = 0
    for _ in range(6):
        n_coll += _check_step_against_literal(o, p, p["grids"])
    if n0 > 1e20:
        assert n_coll > 20


def test_two_holes_injection_bands_and_step():
    """two_holes: injected ions fall in the two aperture bands (one per hole,
    chosen by a fair draw), and a step passes the literal restatement."""
    p = c1_params(nx=128, ny=120, ions=500, electrons=500, geometry="two_holes")
    p["pitch_mm"] = 3.0
    p["n0_plasma"] = 1e17
    o = _sim(p, Capture)
    sr, pitch, dy, Ly = 0.95, 3.0, o.dy, o.Ly
    bands = [(max(dy, 0.5 * sr), min(2.5 * sr, Ly - dy)),
             (max(dy, 0.5 * sr + pitch), min(2.5 * sr + pitch, Ly - dy))]
    ys = []
    for _ in range(4):
        n_before = o.next_ion_pid
        _check_step_against_literal(o, p, p["grids"])
        new = o.cap["pid"] >= n_before
        ys.extend(float(v) for v in o.cap["y"][new])
    assert len(ys) > 100
    ys = np.array(ys)
    # injected positions moved for one step (|vy| dt ~ 1e-4 mm)
    in0 = (ys >= bands[0][0] - 1e-3) & (ys <= bands[0][1] + 1e-3)
    in1 = (ys >= bands[1][0] - 1e-3) & (ys <= bands[1][1] + 1e-3)
    assert np.all(in0 | in1)
    frac = in1.mean()
    assert abs(frac - 0.5) < 4 * math.sqrt(0.25 / len(ys))


def test_half_hole_injection_band():
    p = c1_params(nx=128, ny=48, ions=0, electrons=0, neut_rate=0)
    o = _sim(p, Capture)
    o.Ex = np.zeros_like(o.Ex)
    o.Ey = np.zeros_like(o.Ey)
    o.step(p)                                                 # odd: no solve, E = 0
    n = o.cap_injected
    lam = (o.q * 0.61 * 1e17 * math.sqrt(o.q * 3.0 / o.m_ion) * (o.Ly * 1e-3 * 0.005) * o.dt) / (
        o.q * o.macro_weight)
    assert n == rng.poisson_count(o.seed, rng.P_INJ_COUNT, 1, lam)
    assert n > 0
    y0 = o.cap["y_old"]
    assert np.all(y0 >= 0.0) and np.all(y0 <= min(0.95, o.Ly - o.dy) + 1e-6)
    assert np.all(o.cap["x_old"] == F32(o.dx * 1.5))


# ------------------------------------------------ preload and rasterisation
def test_preload_placement_literal():
    """Each preloaded particle sits at the first of its 32 keyed candidates
    whose nearest node is not an electrode; ions within x <= x_plume."""
    p = c1_params(nx=128, ny=48, ions=400, electrons=400, neut_rate=0)
    o = _sim(p)
    x_plume = o.grid_x_ends[-1] + p["grids"][-1]["gap"]
    for purpose, xs, ys, xmax in ((rng.P_INIT_ION, o.p_x, o.p_y, x_plume),
                                  (rng.P_INIT_ELEC, o.e_x, o.e_y, o.Lx)):
        for k in range(len(xs)):
            for t in range(32):
                u0, u1 = rng.uniforms(o.seed, purpose, np.uint64(k), t, 0)
                cx, cy = F32(xmax * float(u0)), F32(o.Ly * float(u1))
                if not o.isBound[nearest(cy, o.dy, o.ny), nearest(cx, o.dx, o.nx)] or t == 31:
                    break
            assert xs[k] == cx and ys[k] == cy
    # velocity moments: ions drift at the Bohm speed with the Ti spread
    vb = math.sqrt(o.q * 3.0 / o.m_ion)
    vs = math.sqrt(o.q * 0.2 / o.m_ion)
    assert abs(np.mean(o.p_vx) - vb) < 5 * vs / math.sqrt(400)
    assert abs(np.std(o.p_vy) / vs - 1) < 0.15


@pytest.mark.parametrize("geometry,cham,defl", [
    ("half_hole", 0.0, 0.0), ("half_hole", 30.0, 0.02), ("one_hole", 15.0, -0.03),
    ("two_holes", 20.0, 0.04)])
def test_electrode_rasterisation_literal(geometry, cham, defl):
    p = c1_params(nx=120, ny=140, ions=0, electrons=0, geometry=geometry)
    p["pitch_mm"] = 3.0
    for g in p["grids"]:
        g["cham"] = cham
    o = OracleSimulator()
    o.m_e = o.m_ion / 200
    o.build_domain(p)
    if defl:
        o.grid_deflections = [defl, -0.5 * defl]
        o.build_domain(p, preserve_state=True)
    lit = electrode_mask(o, p["grids"], geometry, 3.0)
    assert np.array_equal(lit, o.isBound)
    assert lit.sum() > 500


def test_population_hold_tops_up_both_species():
    """preload {"hold": true}: every step pushes exactly the preloaded
    populations (after the step's injections the counts are topped up), the
    added ions carry fresh ids with bit 30 set, and their positions follow the
    preload's placement law from their own streams."""
    p = c1_params(nx=128, ny=48, ions=3000, electrons=3000)
    p["preload"]["hold"] = True
    o = _sim(p, Capture)
    seen = set(int(v) for v in o.p_pid)
    for _ in range(4):
        before_ids = list(o.hold_ids)
        o.step(p)
        assert o.pushed >= 3000 + 3000
        assert len(o.cap["x"]) >= 3000
        new = o.cap["pid"] >= 0x40000000
        ids = (o.cap["pid"][new] & 0x3FFFFFFF).astype(np.int64)
        assert len(ids) == o.hold_ids[0] - before_ids[0]
        assert np.array_equal(np.sort(ids), np.arange(before_ids[0], o.hold_ids[0]))
        assert not (set(int(v) for v in o.cap["pid"][new]) & seen)
        seen |= set(int(v) for v in o.cap["pid"])
    # placement of the last top-up, literal
    x_plume = o.grid_x_ends[-1] + p["grids"][-1]["gap"]
    k_new = np.where(new)[0][:50]
    for k in k_new:
        i = int(o.cap["pid"][k]) & 0x3FFFFFFF
        for t in range(32):
            u0, u1 = rng.uniforms(o.seed, rng.P_HOLD_ION, np.uint64(i), t, 0)
            cx, cy = F32(x_plume * float(u0)), F32(o.Ly * float(u1))
            if not o.isBound[nearest(cy, o.dy, o.ny), nearest(cx, o.dx, o.nx)] or t == 31:
                break
        assert o.cap["x_old"][k] == cx and o.cap["y_old"][k] == cy
