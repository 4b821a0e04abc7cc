"""OracleSimulator: the reference time step restated in NumPy
(test infrastructure, oracle side).

Follows ``DigitalTwinSimulator`` (Python/physics_engine.py:309-1784): domain
build (:629-847), step (:969-1784), with two deliberate, documented changes
that the MI355X backend shares:

* every random draw comes from the keyed Philox streams of ``oracle.rng``
  instead of the global MT19937 (``np.random``);
* the CEX probability and the SEE yield test use the float64 deterministic
  formulas of ``oracle.kernels.mcc_prob_det64`` (``mcc_math="det64"``); the
  reference's own float32 formula is available as ``mcc_math="ref32"``.

Electrode-surface physics (``sim_mode`` Thermal / Erosion / Both: impact
heating, sputter damage and cell failure, cooling, conduction, thermal bow,
:1359-1449 and :1563-1647) and the RF modulation (:984-991) follow
``oracle.surface``.  Heat and damage are added one ion at a time in array
order with ``np.add.at``, as the reference does; the device sums each node's
increments exactly (fixed point, independent of particle order) and adds the
sum once, which differs from this only by the rounding of the per-node adds.

Additive options (absent keys give the reference's behaviour):
``mesh = {"nx", "ny"}`` fixes the node count; ``preload = {"ions", "electrons"}``
seeds particles uniformly over non-electrode nodes (ions in x <= x_plume, where
the reference keeps them; electrons in x <= Lx); ``seed``; ``preload.hold`` tops both species up to their preloaded counts
after each step's injections (population hold);
``neutral_model in {"roy", "uniform"}``; ``cx_model in {"reference", "loglinear"}``;
``discharge_chamber.pitch_mm`` is honoured as the GUI does
(Python/main.py:1689).

Pins: each kernel is pinned in its own module.  The composition of the step
(injection bands, the order SEE spawn -> electron purge -> CEX -> PPC, the
entered / crossed / hit / out-of-bounds / keep masks, the bookkeeping of both
species, the saddle-point index, the preload placement and the electrode
rasterisation) is pinned by tests/test_oracle_step.py: closed-form
single-particle scenarios in a field-free domain, a scalar per-particle
restatement of sections D, E, H and the PPC map run against this class's
intermediate states over many steps (tests/literal_step.py), and planted
potentials.  Mistakes planted in this file (a mask threshold, a comparison, a
boundary, the order of the purge) fail those tests.  What stays open: the
reference's own MT19937 draws (replaced by Philox by design; the reference
cannot be imported here, it needs Taichi).
"""
import math
import numpy as np

from . import rng
from .kernels import (push_boris, deposit_counts, rho_from_counts, deposit_cic, rho_from_weights, substeps,
                      segment_hits, nearest_index, ppc_map, mcc_prob_det64,
                      mcc_prob_ref32)
from .poisson import SuperLUPoisson, efield
from .dmath import det_exp
from . import surface
from . import spline

F32 = np.float32
HOLD_ID_BIT = 0x40000000


class OracleSimulator:
    def __init__(self, dt=5e-10, mode="ref32", mcc_math="det64", quiet=True):
        self.mode = mode
        self.mcc_math = mcc_math
        self.quiet = quiet
        self.dt = dt
        self.q = 1.602e-19
        self.m_XE = 131.293 * 1.6605e-27
        self.m_ion = self.m_XE
        self.q_ion = self.q
        self.kB = 1.380649e-23
        self.eps0 = 8.854e-12
        self.m_e = self.m_XE / 100.0
        self.macro_weight = 3e5
        self.Lx, self.Ly = 3, 3
        self.dx, self.dy = 0.02, 0.02
        self.seed = 0
        self.iteration = 0
        self.injection_enabled = True
        self.injection_stop_time = None
        self.poisson = None
        self.next_ion_pid = 0
        self._reset_particles()
        self.transmitted_ions = 0.0
        self.entered_optics = 0.0
        self.injected_ions = 0.0
        # electrode material and thermal constants (:321-349); the cell heat
        # capacity and radiating area use dx, dy at set_material time (:450-452)
        self.sb_sigma = 5.67e-8
        self.thermal_accel = 1e7
        self.user_cs = {}
        self.deposit_shape = 'ngp'   # 'cic': the model manual's bilinear deposit
        self.set_material('Molybdenum')
        self.T_grids = []
        self.mask_grids = []
        self.grid_deflections = []
        self.Tmap = None

    MATERIAL_PRESETS = {
        'Molybdenum': {'k': 138.0, 'rho': 10280.0, 'cp': 250.0, 'emissivity': 0.80,
                       'alpha': 4.8e-6, 'E_mod': 329e9, 'Y_coeff': 1.05e-4, 'E_th': 30.0},
    }

    def set_material(self, name=None, props=None):
        mat = self.MATERIAL_PRESETS.get(name) if name else props
        if mat is None:
            return
        self.mat_k, self.mat_rho, self.mat_cp = mat['k'], mat['rho'], mat['cp']
        self.emissivity, self.alpha_thermal = mat['emissivity'], mat['alpha']
        self.sputter_Y_coeff, self.sputter_E_th = mat['Y_coeff'], mat['E_th']
        self.C_cell = self.mat_rho * (self.dx * 1e-3) * (self.dy * 1e-3) * 1e-3 * self.mat_cp
        self.A_cell = 2 * (self.dx * 1e-3) * 1e-3

    # ------------------------------------------------------------------
    def _fdt(self):
        return np.float32 if self.mode == "ref32" else np.float64

    def _reset_particles(self):
        f = self._fdt()
        self.p_x = np.zeros(0, f); self.p_y = np.zeros(0, f)
        self.p_vx = np.zeros(0, f); self.p_vy = np.zeros(0, f); self.p_vz = np.zeros(0, f)
        self.p_isCEX = np.zeros(0, bool); self.p_pid = np.zeros(0, np.uint32)
        self.e_x = np.zeros(0, f); self.e_y = np.zeros(0, f)
        self.e_vx = np.zeros(0, f); self.e_vy = np.zeros(0, f); self.e_vz = np.zeros(0, f)

    @property
    def num_p(self):
        return len(self.p_x)

    @property
    def num_e(self):
        return len(self.e_x)

    def _add_ions(self, x, y, vx, vy, vz, cex, pid):
        f = self._fdt()
        self.p_x = np.concatenate([self.p_x, np.asarray(x, f)])
        self.p_y = np.concatenate([self.p_y, np.asarray(y, f)])
        self.p_vx = np.concatenate([self.p_vx, np.asarray(vx, f)])
        self.p_vy = np.concatenate([self.p_vy, np.asarray(vy, f)])
        self.p_vz = np.concatenate([self.p_vz, np.asarray(vz, f)])
        self.p_isCEX = np.concatenate([self.p_isCEX, np.asarray(cex, bool)])
        self.p_pid = np.concatenate([self.p_pid, np.asarray(pid, np.uint32)])

    def _add_electrons(self, x, y, vx, vy, vz):
        f = self._fdt()
        self.e_x = np.concatenate([self.e_x, np.asarray(x, f)])
        self.e_y = np.concatenate([self.e_y, np.asarray(y, f)])
        self.e_vx = np.concatenate([self.e_vx, np.asarray(vx, f)])
        self.e_vy = np.concatenate([self.e_vy, np.asarray(vy, f)])
        self.e_vz = np.concatenate([self.e_vz, np.asarray(vz, f)])

    # ------------------------------------------------------------------
    def build_domain(self, params, preserve_state=False):
        """Python/physics_engine.py:629-847 (geometry, checks, masks, operator)."""
        grids = params.get("grids", [])
        total_t = sum(g['t'] + g['gap'] for g in grids) if grids else 0.0
        dc = params.get("discharge_chamber")
        pitch = params.get("pitch_mm", dc.get("pitch_mm", 0.0) if isinstance(dc, dict) else 0.0)
        Te_up = params.get('Te_up', 3.0)
        n0_n = params.get("n0_plasma", 1e17)
        bulk = params.get("entire_bulk_plasma", False)
        self.seed = int(params.get("seed", 0))
        self.deposit_shape = params.get("deposit_shape", "ngp")
        debye = np.sqrt(self.eps0 * Te_up / (self.q * n0_n * 0.61))
        if bulk:
            n_debye = 80 if n0_n <= 1e17 else (40 if n0_n <= 4e17 else 30)
            rec_gap = n_debye * debye * 1e3
        else:
            rec_gap = 0.75 * (grids[0]['r'] if grids else 0.80)
        user_gap = float(params.get('upstream_gap_mm', 0.0))
        gap = user_gap if user_gap > 0.0 else rec_gap
        gap = max(gap, self.dx * 2)
        self.upstream_gap_mm = gap
        self.Lx = gap + total_t + 3.0 if grids else params.get('Lx', self.Lx)
        geometry = params.get('geometry', 'half_hole')
        self.geometry = geometry
        self.periodic_y = geometry in ('one_hole', 'two_holes')
        if grids:
            sr = grids[0]['r']
            if geometry == 'two_holes':
                self.Ly = sr + 2.0 * sr + pitch
            elif geometry == 'one_hole':
                self.Ly = 3.0 * sr
            else:
                self.Ly = 0.5 * sr + 0.30 * pitch
        else:
            self.Ly = params.get('Ly', self.Ly)
        mesh = params.get('mesh')
        if mesh:
            self.nx, self.ny = int(mesh['nx']), int(mesh['ny'])
        else:
            self.nx = int(self.Lx / self.dx) + 1
            self.ny = int(self.Ly / self.dy) + 1
        self.dx = self.Lx / (self.nx - 1)
        self.dy = self.Ly / (self.ny - 1)
        if self.dx * 1e-3 > debye or self.dy * 1e-3 > debye:
            raise ValueError("Grid spacing too large for Debye resolution")
        wpi = np.sqrt(n0_n * self.q ** 2 / (self.m_ion * self.eps0))
        if 2 * 3.14159 / wpi < self.dt:
            raise ValueError("Too low time step")
        Ti = params.get('Ti', 0.1)
        vmax = np.sqrt(self.q_ion * Te_up / self.m_ion) + 4 * np.sqrt(self.q_ion * Ti / self.m_ion)
        if self.dx * 1e-3 / self.dt < vmax or self.dy * 1e-3 / self.dt < vmax:
            raise ValueError("Grid spacing and time step too large for velocity resolution")

        self.xpts = np.linspace(0, self.Lx, self.nx)
        self.ypts = np.linspace(0, self.Ly, self.ny)
        X, Y = np.meshgrid(self.xpts, self.ypts)
        if not preserve_state:
            self._reset_particles()
            self.iteration = 0
            self.next_ion_pid = 0
            self.injected_ions = self.transmitted_ions = self.entered_optics = 0.0
            self.V = np.zeros((self.ny, self.nx))
            self.rho = np.zeros((self.ny, self.nx), F32)
            self.Tmap = np.full((self.ny, self.nx), 300.0, dtype=F32)
            self.damage_map = np.zeros((self.ny, self.nx))
            self.eroded_depth = np.zeros((self.ny, self.nx))
        if len(self.grid_deflections) != len(grids):
            self.grid_deflections = [0.0] * len(grids)
        inj_time = params.get("inj_time", 0.0)
        self.injection_stop_time = inj_time if inj_time > 0.0 else None
        n0 = params.get('n0_plasma', 1e17)
        cell_vol = (self.dx * 1e-3) * (self.dy * 1e-3) * 1e-3
        self.macro_weight = max(n0 * cell_vol / 80.0, 1e3)
        self.isBound = np.zeros((self.ny, self.nx), bool)
        self.V_fixed = np.zeros((self.ny, self.nx))
        self.mask_grids = []
        self.T_grids = []
        if grids:
            sr = grids[0]['r']
            if geometry == 'two_holes':
                centers = [1.5 * sr, 1.5 * sr + pitch]
            elif geometry == 'one_hole':
                centers = [1.5 * sr]
            else:
                centers = [0.0]
        else:
            centers = []
        self.hole_centers = centers
        self.grid_x_starts, self.grid_x_ends = [], []
        cur = self.upstream_gap_mm
        for i, g in enumerate(grids):
            gs, ge = cur, cur + g["t"]
            self.grid_x_starts.append(gs)
            self.grid_x_ends.append(ge)
            dxb = 0.0
            d = self.grid_deflections[i] if i < len(self.grid_deflections) else 0.0
            if abs(d) > 1e-6:
                if geometry == 'two_holes':
                    eta = np.clip(1.0 - np.abs(Y - (1.5 * sr + 0.5 * pitch)) / max(pitch * 0.5, 1e-6), 0.0, 1.0)
                    dxb = d * eta ** 2
                elif geometry == 'one_hole':
                    eta = np.clip(1.0 - np.abs(Y - 1.5 * sr) / max(self.Ly * 0.5, 1e-6), 0.0, 1.0)
                    dxb = d * eta ** 2
                else:
                    Lc = self.Ly - g["r"]
                    dxb = d * np.clip((self.Ly - Y) / Lc, 0.0, 1.0) ** 2 if Lc > 0 else 0.0
            inx = (X >= gs + dxb) & (X <= ge + dxb)
            mask = inx.copy()
            for yc in centers:
                lr = np.maximum(g["r"] - np.maximum(0.0, X - gs - dxb) * np.tan(np.radians(g["cham"])), 0.0)
                mask &= ~(inx & (np.abs(Y - yc) <= lr))
            self.isBound[mask] = True
            self.V_fixed[mask] = g["V"]
            self.mask_grids.append(mask)
            self.T_grids.append(float(np.mean(self.Tmap[mask])) if np.any(mask) else 300.0)
            cur = ge + g["gap"]
        self.Vdc = self.V_fixed.copy()
        vpb = (grids[0]["V"] if grids else 1000.0) + params.get("V_plasma_offset", 20.0)
        self.V_fixed[:, 0] = vpb
        self.isBound[:, 0] = True
        if not preserve_state:
            self.Tmap[self.isBound] = 300.0
        self.poisson = SuperLUPoisson(self.isBound, self.periodic_y)
        self.recalc_poisson(30 if not preserve_state else 10, params)
        pre = params.get('preload')
        if not preserve_state:
            self.hold_ids = [0, 0]
            self.hold_targets = None
        if pre and not preserve_state:
            self._preload(params, int(pre.get('ions', 0)), int(pre.get('electrons', 0)))
            if pre.get('hold'):
                self.hold_targets = (int(pre.get('ions', 0)), int(pre.get('electrons', 0)))

    # ------------------------------------------------------------------
    def _preload_positions(self, purpose, count, x_max, first_id=0):
        pid = first_id + np.arange(count, dtype=np.uint64)
        x = np.zeros(count, F32)
        y = np.zeros(count, F32)
        todo = np.arange(count)
        for t in range(32):
            if len(todo) == 0:
                break
            u0, u1 = rng.uniforms(self.seed, purpose, pid[todo], t, 0)
            cx = (x_max * u0).astype(F32)
            cy = (self.Ly * u1).astype(F32)
            ok = ~self.isBound[nearest_index(cy, self.dy, self.ny), nearest_index(cx, self.dx, self.nx)]
            if t == 31:
                ok[:] = True
            x[todo[ok]], y[todo[ok]] = cx[ok], cy[ok]
            todo = todo[~ok]
        return pid, x, y

    def _preload(self, params, n_i, n_e, hold=False):
        """Preloaded particles; with hold, the top-up of the population hold
        (ids from the hold counters, streams P_HOLD_ION / P_HOLD_ELEC, ion
        pids 0x40000000 | id)."""
        Te_up = params.get('Te_up', 3.0)
        Ti = params.get('Ti', 0.1)
        if n_i > 0:
            grids = params.get('grids', [])
            x_max = self.grid_x_ends[-1] + grids[-1]["gap"] if grids else self.Lx
            purpose = rng.P_HOLD_ION if hold else rng.P_INIT_ION
            first = self.hold_ids[0] if hold else 0
            pid, x, y = self._preload_positions(purpose, n_i, x_max, first)
            vb = np.sqrt(self.q_ion * Te_up / self.m_ion)
            vs = np.sqrt(self.q_ion * Ti / self.m_ion)
            z0, z1 = rng.normals(self.seed, purpose, pid, 0, 1)
            z2, _ = rng.normals(self.seed, purpose, pid, 0, 2)
            if hold:
                tag = (np.uint64(HOLD_ID_BIT) | (pid & np.uint64(HOLD_ID_BIT - 1))).astype(np.uint32)
                self.hold_ids[0] = (self.hold_ids[0] + n_i) & 0xFFFFFFFF
            else:
                tag = pid.astype(np.uint32)
                self.next_ion_pid = n_i
            self._add_ions(x, y, (vb + z0 * vs).astype(F32), (z1 * vs).astype(F32),
                           (z2 * vs).astype(F32), np.zeros(n_i, bool), tag)
        if n_e > 0:
            purpose = rng.P_HOLD_ELEC if hold else rng.P_INIT_ELEC
            first = self.hold_ids[1] if hold else 0
            pid, x, y = self._preload_positions(purpose, n_e, self.Lx, first)
            vt = np.sqrt(2.0 * self.q * params.get('Te', 5.0) / self.m_e)
            z0, z1 = rng.normals(self.seed, purpose, pid, 0, 1)
            z2, _ = rng.normals(self.seed, purpose, pid, 0, 2)
            if hold:
                self.hold_ids[1] = (self.hold_ids[1] + n_e) & 0xFFFFFFFF
            self._add_electrons(x, y, (z0 * vt).astype(F32), (z1 * vt).astype(F32),
                                (z2 * vt).astype(F32))

    def _hold(self, params):
        """Population hold (preload {"hold": true}, a synthetic workload
        option with no reference counterpart): after the step's injections,
        top both species up to their preloaded counts."""
        if not self.hold_targets:
            return
        ti, te = self.hold_targets
        self._preload(params, max(ti - self.num_p, 0), max(te - self.num_e, 0), hold=True)

    # ------------------------------------------------------------------
    def recalc_poisson(self, iterations=5, params=None):
        params = params or {}
        coeff = (self.dx * 1e-3) ** 2 / self.eps0
        grids = params.get('grids', [{'V': 1000}])
        Vp = grids[0]['V'] + params.get('V_plasma_offset', 20.0)
        self.V = self.poisson.picard(self.V, self.rho, self.V_fixed, coeff, Vp,
                                     params.get('Te_up', 3.0), params.get('n0_plasma', 1e17),
                                     self.q, iterations)
        self.Ex, self.Ey = efield(self.V, self.dx, self.dy, self.periodic_y)
        self.Bx = np.zeros_like(self.Ex); self.By = np.zeros_like(self.Ex); self.Bz = np.zeros_like(self.Ex)

    # ------------------------------------------------------------------
    def _inject(self, params, t_current):
        grids = params.get('grids', [])
        if not (self.injection_enabled and (self.injection_stop_time is None
                                            or t_current <= self.injection_stop_time)):
            return
        n0 = params.get('n0_plasma', 1e17)
        Te_up = params.get('Te_up', 3.0)
        Ti = params.get('Ti', 0.1)
        vb = np.sqrt(self.q_ion * Te_up / self.m_ion)
        area = self.Ly * 1e-3 * 0.005
        bf = 1.0 if params.get('entire_bulk_plasma', False) else 0.61
        I_ion = self.q_ion * bf * n0 * vb * area
        cpm = self.q_ion * self.macro_weight
        lam = (I_ion * self.dt) / cpm
        n = rng.poisson_count(self.seed, rng.P_INJ_COUNT, self.iteration, lam)
        self.injected_ions_step = float(n)
        if n > 0:
            i = np.arange(n, dtype=np.uint64)
            u0, u1 = rng.uniforms(self.seed, rng.P_INJ_ION, i, self.iteration, 0)
            if grids:
                sr = grids[0]['r']
                if self.geometry == 'two_holes':
                    pitch = params.get('pitch_mm', 0.0)
                    lim = np.array([[max(self.dy, 0.5 * sr), min(2.5 * sr, self.Ly - self.dy)],
                                    [max(self.dy, 0.5 * sr + pitch), min(2.5 * sr + pitch, self.Ly - self.dy)]],
                                   dtype=F32)
                    w = (u1 >= 0.5).astype(np.int64)
                    lo, hi = lim[w, 0], lim[w, 1]
                    y = (lo + (hi - lo) * u0).astype(F32)
                elif self.geometry == 'one_hole':
                    lo, hi = max(self.dy, 0.5 * sr), min(2.5 * sr, self.Ly - self.dy)
                    y = (lo + (hi - lo) * u0).astype(F32)
                else:
                    lo, hi = 0.0, min(sr, self.Ly - self.dy)
                    y = (lo + (hi - lo) * u0).astype(F32)
            else:
                lo, hi = self.dy, self.Ly - self.dy
                y = (lo + (hi - lo) * u0).astype(F32)
            x = np.full(n, self.dx * 1.5, dtype=F32)
            vs = np.sqrt(self.q_ion * Ti / self.m_ion)
            z0, z1 = rng.normals(self.seed, rng.P_INJ_ION, i, self.iteration, 1)
            z2, _ = rng.normals(self.seed, rng.P_INJ_ION, i, self.iteration, 2)
            pid = (self.next_ion_pid + np.arange(n)).astype(np.uint32)
            self.next_ion_pid += n
            self._add_ions(x, y, (vb + z0 * vs).astype(F32), (z1 * vs).astype(F32),
                           (z2 * vs).astype(F32), np.zeros(n, bool), pid)
            self.injected_ions += float(n)
        if params.get('rf_enable'):
            vth_s = np.sqrt(2.0 * self.q * Te_up / self.m_e)
            I_e = self.q * 0.25 * n0 * vth_s * area
            nef = (I_e * self.dt) / cpm
            ne = int(nef)
            u, _ = rng.uniforms(self.seed, rng.P_RF_COUNT, 0, self.iteration, 0)
            if float(u) < (nef - ne):
                ne += 1
            if ne > 0:
                i = np.arange(ne, dtype=np.uint64)
                ylo = max(self.dy, 1e-9)
                yhi = max(ylo, self.Ly - ylo)
                u0, _ = rng.uniforms(self.seed, rng.P_RF_ELEC, i, self.iteration, 0)
                z0, z1 = rng.normals(self.seed, rng.P_RF_ELEC, i, self.iteration, 1)
                z2, _ = rng.normals(self.seed, rng.P_RF_ELEC, i, self.iteration, 2)
                self._add_electrons(np.full(ne, 0.1, F32), (ylo + (yhi - ylo) * u0).astype(F32),
                                    (np.abs(z0) * vth_s + vb).astype(F32), (z1 * vth_s).astype(F32),
                                    (z2 * vth_s).astype(F32))

    def _neutralizer(self, params):
        ne = int(params.get('neut_rate', 30))
        Te = params.get('Te', 5.0)
        nxp = float(np.clip(params.get('neut_x', self.Lx - 0.5), self.dx, self.Lx - self.dx))
        nrp = float(np.clip(params.get('neut_r', self.Ly), self.dy, self.Ly))
        if ne > 0:
            i = np.arange(ne, dtype=np.uint64)
            u0, _ = rng.uniforms(self.seed, rng.P_NEUT, i, self.iteration, 0)
            vt = np.sqrt(2.0 * self.q * Te / self.m_e)
            z0, z1 = rng.normals(self.seed, rng.P_NEUT, i, self.iteration, 1)
            z2, _ = rng.normals(self.seed, rng.P_NEUT, i, self.iteration, 2)
            self._add_electrons(np.full(ne, nxp, F32), (0.0 + (nrp - 0.0) * u0).astype(F32),
                                (z0 * vt).astype(F32), (z1 * vt).astype(F32), (z2 * vt).astype(F32))

    # ------------------------------------------------------------------
    def deposit(self):
        cell_vol = (self.dx * 1e-3) * (self.dy * 1e-3) * 1e-3
        cd = self.q * self.macro_weight / cell_vol
        if self.deposit_shape == 'cic':
            # bilinear deposit of the model manual (PY-BEMCS_Manual.tex:215-230)
            self.w_i = deposit_cic(self.p_x, self.p_y, self.dx, self.dy, self.nx, self.ny)
            self.w_e = deposit_cic(self.e_x, self.e_y, self.dx, self.dy, self.nx, self.ny)
            self.rho = rho_from_weights(self.w_i, cd, self.w_e, -cd).reshape(self.ny, self.nx)
            return
        self.cnt_i = deposit_counts(self.p_x, self.p_y, self.dx, self.dy, self.nx, self.ny)
        self.cnt_e = deposit_counts(self.e_x, self.e_y, self.dx, self.dy, self.nx, self.ny)
        self.rho = rho_from_counts(self.cnt_i, cd, self.cnt_e, -cd).reshape(self.ny, self.nx)

    def step(self, params):
        if self.poisson is None:
            return False, np.nan, np.nan, [], 0.0
        self.iteration += 1
        t_current = self.iteration * self.dt
        grids = params.get('grids', [])
        self.transmitted_ions_step = 0.0
        self.transmitted3_step = 0.0
        self.entered_optics_step = 0.0
        self.injected_ions_step = 0.0
        self.lost_to_grid_step = 0.0
        self.remeshed = False
        sim_mode = params.get('sim_mode', 'Both')

        # RF co-extraction (:984-991): the RF grid follows Vdc + A sin(2 pi f t)
        if params.get('rf_enable') and grids:
            rf_idx = params.get('rf_grid_idx', 0)
            if rf_idx < len(grids):
                v_rf = surface.rf_potential(params.get('rf_amp', 100.0), params.get('rf_freq', 13.56),
                                            t_current)
                m = self.mask_grids[rf_idx]
                self.V_fixed[m] = self.Vdc[m] + v_rf
                self.recalc_poisson(2, params)

        # A. injection
        self._inject(params, t_current)
        self._neutralizer(params)
        self._hold(params)
        # B. deposit + Poisson
        self.deposit()
        if self.iteration % 2 == 0:
            self.recalc_poisson(5, params)
        # C. substeps + push
        nps = self.num_p
        self.pushed = nps + self.num_e
        x_old, y_old = self.p_x.copy(), self.p_y.copy()
        qm_i = self.q_ion / self.m_ion
        qm_e = -self.q / self.m_e
        self.n_sub_ion = substeps(self.p_x, self.p_y, self.p_vx, self.p_vy, self.p_vz, self.Ex, self.Ey,
                                  self.dx, self.dy, self.nx, self.ny, qm_i, self.dt)[0] if nps else 1
        nes = self.num_e
        self.n_sub_e = substeps(self.e_x, self.e_y, self.e_vx, self.e_vy, self.e_vz, self.Ex, self.Ey,
                                self.dx, self.dy, self.nx, self.ny, qm_e, self.dt)[0] if nes else 1
        for _ in range(self.n_sub_ion if nps else 0):
            push_boris(self.p_x, self.p_y, self.p_vx, self.p_vy, self.p_vz, self.Ex, self.Ey, self.Bx,
                       self.By, self.Bz, nps, self.dx, self.dy, self.nx, self.ny,
                       self.dt / self.n_sub_ion, qm_i, mode=self.mode)
        for _ in range(self.n_sub_e if nes else 0):
            push_boris(self.e_x, self.e_y, self.e_vx, self.e_vy, self.e_vz, self.Ex, self.Ey, self.Bx,
                       self.By, self.Bz, nes, self.dx, self.dy, self.nx, self.ny,
                       self.dt / self.n_sub_e, qm_e, mode=self.mode)
        if self.periodic_y:
            if nps:
                self.p_y = np.mod(self.p_y, self.Ly)
            if self.num_e:
                self.e_y = np.mod(self.e_y, self.Ly)
        self._ion_diagnostics(params, grids, x_old, y_old, sim_mode)
        self._electron_purge()
        if sim_mode in ('Thermal', 'Both'):
            self._thermal(params, grids)
        min_pot = self._saddle(grids)
        self._mcc(params)
        self._ppc()
        trans = (self.transmitted3_step / self.entered_optics_step) if self.entered_optics_step > 0 else 0.0
        return self.remeshed, min_pot, self.current_div, self.T_grids, trans

    # ------------------------------------------------------------------
    def _surface(self, params, sim_mode, hit, p_x, p_vx, p_vy, p_vz):
        """Impact heating (:1359-1388) and sputter damage with cell failure
        (:1419-1449) of the ions that hit an electrode with x > 0.25 mm."""
        surf = hit & (p_x > 0.25)
        self.surface_hits = int(np.count_nonzero(surf))
        if not np.any(surf):
            return
        iy, ix = self.impact_iy[surf], self.impact_ix[surf]
        vx, vy, vz = p_vx[surf], p_vy[surf], p_vz[surf]
        if sim_mode in ('Thermal', 'Both'):
            dT = surface.impact_heat(vx, vy, vz, self.m_ion, self.macro_weight, self.C_cell,
                                     self.thermal_accel)
            np.add.at(self.Tmap, (iy, ix), dT)
        if sim_mode in ('Erosion', 'Both'):
            valid, dmg = surface.sputter_damage(vx, vy, vz, self.m_ion, self.q, self.sputter_E_th,
                                                self.sputter_Y_coeff, self.macro_weight)
            if np.any(valid):
                np.add.at(self.damage_map, (iy[valid], ix[valid]), dmg)
            broken = (self.damage_map > params.get('Thresh', 1e5)) & self.isBound
            if np.any(broken):
                self.eroded_depth[broken] += self.dy
                self.isBound[broken] = False
                self.damage_map[broken] = 0.0
                self.poisson = SuperLUPoisson(self.isBound, self.periodic_y)
                self.remeshed = True

    def _thermal(self, params, grids):
        """Radiative cooling, 10 conduction sweeps, 300 K floor, grid means and
        thermal bow with remesh (:1563-1647)."""
        cf = surface.cooling_factor(self.emissivity, self.sb_sigma, self.A_cell, self.dt,
                                    self.thermal_accel, self.C_cell)
        surface.radiative_cooling(self.Tmap, self.isBound, cf)
        Fo_x, Fo_y = surface.fourier_numbers(self.mat_k, self.mat_rho, self.mat_cp, self.dt,
                                             self.thermal_accel, self.dx, self.dy)
        self.Tmap = surface.conduction(self.Tmap, self.isBound, Fo_x, Fo_y, 10)
        surface.clamp_bound(self.Tmap, self.isBound)
        self.T_grids = surface.grid_means(self.Tmap, self.mask_grids, self.T_grids)
        dc = params.get("discharge_chamber")
        pitch = params.get("pitch_mm", dc.get("pitch_mm", 3.0) if isinstance(dc, dict) else 3.0)
        remesh = False
        for i, g in enumerate(grids):
            d = surface.bow_deflection(self.T_grids[i], g, self.geometry, self.Ly, pitch,
                                       self.alpha_thermal)
            if abs(d - self.grid_deflections[i]) > 0.005:
                self.grid_deflections[i] = d
                remesh = True
        if remesh:
            self.build_domain(params, preserve_state=True)
            self.remeshed = True

    # ------------------------------------------------------------------
    def _ion_diagnostics(self, params, grids, x_old, y_old, sim_mode='Trajectories'):
        p_x, p_y = self.p_x, self.p_y
        p_vx, p_vy, p_vz = self.p_vx, self.p_vy, self.p_vz
        if grids:
            x_entry = self.grid_x_starts[0]
            x_exit = self.grid_x_ends[-1]
            x_plume = self.grid_x_ends[-1] + grids[-1]["gap"]
        else:
            x_entry, x_exit, x_plume = 0.5, 3.0, 3.0
        entered = (x_old < x_entry) & (p_x >= x_entry) & (p_vx > 0.0)
        self.entered_optics_step = float(np.count_nonzero(entered))
        self.entered_optics += self.entered_optics_step
        ix = nearest_index(p_x, self.dx, self.nx)
        iy = nearest_index(p_y, self.dy, self.ny)
        hit_final = self.isBound[iy, ix]
        hit_path, hix, hiy = segment_hits(self.isBound, x_old, y_old, p_x, p_y, self.dx, self.dy,
                                          self.nx, self.ny)
        hit = hit_final | hit_path
        self.impact_ix = np.where(hit_path, hix, ix)
        self.impact_iy = np.where(hit_path, hiy, iy)
        oob = (p_x < 0.0) | (p_x > x_plume) | np.isnan(p_x)
        if not self.periodic_y:
            oob = oob | (p_y < 0.0) | (p_y > self.Ly)
        self.hit_mask = hit
        self.oob_mask = oob
        # SEE (all modes)
        see = hit & (p_x > 0.25)
        self.see_count = 0
        if np.any(see):
            sx, sy = p_x[see], p_y[see]
            svx, svy, svz = p_vx[see], p_vy[see], p_vz[see]
            pid = self.p_pid[see].astype(np.uint64)
            if self.mcc_math == "det64":
                v2 = (svx.astype(np.float64) ** 2 + svy.astype(np.float64) ** 2
                      + svz.astype(np.float64) ** 2)
                E_eV = (0.5 * self.m_ion * v2) / self.q
                see_table = spline.lookup(self.user_cs, 'SEE')
                if see_table is not None:                  # user yield table (:1398-1399)
                    gamma = np.clip(spline.table_value(see_table, E_eV), 0.0, 1.0)
                else:
                    gamma = np.clip(0.05 + 1e-4 * E_eV, 0.0, 1.0)
            else:
                E_eV = (0.5 * self.m_ion * (svx ** 2 + svy ** 2 + svz ** 2)) / self.q
                gamma = np.clip(0.05 + 1e-4 * E_eV, 0.0, 1.0)
            u, _ = rng.uniforms(self.seed, rng.P_SEE, pid, self.iteration, 0)
            spawn = u < gamma
            if np.any(spawn):
                ns = int(np.count_nonzero(spawn))
                ex = (sx[spawn] - svx[spawn] * self.dt * 1000.0 / 1.5).astype(F32)
                ey = (sy[spawn] - svy[spawn] * self.dt * 1000.0 / 1.5).astype(F32)
                vt = np.sqrt(2.0 * self.q * 2.0 / self.m_e)
                z0, z1 = rng.normals(self.seed, rng.P_SEE, pid[spawn], self.iteration, 1)
                z2, _ = rng.normals(self.seed, rng.P_SEE, pid[spawn], self.iteration, 2)
                self._add_electrons(ex, ey, (z0 * vt).astype(F32), (z1 * vt).astype(F32),
                                    (z2 * vt).astype(F32))
                self.see_count = ns
        self._surface(params, sim_mode, hit, p_x, p_vx, p_vy, p_vz)
        exited = (p_x > x_exit) & (p_vx > 0.0)
        crossed = (x_old <= x_exit) & (p_x > x_exit) & (p_vx > 0.0)
        self.crossed_mask = crossed
        nc = int(np.count_nonzero(crossed))
        self.exit_count_step = nc
        if nc > 0:
            vxe, vye, vze = p_vx[crossed], p_vy[crossed], p_vz[crossed]
            ve = np.sqrt(vxe ** 2 + vye ** 2 + vze ** 2)
            self.exit_vx_mean = float(np.mean(vxe))
            self.exit_v_mean = float(np.mean(ve))
            self.exit_vx_std = float(np.std(vxe))
            self.exit_v_std = float(np.std(ve))
            self.exit_energy_mean_eV = float(np.mean(0.5 * self.m_ion * ve ** 2 / self.q))
            cex = self.p_isCEX[crossed]
            self.exit_v_mean_primary = float(np.mean(ve[~cex])) if np.any(~cex) else np.nan
            self.exit_v_mean_cex = float(np.mean(ve[cex])) if np.any(cex) else np.nan
        else:
            self.exit_vx_mean = self.exit_v_mean = self.exit_vx_std = np.nan
            self.exit_v_std = self.exit_energy_mean_eV = np.nan
            self.exit_v_mean_primary = self.exit_v_mean_cex = np.nan
        self.crossed_angles = (np.abs(np.arctan2(p_vy[crossed], p_vx[crossed])) * 180.0 / np.pi)
        self.current_div = (float(np.percentile(self.crossed_angles, 95)) if nc > 5 else np.nan)
        if np.any(exited):
            self.transmitted_ions_step = float(nc)
            self.transmitted_ions += float(nc)
            self.transmitted3_step = float(nc)
        self.lost_to_grid_step = float(np.count_nonzero(hit & ~oob))
        alive = ~(hit | oob)
        keep = alive.copy()
        keep[alive] = (p_x[alive] <= x_plume) | (p_vx[alive] < 0.0)
        self.p_x, self.p_y = p_x[keep], p_y[keep]
        self.p_vx, self.p_vy, self.p_vz = p_vx[keep], p_vy[keep], p_vz[keep]
        self.p_isCEX = self.p_isCEX[keep]
        self.p_pid = self.p_pid[keep]

    def _electron_purge(self):
        if self.num_e == 0:
            return
        ix = nearest_index(self.e_x, self.dx, self.nx)
        iy = nearest_index(self.e_y, self.dy, self.ny)
        hit = self.isBound[iy, ix]
        out = ((self.e_x < 0.0) | (self.e_x > self.Lx) | (self.e_y < 0.0) | (self.e_y > self.Ly)
               | np.isnan(self.e_x))
        alive = ~(hit | out)
        self.e_x, self.e_y = self.e_x[alive], self.e_y[alive]
        self.e_vx, self.e_vy, self.e_vz = self.e_vx[alive], self.e_vy[alive], self.e_vz[alive]

    def _saddle(self, grids):
        if len(grids) >= 2:
            xg = 0.5 * (self.grid_x_starts[1] + self.grid_x_ends[1])
        elif len(grids) == 1:
            xg = 0.5 * (self.grid_x_starts[0] + self.grid_x_ends[0])
        else:
            xg = self.Lx * 0.5
        xi = int(np.clip(round(xg / self.dx), 0, self.nx - 1))
        if self.hole_centers:
            return min(self.V[int(np.clip(round(yc / self.dy), 0, self.ny - 1)), xi]
                       for yc in self.hole_centers)
        return self.V[self.ny // 2, xi]

    def _mcc(self, params):
        self.mcc_mask = np.zeros(self.num_p, bool)
        if self.num_p == 0:
            return
        prim = (~self.p_isCEX) & (self.p_x >= 1.0) & (self.p_x <= self.Lx)
        if not np.any(prim):
            return
        n0 = params.get('n0', 1e20)
        nm = params.get('neutral_model', 'roy')
        cm = params.get('cx_model', 'reference')
        xs, ys = self.p_x[prim], self.p_y[prim]
        vx, vy, vz = self.p_vx[prim], self.p_vy[prim], self.p_vz[prim]
        if self.mcc_math == "det64":
            P = mcc_prob_det64(xs, ys, vx, vy, vz, self.Ly, n0, self.dt, self.m_ion, self.q, nm, cm,
                               spline.lookup(self.user_cs, 'CX'))
        else:
            P = mcc_prob_ref32(xs, ys, vx, vy, vz, self.Ly, n0, self.dt, self.m_ion, self.q)
        pid = self.p_pid[prim].astype(np.uint64)
        u, _ = rng.uniforms(self.seed, rng.P_MCC, pid, self.iteration, 0)
        coll = u < P
        idx = np.where(prim)[0][coll]
        self.mcc_mask[idx] = True
        if len(idx) == 0:
            return
        pc = pid[coll]
        U = [rng.uniforms(self.seed, rng.P_MCC, pc, self.iteration, c) for c in range(1, 6)]
        R = [U[k // 2][k % 2] for k in range(10)]
        vth = np.sqrt(2.0 * self.kB * params.get('Tn', 300.0) / self.m_ion)
        fx = 2.0 * (R[0] + R[1] + R[2] - 1.5)
        fy = 2.0 * (R[3] + R[4] + R[5] - 1.5)
        fz = 2.0 * (R[6] + R[7] + R[8] - 1.5)
        self.p_vx[idx] = (vth * fx).astype(F32)
        self.p_vy[idx] = (vth * fy).astype(F32)
        self.p_vz[idx] = (vth * fz).astype(F32)
        self.p_isCEX[idx] = True

    def _ppc(self):
        if self.num_p:
            self.current_ppc_map = ppc_map(self.p_x, self.p_y, self.dx, self.dy, self.nx, self.ny)
        else:
            self.current_ppc_map = np.zeros((self.ny, self.nx), np.int64)
        self.total_active_cells = int(np.count_nonzero(self.current_ppc_map > 0))
        self.low_ppc_cells = int(np.count_nonzero((self.current_ppc_map > 0) & (self.current_ppc_map < 3)))
