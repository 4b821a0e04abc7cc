"""Sweep workloads on the batched engine: case generation, the autotuner (sort
intervals and deposit tile), and batched vs one-at-a-time timing.

The reference sweeps (Python/benchmarks/benchmark_perveance.py:11-75,
benchmark_ebs.py, benchmark_impingement.py) build one DigitalTwinSimulator
per operating point and step it to steady state, one process per case.
Here a sweep is a list of params dicts; the cases advance in lockstep on one
GPU with one launch per step phase for all of them (batch.BatchedSimulator).
"""
import math
import time

import numpy as np
import torch


def two_grid_params(nx, ny, V_screen=1000.0, V_accel=-200.0, r_accel=0.55, ions=0, electrons=0,
                    seed=0):
    """Two-grid Xe+ extraction case: screen t 0.38 mm, r 0.95 mm; accel
    t 0.5 mm; 0.6 mm gap; Ti 0.2 eV, neutraliser Te 5 eV, uniform neutral Xe
    1e18 m^-3 at 300 K, log-linear CEX cross-section, half_hole geometry."""
    return {
        "grids": [{"V": float(V_screen), "t": 0.38, "gap": 0.6, "r": 0.95, "cham": 0.0},
                  {"V": float(V_accel), "t": 0.5, "gap": 1.0, "r": float(r_accel), "cham": 0.0}],
        "pitch_mm": 6.1447, "geometry": "half_hole", "V_plasma_offset": 20.0,
        "n0_plasma": 1e17, "Te_up": 3.0, "Ti": 0.2, "Te": 5.0, "Tn": 300.0,
        "n0": 1e18, "neutral_model": "uniform", "cx_model": "loglinear",
        "neut_rate": 30, "sim_mode": "Trajectories", "mesh": {"nx": int(nx), "ny": int(ny)},
        "preload": {"ions": int(ions), "electrons": int(electrons)}, "seed": int(seed),
    }


def uneven_sweep(n_cases=64, seed=0, n_min=200_000, n_max=4_000_000):
    """Uneven-mesh sweep: nx ~ U{256..1024 step 64}, ny = 3 nx / 8, accel
    voltage ~ U[-400, -100] V, accel aperture radius ~ U[0.45, 0.75] mm,
    screen voltage ~ U[800, 1600] V, particles per species log-uniform in
    [n_min, n_max]; case seed = case index."""
    r = np.random.default_rng(seed)
    out = []
    for c in range(n_cases):
        nx = int(r.integers(4, 17)) * 64
        ny = 3 * nx // 8
        n = int(round(math.exp(r.uniform(math.log(n_min), math.log(n_max)))))
        out.append(two_grid_params(nx, ny, V_screen=r.uniform(800, 1600), V_accel=r.uniform(-400, -100),
                                   r_accel=r.uniform(0.45, 0.75), ions=n, electrons=n, seed=c))
    return out


def build_case(params, m_e_ratio=200.0, coarse_max=None, level0_tile=None):
    """A built case; coarse_max bounds the coarsest multigrid level (nodes,
    None: the solver's default); level0_tile sets the level-0 tile edge of its
    Poisson solves (None: by mesh size; 16 suits cases stepped in large
    batches, and gives the same bits alone and batched)."""
    from .physics_engine import DigitalTwinSimulator
    s = DigitalTwinSimulator()
    s.m_e = s.m_ion / m_e_ratio
    if coarse_max is not None:
        s.poisson_coarse_max = int(coarse_max)
    if level0_tile is not None:
        s.poisson_level0_tile = int(level0_tile)
    s.build_domain(params)
    return s


class Autotuner:
    """Project autotuner of a batch: the ion and electron cell-sort intervals
    and the deposit tile, i.e. the LDS window (width x height nodes) in which
    each work-group privatises its NGP and PPC counts.  None of them changes a
    result (every result is independent of particle order, and counts are
    integers), only gather locality, sorting cost and the number of global
    atomics, so candidates are tried on the run's own steps: each candidate
    runs `trial` timed steps, one parameter at a time (ion sort interval,
    electron sort interval, then window), and the fastest is kept until the
    next re-tune (`every` steps), since populations change as the preload
    decays."""

    def __init__(self, batch, ion_candidates=(1, 2, 5, 10, 20), electron_candidates=(1, 2, 4),
                 window_candidates=((64, 32), (128, 16), (256, 8), (512, 4), (32, 16)),
                 trial=6, every=500):
        self.batch = batch
        self.ion_candidates = tuple(ion_candidates)
        self.electron_candidates = tuple(electron_candidates)
        self.window_candidates = tuple(tuple(w) for w in window_candidates)
        self.trial = int(trial)
        self.every = int(every)
        self.history = []
        self._since = None

    def _set(self, si, se, win):
        for s in self.batch.sims:
            s.sort_interval, s.electron_sort_interval = int(si), int(se)
            s.count_window = win

    def _time(self, si, se, win):
        self._set(si, se, win)
        # a whole number of Poisson periods, so candidates see the same work
        steps = self.trial + (self.trial % 2)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        self.batch.advance(steps)
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / steps, steps

    def tune(self):
        done = 0
        s0 = self.batch.sims[0]
        best = [s0.sort_interval, s0.electron_sort_interval, tuple(s0.count_window)]
        best_t = math.inf
        for k, cands in enumerate((self.ion_candidates, self.electron_candidates, self.window_candidates)):
            for c in cands:
                trial = list(best)
                trial[k] = c
                t, n = self._time(*trial)
                done += n
                if t < best_t:
                    best, best_t = trial, t
        self._set(*best)
        self.history.append({"ion_sort": best[0], "electron_sort": best[1], "count_window": list(best[2]),
                             "ms_per_step": best_t * 1e3})
        self._since = 0
        return done

    def tune_cost(self):
        per = self.trial + (self.trial % 2)
        return per * (len(self.ion_candidates) + len(self.electron_candidates) + len(self.window_candidates))

    def advance(self, n_steps):
        left = int(n_steps)
        while left > 0:
            due = self._since is None or self._since >= self.every
            if due and left >= self.tune_cost():
                left -= self.tune()
                continue
            k = left if due else min(left, self.every - self._since)
            self.batch.advance(k)
            self._since = (self._since or 0) + k
            left -= k

