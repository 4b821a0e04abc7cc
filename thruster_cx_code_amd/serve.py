"""Served sweep: a request loop over the batched engine with continuous
batching.

Requests (operating points) arrive over time; each becomes a case that is
built on the device, admitted into the running batch when its estimated
device-memory footprint fits, advanced with every other active case (one
launch per step phase for all of them), and retired when it reaches steady
state or its step cap.  Cases admitted while others run overlap with them on
the device; a case's results do not depend on what else is in the batch.

Steady state follows the reference's sweep rule
(Python/benchmarks/benchmark_perveance.py:40-62): keep the last 50 non-NaN
divergence (theta95) values; after `min_steps` steps, stop when their
standard deviation is below 0.1 degree; the result is their mean.
"""
import math
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from .batch import BatchedSimulator
from .physics_engine import RING
from .sweep import build_case, two_grid_params


@dataclass
class Request:
    id: int
    params: dict
    arrival: float = 0.0  # seconds after the start of the run


@dataclass
class Result:
    id: int
    steps: int            # step at which the steady-state rule held (or the cap)
    converged: bool
    divergence: float     # mean theta95 of the final window (deg)
    window: list          # the final window of theta95 values
    arrival: float
    admitted: float
    finished: float
    error: str = ""       # set when the request was rejected (invalid case)

    @property
    def latency(self):
        return self.finished - self.arrival


def beam_request(rid, mesh, V_accel, current_mA, seed):
    """Perveance / backstreaming point: two-grid case whose upstream density
    follows from the beam current per aperture by inverting the injection
    current I = 0.61 q n0 v_B A (Python/physics_engine.py:1005-1015)."""
    nx, ny = mesh
    p = two_grid_params(nx, ny, V_accel=V_accel, seed=seed)
    q, m = 1.602e-19, 131.293 * 1.6605e-27
    v_b = math.sqrt(q * p["Te_up"] / m)
    Ly = 0.5 * p["grids"][0]["r"] + 0.30 * p["pitch_mm"]  # half_hole domain height (mm)
    area = Ly * 1e-3 * 0.005
    p["n0_plasma"] = current_mA * 1e-3 / (0.61 * q * v_b * area)
    return Request(rid, p)


def poisson_stream(n_requests, rate_hz, seed=0, meshes=((256, 96), (512, 192), (1024, 384))):
    """Poisson arrivals at rate_hz; mesh uniform over `meshes`, accel voltage
    ~ U[-400, -100] V, beam current ~ U[0.5, 3] mA per aperture."""
    r = np.random.default_rng(seed)
    t = 0.0
    out = []
    for k in range(n_requests):
        t += r.exponential(1.0 / rate_hz)
        mesh = meshes[int(r.integers(0, len(meshes)))]
        req = beam_request(k, mesh, r.uniform(-400, -100), r.uniform(0.5, 3.0), seed=k)
        req.arrival = t
        out.append(req)
    return out


def footprint_bytes(params):
    """Device-memory estimate of one case at admission: mesh arrays and the
    Poisson hierarchy (about 200 B per node) and the initial particle
    buffers (both species, ping-pong, at least 1e5 per species)."""
    nx, ny = params["mesh"]["nx"], params["mesh"]["ny"]
    pre = params.get("preload") or {}
    ni, ne = int(pre.get("ions", 0)), int(pre.get("electrons", 0))
    cap_i = max(100_000, int(ni * 1.25) + 65536)
    cap_e = max(100_000, int(ne * 1.25) + 65536) + cap_i
    return 200 * nx * ny + 2 * (24 * cap_i + 20 * cap_e)


class _Case:
    def __init__(self, req, sim, admitted):
        self.req = req
        self.sim = sim
        self.admitted = admitted
        self.start_it = sim.iteration
        self.checked_it = sim.iteration
        self.window = []
        self.stop_step = None
        self.converged = False


class SweepServer:
    """Continuous-batching server over BatchedSimulator.

    hbm_budget: bytes the active cases may occupy (admission control, by
    footprint_bytes against memory already allocated); max_active: slot
    limit; max_steps: step cap per request; min_steps: earliest steady-state
    stop (the reference uses 1500); check_every: steps between completion
    checks (the stop step itself is found exactly from the per-step records);
    level0_tile: level-0 tile edge of the cases' Poisson solves (16: the
    large tiles, faster in large batches); coarse_max: bound on the nodes of
    the coarsest multigrid level of the cases' solves (None: the solver's).
    """

    def __init__(self, hbm_budget=None, max_active=64, max_steps=4000, min_steps=1500,
                 check_every=25, tol_deg=0.1, window=50, chunk=5, level0_tile=16, coarse_max=None):
        if hbm_budget is None:
            free, total = torch.cuda.mem_get_info()
            hbm_budget = int(0.8 * total)
        self.hbm_budget = int(hbm_budget)
        self.max_active = int(max_active)
        self.max_steps = int(max_steps)
        self.min_steps = int(min_steps)
        self.check_every = int(check_every)
        self.tol_deg = float(tol_deg)
        self.window = int(window)
        self.chunk = int(chunk)
        self.level0_tile = level0_tile
        self.coarse_max = coarse_max
        self.batch = BatchedSimulator()
        self.log = []

    def _fits(self, req):
        used = torch.cuda.memory_allocated()
        return used + footprint_bytes(req.params) <= self.hbm_budget

    def _scan(self, case):
        """Feed the theta95 records since the last check into the case's
        window; return True when the case is done.  The records sit in pinned
        host memory: the caller synchronises the stream once before scanning
        the cases."""
        s = case.sim
        first, last = case.checked_it + 1, s.iteration
        if last < first:
            return False
        slots = [(k % RING) for k in range(first, last + 1)]
        vals = s._ring_theta_np[slots]
        for k, v in zip(range(first, last + 1), vals):
            steps = k - case.start_it
            if not math.isnan(v):
                case.window.append(float(v))
                if len(case.window) > self.window:
                    case.window.pop(0)
            if steps >= self.min_steps and len(case.window) == self.window and \
                    float(np.std(case.window)) < self.tol_deg:
                case.stop_step, case.converged = steps, True
                break
            if steps >= self.max_steps:
                case.stop_step, case.converged = steps, False
                break
        case.checked_it = last
        return case.stop_step is not None

    def run(self, requests, realtime=True, progress=None):
        """Serve `requests` (sorted by arrival).  realtime=True admits a
        request once its arrival time has passed on the wall clock; False
        admits every request immediately (for equality tests).  progress, if
        given, is called after every completion check with (seconds since
        start, active cases, requests admitted or rejected, results).
        Returns the results in completion order."""
        queue = sorted(requests, key=lambda r: r.arrival)
        active, results = [], []
        t0 = time.perf_counter()
        steps_since_check = 0
        qi = 0
        while qi < len(queue) or active:
            now = time.perf_counter() - t0
            # admission: arrived requests, while slots and device memory allow
            while qi < len(queue) and (not realtime or queue[qi].arrival <= now) and \
                    len(active) < self.max_active and (not active or self._fits(queue[qi])):
                req = queue[qi]
                qi += 1
                try:
                    sim = build_case(req.params, coarse_max=self.coarse_max, level0_tile=self.level0_tile)
                except ValueError as exc:  # the reference's build checks reject it
                    t = time.perf_counter() - t0
                    results.append(Result(req.id, 0, False, float("nan"), [], req.arrival, t, t, str(exc)))
                    continue
                self.batch.add(sim, req.params)
                active.append(_Case(req, sim, time.perf_counter() - t0))
            if not active:
                if qi >= len(queue):
                    break
                wait = queue[qi].arrival - (time.perf_counter() - t0)
                if wait > 0:
                    time.sleep(min(wait, 0.05))
                continue
            self.batch.advance(self.chunk)
            steps_since_check += self.chunk
            if steps_since_check >= self.check_every:
                steps_since_check = 0
                torch.cuda.current_stream().synchronize()
                done = [c for c in active if self._scan(c)]
                now = time.perf_counter() - t0
                for c in done:
                    self.batch.remove(c.sim)
                    active.remove(c)
                    w = list(c.window)
                    results.append(Result(c.req.id, c.stop_step, c.converged,
                                          float(np.mean(w)) if w else float("nan"), w,
                                          c.req.arrival, c.admitted, now))
                    del c.sim
                self.log.append((now, len(active), qi))
                if progress is not None:
                    progress(now, len(active), qi, len(results))
        return results


def latency_summary(results, wall_s):
    """p50/p99 latency (arrival to result) of the served requests; rejected
    requests (invalid cases) are counted separately."""
    ok = [r for r in results if not r.error]
    lat = np.array([r.latency for r in ok]) if ok else np.array([float("nan")])
    return {"served": len(ok), "rejected": len(results) - len(ok),
            "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mean_s": float(lat.mean()), "throughput_req_per_s": len(ok) / wall_s,
            "converged": int(sum(r.converged for r in ok))}
