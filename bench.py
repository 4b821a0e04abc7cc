"""Benchmark: particle-steps/s of the PY-BEMCS time step (Poisson included) on
one MI355X.

Workload (default "c2"): two-grid Xe+ extraction (screen +1000 V, 0.38 mm,
r 0.95 mm; accel -200 V, 0.5 mm, r 0.55 mm; 0.6 mm gap), 1024 x 384 nodes,
8M macro-ions + 8M macro-electrons (synthetic, seed 0), Ti 0.2 eV, electrons
Maxwellian 5 eV with m_e = M_Xe/200, uniform neutral Xe 1e18 m^-3 at 300 K,
log-linear CEX cross-section, dt 5e-10 s, MG-PCG Poisson to 1e-9 relative
residual, 5 Picard iterations every 2nd step.  The population is held at
8M + 8M: after each step's beam injection and neutraliser the device tops
both species up with preload-distributed particles (the reference's own
injection at this mesh holds about 1M of each: the macro-weight rule caps the
ion source and electron space charge expels electrons).  Ions are sorted by
cell every 20 steps, electrons at the start of every step.

Prints one JSON line.  The timed region is `--steps` steps (default 2000)
after `--warmup` steps (default 50), bracketed by device synchronisation;
`value` counts the particles pushed in each timed step (after injection)
divided by the wall time; the pushed populations are reported in `config`.
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
WINDOW = 4000  # timed steps per synchronised window (below the per-step record ring)


def workload_params(name, ions=None, electrons=None):
    if name == "c5":
        return c5_params(ions, electrons)
    mesh = {"c2": (1024, 384, 8_000_000), "c1": (256, 96, 100_000)}[name]
    nx, ny, npre = mesh
    return {
        "grids": [{"V": 1000.0, "t": 0.38, "gap": 0.6, "r": 0.95, "cham": 0.0},
                  {"V": -200.0, "t": 0.5, "gap": 1.0, "r": 0.55, "cham": 0.0}],
        "pitch_mm": 6.1447, "geometry": "half_hole", "V_plasma_offset": 20.0,
        "n0_plasma": 1e17, "Te_up": 3.0, "Ti": 0.2, "Te": 5.0, "Tn": 300.0,
        "n0": 1e18, "neutral_model": "uniform", "cx_model": "loglinear",
        "neut_rate": 30, "sim_mode": "Trajectories", "mesh": {"nx": nx, "ny": ny},
        "preload": {"ions": npre if ions is None else ions,
                    "electrons": npre if electrons is None else electrons,
                    "hold": name == "c2"},
        "seed": 0,
    }


def c5_params(ions=None, electrons=None):
    """Three-grid optics with sputter erosion, thermal conduction and bow, and
    RF co-extraction: screen +1200 V / accel -250 V / decel 0 V, each 0.5 mm
    thick with 0.6 mm gaps (apertures 0.95 / 0.55 / 0.8 mm), 50 V at 13.56 MHz
    on the screen, 2048 x 768 nodes, 32M + 32M preloaded."""
    n = 32_000_000
    return {
        "grids": [{"V": 1200.0, "t": 0.5, "gap": 0.6, "r": 0.95, "cham": 0.0},
                  {"V": -250.0, "t": 0.5, "gap": 0.6, "r": 0.55, "cham": 0.0},
                  {"V": 0.0, "t": 0.5, "gap": 0.6, "r": 0.8, "cham": 0.0}],
        "pitch_mm": 6.1447, "geometry": "half_hole", "V_plasma_offset": 20.0,
        "n0_plasma": 1e17, "Te_up": 3.0, "Ti": 0.2, "Te": 5.0, "Tn": 300.0,
        "n0": 1e18, "neutral_model": "uniform", "cx_model": "loglinear",
        "neut_rate": 30, "sim_mode": "Both", "Thresh": 1e5,
        "rf_enable": True, "rf_grid_idx": 0, "rf_amp": 50.0, "rf_freq": 13.56,
        "mesh": {"nx": 2048, "ny": 768},
        "preload": {"ions": n if ions is None else ions, "electrons": n if electrons is None else electrons},
        "seed": 0,
    }


# ------------------------------------------------------------- CPU baseline
def cpu_worker(workload, ions, electrons, max_seconds, max_steps):
    """Time the NumPy oracle (reference dtype semantics) on one core."""
    from oracle.sim import OracleSimulator
    p = workload_params(workload, ions, electrons)
    o = OracleSimulator(mode="ref32", mcc_math="ref32")
    o.m_e = o.m_ion / 200.0
    o.build_domain(p)
    pushed = 0
    steps = 0
    t0 = time.perf_counter()
    last = t0
    while steps < max_steps and time.perf_counter() - t0 < max_seconds:
        o.step(p)
        pushed += o.pushed
        steps += 1
        if time.perf_counter() - last >= 30.0:
            last = time.perf_counter()
            print(f"[cpu] {steps} steps, {last - t0:.0f} s", file=sys.stderr, flush=True)
    wall = time.perf_counter() - t0
    print(json.dumps({"value": pushed / wall, "steps": steps, "wall_s": wall, "pushed": pushed}))


def cpu_worker_composed(workload, n_sub_i, n_sub_e):
    """Per-step time of the NumPy oracle (reference dtype semantics) at the
    GPU line's own population, composed from measured pieces on one core:
    one full odd step with every species pushed once (injection, hold, NGP
    deposit, substep counts, diagnostics, purge, CEX, PPC), one push substep
    per species, and one 5-iteration SuperLU Picard call (every 2nd step):
        t_step = t_odd + t_picard / 2 + (n_sub_i - 1) t_push_i + (n_sub_e - 1) t_push_e
    with the mean substep counts the GPU measured over its timed steps."""
    from oracle import kernels as K
    from oracle.sim import OracleSimulator

    class OneSubstep(OracleSimulator):
        """Runs the push of each species once per step (timing only)."""
        def step(self, params):
            real = K.push_boris
            calls = {}

            def once(x, *a, **kw):
                key = id(x)
                if key not in calls:
                    calls[key] = True
                    real(x, *a, **kw)
            import oracle.sim as osim
            osim.push_boris = once
            try:
                return super().step(params)
            finally:
                osim.push_boris = real

    p = workload_params(workload)
    o = OneSubstep(mode="ref32", mcc_math="ref32")
    o.m_e = o.m_ion / 200.0
    t0 = time.perf_counter()
    o.build_domain(p)
    t_build = time.perf_counter() - t0
    t0 = time.perf_counter()
    o.step(p)                       # iteration 1: odd, no Poisson call
    t_odd = time.perf_counter() - t0
    pushed = o.pushed
    t0 = time.perf_counter()
    o.recalc_poisson(5, p)
    t_picard = time.perf_counter() - t0
    qm_i, qm_e = o.q_ion / o.m_ion, -o.q / o.m_e
    t0 = time.perf_counter()
    K.push_boris(o.p_x, o.p_y, o.p_vx, o.p_vy, o.p_vz, o.Ex, o.Ey, o.Bx, o.By, o.Bz, o.num_p, o.dx, o.dy,
                 o.nx, o.ny, o.dt / max(n_sub_i, 1), qm_i)
    t_push_i = time.perf_counter() - t0
    t0 = time.perf_counter()
    K.push_boris(o.e_x, o.e_y, o.e_vx, o.e_vy, o.e_vz, o.Ex, o.Ey, o.Bx, o.By, o.Bz, o.num_e, o.dx, o.dy,
                 o.nx, o.ny, o.dt / max(n_sub_e, 1), qm_e)
    t_push_e = time.perf_counter() - t0
    t_step = t_odd + 0.5 * t_picard + (n_sub_i - 1) * t_push_i + (n_sub_e - 1) * t_push_e
    print(json.dumps({"value": pushed / t_step, "pushed": pushed, "t_step_s": t_step, "t_odd_s": t_odd,
                      "t_picard_s": t_picard, "t_push_i_s": t_push_i, "t_push_e_s": t_push_e,
                      "n_sub": [n_sub_i, n_sub_e], "build_s": t_build}))


def cpu_worker_sweep(path):
    """CPU time of a sweep's cases on the NumPy oracle (reference dtype
    semantics, SciPy SuperLU), one core, composed per case from measured
    pieces and the populations the GPU run recorded:
        t_case = t_build(m) + c_pre N_pre + steps t0(m) + a S + b P + (steps / 2) t_picard(m)
    where, for the case's mesh m, t_build(m) and t0(m) (one odd step with
    only the step's own injections) and t_picard(m) (a 5-iteration Picard
    call) are measured on that mesh; b (one push substep per particle) and a
    (the rest of a step per particle: deposit, masks, purge, CEX, PPC) are
    measured once with 2e5 + 2e5 particles, c_pre from the same build; S is
    the sum over steps of the pushed particles and P of the pushed particles
    times their substep counts, both from the GPU's per-step records."""
    from oracle import kernels as K
    from oracle.sim import OracleSimulator
    spec = json.load(open(path))
    cases, stats = spec["cases"], spec["stats"]

    class OneSubstep(OracleSimulator):
        """Runs the push of each species once per step (timing only)."""
        def step(self, params):
            real = K.push_boris
            calls = {}

            def once(x, *a, **kw):
                if id(x) not in calls:
                    calls[id(x)] = True
                    real(x, *a, **kw)
            import oracle.sim as osim
            osim.push_boris = once
            try:
                return super().step(params)
            finally:
                osim.push_boris = real

    def build(params):
        o = OneSubstep(mode="ref32", mcc_math="ref32")
        o.m_e = o.m_ion / 200.0
        t0 = time.perf_counter()
        o.build_domain(params)
        return o, time.perf_counter() - t0

    def empty(params):
        q = json.loads(json.dumps(params))
        q["preload"] = {"ions": 0, "electrons": 0}
        return q

    per_mesh = {}
    for p in cases:
        key = (p["mesh"]["nx"], p["mesh"]["ny"])
        if key in per_mesh:
            continue
        q = empty(p)
        o, t_build = build(q)
        t0 = time.perf_counter()
        o.step(q)
        t_odd = time.perf_counter() - t0
        t0 = time.perf_counter()
        o.recalc_poisson(5, q)
        per_mesh[key] = {"t_build": t_build, "t0": t_odd, "t_picard": time.perf_counter() - t0}
    ref = sorted(per_mesh)[len(per_mesh) // 2]
    q = empty(next(p for p in cases if (p["mesh"]["nx"], p["mesh"]["ny"]) == ref))
    nc = 200_000
    q["preload"] = {"ions": nc, "electrons": nc}
    o, t_build = build(q)
    t0 = time.perf_counter()
    o.step(q)
    t_odd = time.perf_counter() - t0
    pushed = o.pushed
    t0 = time.perf_counter()
    K.push_boris(o.e_x, o.e_y, o.e_vx, o.e_vy, o.e_vz, o.Ex, o.Ey, o.Bx, o.By, o.Bz, o.num_e, o.dx, o.dy,
                 o.nx, o.ny, o.dt, -o.q / o.m_e)
    K.push_boris(o.p_x, o.p_y, o.p_vx, o.p_vy, o.p_vz, o.Ex, o.Ey, o.Bx, o.By, o.Bz, o.num_p, o.dx, o.dy,
                 o.nx, o.ny, o.dt, o.q_ion / o.m_ion)
    b = (time.perf_counter() - t0) / max(o.num_e + o.num_p, 1)
    a = max((t_odd - per_mesh[ref]["t0"] - b * pushed) / max(pushed, 1), 0.0)
    c_pre = max((t_build - per_mesh[ref]["t_build"]) / (2 * nc), 0.0)
    total = 0.0
    per_case = []
    for p, st in zip(cases, stats):
        m = per_mesh[(p["mesh"]["nx"], p["mesh"]["ny"])]
        n_pre = p["preload"]["ions"] + p["preload"]["electrons"]
        t = (m["t_build"] + c_pre * n_pre + st["steps"] * m["t0"] + a * st["sum_n"] + b * st["sum_n_sub"]
             + (st["steps"] // 2) * m["t_picard"])
        per_case.append(t)
        total += t
    print(json.dumps({"total_s": total, "per_case_s": per_case, "a_s": a, "b_s": b, "c_pre_s": c_pre,
                      "per_mesh": {f"{k[0]}x{k[1]}": v for k, v in per_mesh.items()}}))


def cpu_sweep_baseline(cases, stats):
    """Cases/hr of the sweep on the host's cores: the one-core composed CPU
    time of every case (cpu_worker_sweep), spread over P = the cores this
    process may use (at most 16, the host share of one GPU) running one case
    each, as the reference's sweep runs one process per case.  P processes
    at the one-core rate is an upper bound for the CPU."""
    import tempfile
    with tempfile.NamedTemporaryFile("w", suffix=".json", delete=False) as f:
        json.dump({"cases": cases, "stats": stats}, f)
        path = f.name
    try:
        rec = _run_cpu(["--cpu-sweep", path], 3600)
    finally:
        os.unlink(path)
    P = min(len(os.sched_getaffinity(0)), 16)
    value = len(cases) * 3600.0 / (rec["total_s"] / P)
    return {
        "value": value, "unit": "cases/hr", "cores": P, "kind": "port",
        "sample": (f"NumPy oracle (reference dtype semantics, SciPy SuperLU) on {_cpu_model()}: per-case "
                   f"one-core time composed from measured pieces (per mesh: build, an odd step with only "
                   f"its injections, a 5-iteration Picard call; per particle: {rec['b_s'] * 1e9:.0f} ns "
                   f"per push substep and {rec['a_s'] * 1e9:.0f} ns per step) and the populations and "
                   f"substep counts the GPU recorded at every step; {rec['total_s'] / 3600:.1f} core-hours "
                   f"for all cases, spread over {P} processes"),
        "breakdown": rec,
    }


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _run_cpu(args, timeout):
    env = dict(os.environ)
    for k in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS", "NUMEXPR_NUM_THREADS"):
        env[k] = "1"
    env["HIP_VISIBLE_DEVICES"] = ""
    cmd = [sys.executable, os.path.abspath(__file__)] + [str(a) for a in args]
    # progress lines of long runs pass through on stderr
    out = subprocess.run(cmd, env=env, stdout=subprocess.PIPE, text=True, timeout=timeout)
    if out.returncode != 0:
        raise RuntimeError(f"cpu baseline failed (exit {out.returncode})")
    return json.loads(out.stdout.strip().splitlines()[-1])


def cpu_baseline(workload, ions, electrons, max_seconds=20.0, max_steps=200):
    rec = _run_cpu(["--cpu-worker", workload, ions, electrons, max_seconds, max_steps],
                   min(max_seconds * 20 + 600, 7200))
    return {
        "value": rec["value"], "unit": "particle-steps/s", "cores": 1, "kind": "port",
        "sample": (f"NumPy oracle (reference dtype semantics, SciPy SuperLU) on the same "
                   f"{workload} mesh and physics with {ions}+{electrons} preloaded particles, "
                   f"{rec['steps']} steps in {rec['wall_s']:.1f} s on 1 core of {_cpu_model()}"),
    }


def cpu_baseline_composed(workload, n_sub_i, n_sub_e):
    rec = _run_cpu(["--cpu-composed", workload, n_sub_i, n_sub_e], 1800)
    return {
        "value": rec["value"], "unit": "particle-steps/s", "cores": 1, "kind": "port",
        "sample": (f"NumPy oracle (reference dtype semantics, SciPy SuperLU) at the GPU line's population "
                   f"({rec['pushed']} particles pushed per step) on 1 core of {_cpu_model()}; per-step time "
                   f"{rec['t_step_s']:.1f} s composed from measured pieces: one odd step with one push per "
                   f"species {rec['t_odd_s']:.2f} s + half a 5-iteration Picard call "
                   f"({rec['t_picard_s']:.2f} s) + (n_sub - 1) further push substeps per species at "
                   f"{rec['t_push_i_s']:.2f} s (ions) and {rec['t_push_e_s']:.2f} s (electrons), with the "
                   f"GPU's mean n_sub {n_sub_i}/{n_sub_e} over the timed steps"),
        "breakdown": rec,
    }


# ---------------------------------------------------------------- GPU bench
# Algorithmic bytes per unit (counting convention: state read + written once
# per step, field read once, counts written once; sorting, ping-pong,
# iterations and scratch are method costs and are not counted).
BYTES_PER_ION = 45       # 5 x f32 state read + write (40), id u32 (4), CEX flag u8 (1)
BYTES_PER_ELECTRON = 40  # 5 x f32 state read + write
BYTES_PER_NODE_PUSH = 12  # (Ex, Ey) f32 read + u32 deposit counts written, per particle kernel
BYTES_PER_NODE_STEP = 84  # whole step: counts 8 + field 8 + Poisson I/O 60 + E field 8


def algorithmic_bytes(name, n_ions, n_elec, n_nodes):
    """HBM bytes one launch must move at minimum (see the constants above)."""
    if name == "ion_step":
        return BYTES_PER_ION * n_ions + BYTES_PER_NODE_PUSH * n_nodes
    if name == "electron_step":
        return BYTES_PER_ELECTRON * n_elec + BYTES_PER_NODE_PUSH * n_nodes
    return 0


# MI355X vector issue: 256 CUs x 4 SIMDs at 2.4 GHz; a SIMD has 16 lanes, so
# one wave64 VALU instruction (float64 or 32-bit, unpacked) occupies it for 4
# clocks (the published FP64 vector peak, 78.6 TFLOP/s, is that rate with
# FMAs; the FP32 peak of twice that needs packed instructions)
SIMD_CLOCKS_PER_S = 1024 * 2.4e9
CLOCKS_PER_WAVE_VALU = 4


def read_profile_value(name, key):
    """A figure from a committed rocprofv3 summary under profiles/ (a line
    `key: value`); None when the summary is absent."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        with open(path) as fh:
            for ln in fh:
                if ln.startswith(key + ":"):
                    return float(ln.split(":", 1)[1].split()[0])
    except OSError:
        pass
    return None


def read_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` measured with rocprofv3 --pmc
    (FETCH_SIZE and WRITE_SIZE, scaled by a copy-kernel calibration) on the
    driver's bench command; None when no summary is committed."""
    return read_profile_value(f"traffic_{kernel}_{workload}.txt", "traffic_bytes_per_launch")


def _progress(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def run(args):
    import numpy as np
    import torch
    from thruster_cx_code_amd import DigitalTwinSimulator

    p = workload_params(args.workload)
    # c5 keeps millions of particles for hundreds of steps: sorting more often pays
    default_sort = (5, 1) if args.workload == "c5" else (20, 20)
    if args.ion_sort_interval is None:
        args.ion_sort_interval = default_sort[0]
    if args.electron_sort_interval is None:
        args.electron_sort_interval = default_sort[1]
    sim = DigitalTwinSimulator(sort_interval=args.ion_sort_interval,
                               electron_sort_interval=args.electron_sort_interval)
    sim.m_e = sim.m_ion / 200.0
    t_build = time.perf_counter()
    sim.build_domain(p)
    torch.cuda.synchronize()
    t_build = time.perf_counter() - t_build
    _progress(f"built {args.workload} in {t_build:.2f} s")
    sim.advance(p, args.warmup)
    torch.cuda.synchronize()
    _progress(f"warmup {args.warmup} steps done")
    first = sim.iteration + 1
    from thruster_cx_code_amd.physics_engine import RING
    # the timed steps run in windows of at most WINDOW steps (the per-step
    # records live in a ring of RING slots); each window is bracketed by
    # synchronisation and timed with events, and its records and kernel
    # timings are read back between windows, outside the timed time
    wall_ms = 0.0
    nin_l, nsub_l, remesh_l, durs = [], [], [], {}
    done = 0
    while done < args.steps:
        k = min(WINDOW, args.steps - done)
        sim.kernel_timing = {}
        start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        w0 = sim.iteration + 1
        start.record()
        sim.advance(p, k)
        end.record()
        torch.cuda.synchronize()
        wall_ms += start.elapsed_time(end)
        slots = [(w0 + j) % RING for j in range(k)]
        nin_l.append(sim.ring_nin.view(RING, 2).cpu().numpy()[slots].astype(np.int64))
        nsub_l.append(sim.ring_nsub.view(RING, 2).cpu().numpy()[slots].astype(np.int64))
        remesh_l.append(sim.ring_remeshed[slots].copy())
        for name, evs in sim.kernel_timing.items():
            durs.setdefault(name, []).extend(a.elapsed_time(b) for a, b in evs)
        done += k
        _progress(f"timed {done} of {args.steps} steps: {wall_ms:.1f} ms")
    nin = np.concatenate(nin_l)
    nsub = np.concatenate(nsub_l)
    pushed = int(nin.sum())
    value = pushed / (wall_ms * 1e-3)
    n_nodes = sim.nx * sim.ny
    kstats = {}
    for name, d in durs.items():
        kstats[name] = {"launches": len(d), "total_ms": float(sum(d)), "avg_ms": float(np.mean(d))}
    for name, col in (("ion_step", 0), ("electron_step", 1)):
        if name in kstats:
            subs = int((nin[:, col] * nsub[:, col]).sum())
            kstats[name]["particle_substeps"] = subs
            kstats[name]["particle_substeps_per_s"] = subs / (kstats[name]["total_ms"] * 1e-3)
            kstats[name]["n_sub_range"] = [int(nsub[:, col].min()), int(nsub[:, col].max())]
            kstats[name]["particles_range"] = [int(nin[:, col].min()), int(nin[:, col].max())]
            byts = sum(algorithmic_bytes(name, int(r[0]), int(r[1]), n_nodes) for r in nin)
            kstats[name]["algorithmic_bytes_per_launch"] = byts / kstats[name]["launches"]
    n_calls = sum(kstats[k]["launches"] for k in ("poisson_picard", "poisson_rf") if k in kstats)
    if n_calls:
        # every Picard call of the timed steps (RF calls interleave with the
        # every-2nd-step calls); geometry rebuilds add their own calls
        its = np.asarray(sim.poisson_iter_log[-n_calls:])
        kstats["pcg_iterations_per_picard_call"] = {
            "calls": n_calls, "min": int(its.min()), "median": float(np.median(its)),
            "max": int(its.max()), "total": int(its.sum())}
    remeshes = int(np.concatenate(remesh_l).sum())
    # dominant single kernel: the fused particle steps (one launch each); the
    # Poisson call is a sequence of ~30 launches per PCG iteration
    dom = max(("ion_step", "electron_step"), key=lambda k: kstats[k]["total_ms"])
    per_launch = kstats[dom]["algorithmic_bytes_per_launch"]
    achieved = per_launch / (kstats[dom]["avg_ms"] * 1e-3) / 1e9
    traffic = read_traffic(dom, args.workload)
    # second bound of the dominant kernel: its substep loop issues a fixed
    # number of VALU instructions per particle-substep (SQ_INSTS_VALU per
    # substep, from the committed counter summary of the same command)
    vfile = f"valu_{dom}_{args.workload}.txt"
    valu_per_sub = read_profile_value(vfile, "valu_instructions_per_substep")
    valu_busy = read_profile_value(vfile, "valu_busy")
    ta_busy = read_profile_value(vfile, "ta_busy")
    sub_rate = kstats[dom].get("particle_substeps_per_s")
    substep_roofline = None
    if valu_per_sub and sub_rate:
        # wave64 VALU instructions per 64 particle-substeps, each 4 SIMD clocks
        peak_sub = SIMD_CLOCKS_PER_S / CLOCKS_PER_WAVE_VALU / valu_per_sub * 64
        substep_roofline = {"bound": "valu", "kernel": dom, "achieved": sub_rate, "peak": peak_sub,
                            "unit": "particle-substeps/s", "frac": sub_rate / peak_sub,
                            "valu_instructions_per_substep": valu_per_sub, "valu_busy_counters": valu_busy,
                            "texture_addresser_busy_counters": ta_busy,
                            "peak_basis": "1024 SIMDs x 2.4 GHz / 4 clocks per wave64 VALU instruction / "
                                          "VALU instructions per wave-substep x 64 particles"}
    step_bytes = int(sum(BYTES_PER_ION * r[0] + BYTES_PER_ELECTRON * r[1] + BYTES_PER_NODE_STEP * n_nodes
                         for r in nin))
    step_achieved = step_bytes / (wall_ms * 1e-3) / 1e9
    rec = sim.record()
    labels = {
        "c2": ("particle-steps/s incl. Poisson at 1024x384",
               "c2: two-grid Xe+ extraction, 1024x384 nodes, 8M ions + 8M electrons held (preload "
               "topped up to 8M + 8M each step after injection, synthetic), "
               f"steps {first}-{first + args.steps - 1}, dt 5e-10 s, "
               f"MG-PCG to {sim.poisson_rtol:g}, 5 Picard iterations every 2nd step"),
        "c1": ("particle-steps/s incl. Poisson at 256x96", "c1: 256x96 nodes, 1e5 + 1e5 preloaded"),
        "c5": ("particle-steps/s incl. Poisson, thermal, erosion and RF at 2048x768",
               "c5: three-grid (+1200/-250/0 V) with sputter erosion, thermal conduction and bow, "
               "RF 50 V at 13.56 MHz on the screen (2 Picard iterations every step plus 5 every 2nd), "
               f"2048x768 nodes, 32M ions + 32M electrons preloaded, steps {first}-{first + args.steps - 1}"),
    }
    out = {
        "metric": labels[args.workload][0],
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall_ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 state, f64 gather/kick, f64 CG + f32 multigrid",
        "data": "synthetic (seeded preload, Philox streams)",
        "config": {"workload": labels[args.workload][1],
                   "electron_sort": {"presort": sim.electron_presort, "chunk_substeps": sim.electron_chunk},
                   "nodes": n_nodes, "particles_pushed": pushed, "build_s": t_build,
                   "ions_range": kstats["ion_step"]["particles_range"],
                   "electrons_range": kstats["electron_step"]["particles_range"],
                   "n_sub_last": [rec["n_sub_ion"], rec["n_sub_e"]],
                   "num_p_last": rec["num_p"], "num_e_last": rec["num_e"],
                   "sort_interval_steps": {"ions": args.ion_sort_interval,
                                           "electrons": args.electron_sort_interval},
                   "remeshes": remeshes},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic},
        "substep_roofline": substep_roofline,
        "step_roofline": {"bytes_per_step": step_bytes / args.steps, "achieved": step_achieved,
                          "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": step_achieved / HBM_PEAK_GBS},
        "kernels": kstats,
    }
    if not args.no_cpu_baseline:
        _progress("cpu baseline")
        if args.workload == "c1":
            if args.cpu_steps:
                # the whole run: every step of the CPU restatement, no time bound
                out["cpu_baseline"] = cpu_baseline(args.workload, 100_000, 100_000, 1e9, args.cpu_steps)
            else:
                out["cpu_baseline"] = cpu_baseline(args.workload, 100_000, 100_000, args.cpu_seconds)
        else:
            ns_i = int(round(float(np.mean(nsub[:, 0]))))
            ns_e = int(round(float(np.mean(nsub[:, 1]))))
            out["cpu_baseline"] = cpu_baseline_composed(args.workload, ns_i, ns_e)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


def poisson_bench(calls=10):
    """Time recalc_poisson (5 Picard iterations + E field) alone on the c2
    mesh with the c2 preload's charge (for profiling the solver)."""
    import numpy as np
    import torch
    from thruster_cx_code_amd import DigitalTwinSimulator
    p = workload_params("c2")
    sim = DigitalTwinSimulator()
    sim.m_e = sim.m_ion / 200.0
    sim.build_domain(p)
    sim.recalc_poisson(5, p)
    torch.cuda.synchronize()
    times = []
    for _ in range(calls):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        sim.recalc_poisson(5, p)
        b.record()
        torch.cuda.synchronize()
        times.append(a.elapsed_time(b))
    print(json.dumps({"poisson_ms": times, "median_ms": float(np.median(times)),
                      "pcg_iterations": sim.poisson_iter_log[-calls:],
                      "levels": sim.lib.bemcs_poisson_levels(sim._poisson)}))


def sweep_bench(args):
    """Sweep throughput (cases/hr): `--sweep-cases` uneven two-grid cases
    (thruster_cx_code_amd.sweep.uneven_sweep), `--sweep-steps` steps each,
    advanced as one batch with the autotuner (sort intervals and deposit tile), against the same
    GPU running the cases one at a time (timed on every k-th case and scaled to
    all cases).  Build (geometry, preload, initial Poisson) is included.  The
    CPU baseline (cpu_sweep_baseline) composes each case's one-core NumPy
    time from the populations this run recorded."""
    import gc
    import numpy as np
    import torch
    from thruster_cx_code_amd.batch import BatchedSimulator
    from thruster_cx_code_amd.sweep import Autotuner, build_case, uneven_sweep
    cases = uneven_sweep(args.sweep_cases, seed=0)
    n = len(cases)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sims = [build_case(p, coarse_max=args.sweep_coarse or None, level0_tile=args.sweep_tile) for p in cases]
    torch.cuda.synchronize()
    t_build = time.perf_counter() - t0
    _progress(f"built {n} cases in {t_build:.1f} s")
    batch = BatchedSimulator(sims, cases)
    tuner = Autotuner(batch)
    t0 = time.perf_counter()
    done = 0
    while done < args.sweep_steps:
        k = min(100, args.sweep_steps - done)
        tuner.advance(k)
        done += k
        torch.cuda.synchronize()
        _progress(f"batch: {done} steps, {time.perf_counter() - t0:.1f} s")
    t_batch = time.perf_counter() - t0
    captures = batch.graph_captures
    final = [(s.record()["num_p"], s.record()["num_e"]) for s in sims]
    # per-case populations and substep counts of every step, for the CPU model
    from thruster_cx_code_amd.physics_engine import RING
    stats = []
    for s in sims:
        slots = [k % RING for k in range(1, s.iteration + 1)]
        nin = s.ring_nin.view(RING, 2).cpu().numpy()[slots].astype(np.int64)
        nsub = s.ring_nsub.view(RING, 2).cpu().numpy()[slots].astype(np.int64)
        stats.append({"steps": int(s.iteration), "sum_n": int(nin.sum()), "sum_n_sub": int((nin * nsub).sum())})
    tuning = tuner.history
    del tuner, batch, sims
    gc.collect()
    torch.cuda.empty_cache()
    stride = max(1, n // max(1, args.sweep_seq_cases))
    sample = list(range(0, n, stride))
    t_seq_sample = 0.0
    for c in sample:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        s = build_case(cases[c])
        s.advance(cases[c], args.sweep_steps)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        t_seq_sample += dt
        _progress(f"alone: case {c} ({cases[c]['mesh']['nx']}x{cases[c]['mesh']['ny']}) {dt:.1f} s")
        del s
        gc.collect()
        torch.cuda.empty_cache()
    t_seq = t_seq_sample * n / len(sample)
    batched_hr = n / (t_build + t_batch) * 3600.0
    seq_hr = n / t_seq * 3600.0
    out = {
        "metric": "sweep cases/hr (uneven two-grid cases, 256x96..1024x384)",
        "value": batched_hr, "unit": "cases/hr", "n_gpus": 1, "higher_is_better": True,
        "vs_baseline": None, "data": "synthetic (seeded preload, Philox streams)",
        "config": {"workload": f"{n} cases, {args.sweep_steps} steps each: nx ~ U{{256..1024 step 64}}, "
                               "ny = 3nx/8, Va ~ U[-400,-100] V, r_accel ~ U[0.45,0.75] mm, "
                               "Vs ~ U[800,1600] V, particles per species log-uniform [2e5, 4e6]",
                   "cases": n, "steps": args.sweep_steps},
        "batched": {"cases_per_hr": batched_hr, "build_s": t_build, "steps_s": t_batch,
                    "autotuning": tuning, "graph_captures": captures},
        "one_at_a_time": {"cases_per_hr": seq_hr, "timed_cases": sample, "timed_s": t_seq_sample,
                          "scaled_s": t_seq},
        "speedup": batched_hr / seq_hr,
        "final_counts_first_cases": final[:4],
    }
    if not args.no_cpu_baseline:
        _progress("cpu baseline (composed per case)")
        out["cpu_baseline"] = cpu_sweep_baseline(cases, stats)
    print(json.dumps(out))


def serve_bench(args):
    """Served sweep (continuous batching): `--serve-requests` perveance /
    backstreaming requests with Poisson arrivals at `--serve-rate` per second,
    mesh from {256x96, 512x192, 1024x384}, accel voltage ~ U[-400,-100] V,
    beam current ~ U[0.5, 3] mA per aperture, each run to the steady-state rule
    or `--serve-max-steps`; p50/p99 latency (arrival to result) and
    throughput."""
    import torch
    from thruster_cx_code_amd.serve import SweepServer, latency_summary, poisson_stream
    reqs = poisson_stream(args.serve_requests, args.serve_rate, seed=0)
    # the reference sweep's rule: at least 1500 steps, then stop when the last
    # 50 theta95 values have a standard deviation below 0.1 degree
    # (Python/benchmarks/benchmark_perveance.py:56-60), capped at max_steps
    min_steps = min(1500, args.serve_max_steps)
    srv = SweepServer(max_steps=args.serve_max_steps, min_steps=min_steps, max_active=args.serve_slots,
                      coarse_max=args.serve_coarse)
    t0 = time.perf_counter()
    last = [0.0]

    def progress(now, active, admitted, done):
        if now - last[0] >= 20.0:
            last[0] = now
            _progress(f"serve: {now:.0f} s, {done} done, {active} active, {admitted} admitted")
    res = srv.run(reqs, realtime=True, progress=progress)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    summ = latency_summary(res, wall)
    _progress(f"served {len(res)} requests in {wall:.1f} s")
    cap = capacity(res, wall, args.serve_rate)
    alone = None
    if args.serve_seq > 0:
        alone = serve_alone(reqs, res, args.serve_seq)
        alone["served_over_alone"] = cap["case_steps_per_s"] / alone["case_steps_per_s"]
    print(json.dumps({
        "metric": "served sweep p50/p99 latency (continuous batching)",
        "value": summ["p50_s"], "unit": "s", "higher_is_better": False, "n_gpus": 1,
        "vs_baseline": None, "data": "synthetic (Poisson arrivals, seeded cases)",
        "config": {"workload": f"{args.serve_requests} requests at {args.serve_rate}/s, meshes 256x96/512x192/"
                               f"1024x384, Va ~ U[-400,-100] V, I ~ U[0.5,3] mA, steady-state rule after "
                               f"{min_steps} steps, cap {args.serve_max_steps} steps",
                   "slots": args.serve_slots},
        "latency": summ, "wall_s": wall,
        "max_active": max((a for _, a, _ in srv.log), default=0),
        "capacity": cap,
        "one_at_a_time": alone,
    }))


def serve_alone(reqs, results, k):
    """The same GPU serving requests one at a time: k served requests (every
    j-th by id) each built and stepped alone for the steps it took when
    served; case-steps/s and requests/hr over them (build included)."""
    import gc
    import torch
    from thruster_cx_code_amd.sweep import build_case
    steps = {r.id: int(r.steps) for r in results if not r.error}
    ok = [q for q in reqs if q.id in steps]
    sample = ok[::max(1, len(ok) // k)][:k]
    total_s, total_steps = 0.0, 0
    for q in sample:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sim = build_case(q.params)
        sim.advance(q.params, steps[q.id])
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        total_s += dt
        total_steps += steps[q.id]
        _progress(f"alone: request {q.id} ({q.params['mesh']['nx']}x{q.params['mesh']['ny']}, "
                  f"{steps[q.id]} steps) {dt:.1f} s")
        del sim
        gc.collect()
        torch.cuda.empty_cache()
    return {"requests": [q.id for q in sample], "seconds": total_s, "case_steps": total_steps,
            "case_steps_per_s": total_steps / total_s, "req_per_hr": len(sample) / total_s * 3600.0}


def capacity(results, wall, rate):
    """Device throughput in case-steps/s over the run and the request rate it
    sustains at the run's mean steps per request, against the arrival rate."""
    ok = [r for r in results if not r.error]
    steps = sum(int(r.steps) for r in ok)
    mean_steps = steps / len(ok) if ok else float("nan")
    cs = steps / wall
    return {"case_steps": steps, "case_steps_per_s": cs, "mean_steps_per_request": mean_steps,
            "sustainable_req_per_s": cs / mean_steps if ok else float("nan"), "arrival_req_per_s": rate}


def copy_calibration(mbytes=512, reps=20):
    """A plain device-to-device copy of known size, to scale the FETCH_SIZE and
    WRITE_SIZE counters (run it under rocprofv3 --pmc)."""
    import torch
    n = mbytes * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "reps": reps}))


def main():
    if len(sys.argv) > 1 and sys.argv[1] == "--copy-calibration":
        copy_calibration()
        return
    if len(sys.argv) > 1 and sys.argv[1] == "--poisson-bench":
        poisson_bench()
        return
    if len(sys.argv) > 1 and sys.argv[1] == "--serve":
        sp = argparse.ArgumentParser()
        sp.add_argument("--serve", action="store_true")
        sp.add_argument("--serve-requests", type=int, default=512)
        sp.add_argument("--serve-rate", type=float, default=4.0)
        sp.add_argument("--serve-max-steps", type=int, default=4000)
        sp.add_argument("--serve-slots", type=int, default=64)
        sp.add_argument("--serve-coarse", type=int, default=None,
                        help="coarsest multigrid level of the served cases (nodes; default the solver's)")
        sp.add_argument("--serve-seq", type=int, default=0,
                        help="also time this many served requests one at a time")
        serve_bench(sp.parse_args())
        return
    if len(sys.argv) > 1 and sys.argv[1] == "--sweep":
        sp = argparse.ArgumentParser()
        sp.add_argument("--sweep", action="store_true")
        sp.add_argument("--sweep-cases", type=int, default=64)
        sp.add_argument("--sweep-steps", type=int, default=3000)
        sp.add_argument("--sweep-seq-cases", type=int, default=16)
        sp.add_argument("--sweep-coarse", type=int, default=1024,
                        help="coarsest multigrid level of the batched cases (nodes; 0: the solver's default)")
        sp.add_argument("--sweep-tile", type=int, default=16,
                        help="level-0 tile edge of the batched cases' Poisson solves (0: by mesh size)")
        sp.add_argument("--no-cpu-baseline", action="store_true")
        sweep_bench(sp.parse_args())
        return
    if len(sys.argv) > 1 and sys.argv[1] == "--cpu-sweep":
        cpu_worker_sweep(sys.argv[2])
        return
    if len(sys.argv) > 1 and sys.argv[1] == "--cpu-composed":
        cpu_worker_composed(sys.argv[2], int(sys.argv[3]), int(sys.argv[4]))
        return
    if len(sys.argv) > 1 and sys.argv[1] == "--cpu-worker":
        w, ni, ne, secs, steps = sys.argv[2:7]
        cpu_worker(w, int(ni), int(ne), float(secs), int(steps))
        return
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--workload", default="c2", choices=["c2", "c1", "c5"])
    ap.add_argument("--ion-sort-interval", type=int, default=None,
                    help="steps between cell sorts of the ions (0: never after the preload); "
                         "default 20 (the c2 configuration), 5 for c5")
    ap.add_argument("--electron-sort-interval", type=int, default=None,
                    help="steps between cell sorts of the electrons; default 20 (c2), 1 for c5")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--cpu-steps", type=int, default=0,
                    help="c1: time this many whole CPU steps instead of a --cpu-seconds sample")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the time step shards "
              f"across GPUs by case or by particles, which this build does not run.", file=sys.stderr)
        sys.exit(2)
    run(args)


if __name__ == "__main__":
    main()
