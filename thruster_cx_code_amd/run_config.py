"""Headless runner: build a simulator from a config.json and step it,
printing the selected quantities every 10 iterations (reference
Python/run_simulation_from_config.py:197-290), on the MI355X backend.

    python -m thruster_cx_code_amd.run_config config.json [--steps N] [--csv history.csv]

The reference runs until interrupted; ``--steps`` bounds the run.  ``step``
returns five values here (the reference's runner unpacks four and fails).
"""
import argparse
import sys
import time


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("config", nargs="?", default="config.json")
    ap.add_argument("--steps", type=int, default=0, help="stop after N steps (0: until interrupted)")
    ap.add_argument("--csv", default="", help="write the step history to this CSV at the end")
    ap.add_argument("--preserve-reference-pitch", action="store_true",
                    help="drop discharge_chamber.pitch_mm as the reference loader does")
    args = ap.parse_args(argv)
    from .config import parse_config, apply_config
    from .physics_engine import DigitalTwinSimulator
    from .telemetry import StepHistory
    params, out = parse_config(args.config, forward_pitch=not args.preserve_reference_pitch, quiet=False)
    print("Using config file:", args.config)
    sim = DigitalTwinSimulator()
    mass_amu, charge = params.get('beam_mass_amu', 131.293), params.get('beam_charge_state', 1)
    mat = apply_config(sim, params)
    print(f"Beam species: {mass_amu:.3f} amu, charge +{charge}")
    print(f"Grid material: {mat} (k={sim.mat_k:.1f}, rho={sim.mat_rho:.0f}, cp={sim.mat_cp:.0f})")
    if sim.user_cs:
        print(f"Cross-section datasets loaded: {', '.join(sim.user_cs.keys())}")
    print("Building simulation domain...")
    sim.build_domain(params)
    print(f"Domain built with {len(params['grids'])} grids ({sim.nx} x {sim.ny} nodes)")
    for i, g in enumerate(params['grids'], start=1):
        print(f"Built Grid {i}: V={g['V']} V, t={g['t']} mm, gap={g['gap']} mm, r={g['r']} mm, cham={g['cham']} deg")
    print("Starting simulation...")
    hist = StepHistory()
    try:
        while args.steps <= 0 or sim.iteration < args.steps:
            t0 = time.time()
            res = sim.step(params)
            dt_wall = time.time() - t0
            hist.record(sim, res)
            remeshed, min_pot, current_div, T_grids, _ = res
            if sim.iteration % 10 == 0:
                s = f"Iteration: {sim.iteration} | "
                if out.get('grid_temperatures'):
                    s += "Grid Temps: [" + ', '.join(f"G{i + 1}: {T - 273.15:.1f}C" for i, T in enumerate(T_grids)) + "] | "
                if out.get('beam_divergence'):
                    s += f"Divergence: {current_div:.2f} deg | "
                if out.get('saddle_point_potential'):
                    s += f"Saddle Potential: {min_pot:.2f} V | "
                if out.get('mean_particle_energy') and sim.num_p > 0:
                    s += f"Mean Ion Energy: {sim.mean_ion_energy_eV():.2f} eV | "
                if out.get('iteration_time'):
                    s += f"Step Time: {dt_wall:.4f} s"
                print(s, flush=True)
            if remeshed:
                print("Domain has been remeshed due to thermal or erosion effects.")
    except KeyboardInterrupt:
        print("\nSimulation stopped by user.")
    if args.csv and len(hist):
        hist.export_csv(args.csv)
        print(f"CSV saved: {args.csv}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
