"""Configuration files of the headless runner: ``config.json`` -> the
``params`` dict of ``DigitalTwinSimulator`` (reference
Python/run_simulation_from_config.py:30-195) and the set-up the runner applies
to a simulator before ``build_domain`` (:209-232).

Keys, defaults and the material presets are the reference's.  One deliberate
difference: the reference drops ``discharge_chamber``, so ``build_domain``
gets pitch 0 and a half-hole domain of height 0.5 r_screen, whose electrode
masks are empty when r >= Ly; here ``discharge_chamber.pitch_mm`` (or a
top-level ``pitch_mm``) is forwarded as ``params['pitch_mm']``, as the GUI
passes it (Python/main.py:1900).  ``forward_pitch=False`` restores the
reference's behaviour.

Cross-section tables (energy in eV, value in m^2 or yield) are read with
``np.loadtxt`` (whitespace, else comma-separated with a header row), sorted by
energy and fitted as a cubic ``UnivariateSpline`` of log10(value) in
log10(E) with the configured smoothing, into the ``user_cs`` store the
simulator reads (labels 'CX', 'SEE', 'Custom').
"""
import json
import os

import numpy as np

MATERIAL_PRESETS = {
    'Molybdenum': {'k': 138.0, 'rho': 10280.0, 'cp': 250.0, 'emissivity': 0.80, 'alpha': 4.8e-6,
                   'E_mod': 329e9, 'Y_coeff': 1.05e-4, 'E_th': 30.0},
    'Steel (SS316)': {'k': 16.3, 'rho': 8000.0, 'cp': 500.0, 'emissivity': 0.60, 'alpha': 16.0e-6,
                      'E_mod': 193e9, 'Y_coeff': 2.8e-4, 'E_th': 25.0},
    'Titanium': {'k': 21.9, 'rho': 4507.0, 'cp': 520.0, 'emissivity': 0.50, 'alpha': 8.6e-6,
                 'E_mod': 116e9, 'Y_coeff': 1.8e-4, 'E_th': 20.0},
    'Graphite': {'k': 120.0, 'rho': 2200.0, 'cp': 710.0, 'emissivity': 0.85, 'alpha': 3.0e-6,
                 'E_mod': 11e9, 'Y_coeff': 3.5e-4, 'E_th': 15.0},
}

OUTPUT_KEYS = ('grid_temperatures', 'beam_divergence', 'saddle_point_potential',
               'mean_particle_energy', 'iteration_time')


def load_cross_section(path, smoothing=0.0, label='CX'):
    """One table as the loaders store it: {'energy', 'cs', 'spline', 'type'}
    (Python/run_simulation_from_config.py:112-140)."""
    from scipy.interpolate import UnivariateSpline
    try:
        raw = np.loadtxt(path, delimiter=None, comments='#')
    except Exception:
        raw = np.loadtxt(path, delimiter=',', comments='#', skiprows=1)
    if raw.ndim != 2 or raw.shape[1] < 2:
        raise ValueError(f"{path} must have at least 2 columns")
    energy, cs = raw[:, 0], raw[:, 1]
    order = np.argsort(energy)
    energy, cs = energy[order], cs[order]
    log_e = np.log10(np.maximum(energy, 1e-30))
    log_cs = np.log10(np.maximum(cs, 1e-50))
    spline = UnivariateSpline(log_e, log_cs, s=smoothing, k=3)
    return {'energy': energy, 'cs': cs, 'spline': spline, 'type': label}


def parse_config(config_path='config.json', forward_pitch=True, base_dir=None, quiet=True):
    """Returns (params, output_selection) like the reference's parse_config.
    Relative cross-section paths resolve against base_dir (default: the
    working directory, as the reference opens them)."""
    if not os.path.isfile(config_path):
        raise FileNotFoundError(f"Config file '{config_path}' not found.")
    with open(config_path, 'r', encoding='utf-8') as f:
        config = json.load(f)
    base_dir = base_dir if base_dir is not None else os.getcwd()
    params = {}
    beam = config.get('beam_species', {})
    params['beam_mass_amu'] = beam.get('mass_amu', 131.293)
    params['beam_charge_state'] = beam.get('charge_state', 1)

    mat_cfg = config.get('grid_material', {})
    preset = mat_cfg.get('preset', 'Molybdenum').strip()
    if preset in MATERIAL_PRESETS:
        params['_material'] = (preset, MATERIAL_PRESETS[preset])
    else:
        params['_material'] = ('Custom', {
            'k': mat_cfg.get('k', 138.0), 'rho': mat_cfg.get('rho', 10280.0), 'cp': mat_cfg.get('cp', 250.0),
            'emissivity': mat_cfg.get('emissivity', 0.80), 'alpha': mat_cfg.get('alpha', 4.8e-6),
            'E_mod': mat_cfg.get('E_mod', 329e9), 'Y_coeff': mat_cfg.get('Y_coeff', 1.05e-4),
            'E_th': mat_cfg.get('E_th', 30.0)})

    adv = config.get('advanced_settings', {})
    params['neut_x'] = adv.get('neut_x', 19.9)
    params['neut_r'] = adv.get('neut_r', 3.0)
    params['V_plasma_offset'] = adv.get('V_plasma_offset', 20.0)
    params['m_e_ratio'] = adv.get('m_e_ratio', 1000.0)
    params['Lx'] = adv.get('Lx', 20.0)
    params['Ly'] = adv.get('Ly', 3.0)

    cs_cfg = config.get('cross_sections', {})
    smoothing = cs_cfg.get('spline_smoothing', 0.0)
    cs_store = {}
    for label, key in (('CX', 'cx_file'), ('SEE', 'see_file'), ('Custom', 'custom_file')):
        fpath = cs_cfg.get(key, '').strip()
        if not fpath:
            continue
        full = fpath if os.path.isabs(fpath) else os.path.join(base_dir, fpath)
        if not os.path.isfile(full):
            if not quiet:
                print(f"Warning: cross-section file not found for {label}: {fpath}")
            continue
        try:
            cs_store[label] = load_cross_section(full, smoothing, label)
        except Exception as exc:
            if not quiet:
                print(f"Warning: failed to load {label} CS from {fpath}: {exc}")
    params['_cs_store'] = cs_store

    sim_cfg = config.get('simulation', {})
    params['n0_plasma'] = sim_cfg.get('n0_plasma', 1e17)
    params['Te_up'] = sim_cfg.get('Te_up', 3.0)
    params['Ti'] = sim_cfg.get('Ti', 2.0)
    params['Tn'] = sim_cfg.get('Tn', 300.0)
    params['n0'] = sim_cfg.get('n0', 1e20)
    params['Accel'] = sim_cfg.get('Accel', 0.1)
    params['Thresh'] = sim_cfg.get('Thresh', 10000.0)
    params['sim_mode'] = sim_cfg.get('sim_mode', 'Both')
    params['geometry'] = sim_cfg.get('geometry', 'half_hole')

    rf_cfg = config.get('rf_co_extraction', {})
    params['rf_enable'] = rf_cfg.get('rf_enable', False)
    params['rf_grid_idx'] = rf_cfg.get('rf_grid_idx', 0)
    params['rf_freq'] = rf_cfg.get('rf_freq', 13.56)
    params['rf_amp'] = rf_cfg.get('rf_amp', 500.0)

    neut_cfg = config.get('neutralizer', {})
    params['neut_rate'] = neut_cfg.get('neut_rate', 0)
    params['Te'] = neut_cfg.get('Te', 5.0)

    params['grids'] = [{'V': g.get('V', 0.0), 't': g.get('t', 1.0), 'gap': g.get('gap', 1.0),
                        'r': g.get('r', 1.0), 'cham': g.get('cham', 0.0)} for g in config.get('grids', [])]
    if not params['grids']:
        raise ValueError("No grids found in config.json. Add at least one entry in the 'grids' list.")

    if forward_pitch:
        dc = config.get('discharge_chamber')
        pitch = config.get('pitch_mm', dc.get('pitch_mm') if isinstance(dc, dict) else None)
        if pitch is not None:
            params['pitch_mm'] = float(pitch)

    out_cfg = config.get('terminal_output', {})
    output_selection = {k: out_cfg.get(k, True) for k in OUTPUT_KEYS}
    return params, output_selection


def apply_config(sim, params):
    """The runner's set-up of a simulator from parsed params
    (Python/run_simulation_from_config.py:209-232): beam species, electron
    mass ratio, grid material, user cross-sections.  Pops the private keys
    from params as the reference does, and returns the material name."""
    mass_amu = params.pop('beam_mass_amu', 131.293)
    charge = params.pop('beam_charge_state', 1)
    sim.m_ion = mass_amu * 1.6605e-27
    sim.m_XE = sim.m_ion
    sim.Z_ion = charge
    sim.q_ion = charge * sim.q
    sim.m_e = sim.m_ion / params.get('m_e_ratio', 1000.0)
    name, props = params.pop('_material', ('Molybdenum', None))
    if props:
        sim.set_material(props=props)
    else:
        sim.set_material(name='Molybdenum')
    sim.user_cs = params.pop('_cs_store', {})
    return name
