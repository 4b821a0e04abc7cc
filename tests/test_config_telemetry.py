"""Headless-runner configuration and run telemetry (CPU): config.json parsing
against the reference's own configuration files, cross-section table
loading, the runner's simulator set-up, and the history / tracking CSV
layouts the GUI writes (Python/run_simulation_from_config.py:30-232,
Python/main.py:2316-2470)."""
import csv
import json
import math
import os

import numpy as np
import pytest

from thruster_cx_code_amd import config as cfg
from thruster_cx_code_amd import telemetry as tel

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "config")


def test_parse_reference_config():
    params, out = cfg.parse_config(os.path.join(GOLD, "config.json"))
    assert params['beam_mass_amu'] == 131.293 and params['beam_charge_state'] == 1
    assert params['_material'] == ('Molybdenum', cfg.MATERIAL_PRESETS['Molybdenum'])
    assert (params['neut_x'], params['neut_r'], params['Lx'], params['Ly']) == (10.0, 1.5, 20.0, 3.0)
    assert params['n0_plasma'] == 7.72926180e17 and params['Te_up'] == 8.5
    assert params['sim_mode'] == 'Thermal' and params['geometry'] == 'half_hole'
    assert params['rf_enable'] is False and params['rf_freq'] == 13.56
    assert params['neut_rate'] == 30 and params['Te'] == 5.0
    assert [g['V'] for g in params['grids']] == [1300.0, -150.0, 50.0]
    assert params['grids'][0] == {'V': 1300.0, 't': 1.2, 'gap': 0.9, 'r': 1.3, 'cham': 20.0}
    assert params['_cs_store'] == {}
    assert params['pitch_mm'] == 3.0                   # discharge_chamber.pitch_mm forwarded
    assert out == {k: True for k in cfg.OUTPUT_KEYS}
    # the reference loader's behaviour: no pitch key
    p2, _ = cfg.parse_config(os.path.join(GOLD, "config.json"), forward_pitch=False)
    assert 'pitch_mm' not in p2
    p2.pop('_material'), p2.pop('_cs_store'), params.pop('_material'), params.pop('_cs_store')
    params.pop('pitch_mm')
    assert p2 == params


def test_parse_rf_config_and_defaults(tmp_path):
    params, _ = cfg.parse_config(os.path.join(GOLD, "configRF.json"))
    assert len(params['grids']) == 2 and params['grids'][1]['t'] == 2
    # every section missing except grids: the reference's defaults
    p = tmp_path / "min.json"
    p.write_text(json.dumps({"grids": [{}], "terminal_output": {"iteration_time": False}}))
    params, out = cfg.parse_config(str(p))
    assert params['grids'] == [{'V': 0.0, 't': 1.0, 'gap': 1.0, 'r': 1.0, 'cham': 0.0}]
    assert params['Thresh'] == 10000.0 and params['Accel'] == 0.1 and params['sim_mode'] == 'Both'
    assert params['rf_amp'] == 500.0 and params['neut_rate'] == 0 and params['m_e_ratio'] == 1000.0
    assert out['iteration_time'] is False and out['beam_divergence'] is True
    assert 'pitch_mm' not in params


def test_parse_errors(tmp_path):
    with pytest.raises(FileNotFoundError):
        cfg.parse_config(str(tmp_path / "absent.json"))
    p = tmp_path / "nogrids.json"
    p.write_text(json.dumps({"grids": []}))
    with pytest.raises(ValueError):
        cfg.parse_config(str(p))


def test_custom_material(tmp_path):
    p = tmp_path / "c.json"
    p.write_text(json.dumps({"grids": [{"V": 1.0}], "grid_material": {"preset": "Unobtainium", "k": 50.0}}))
    params, _ = cfg.parse_config(str(p))
    name, props = params['_material']
    assert name == 'Custom' and props['k'] == 50.0 and props['rho'] == 10280.0 and props['E_th'] == 30.0


def _write_cs_config(tmp_path, cs_path, smoothing=0.0):
    p = tmp_path / "cs.json"
    p.write_text(json.dumps({"grids": [{"V": 1.0}],
                             "cross_sections": {"cx_file": cs_path, "see_file": "missing.txt",
                                                "spline_smoothing": smoothing}}))
    return str(p)


def test_cross_section_table_interpolates(tmp_path):
    src = os.path.join(GOLD, "sample_cs.txt")
    raw = np.loadtxt(src)
    params, _ = cfg.parse_config(_write_cs_config(tmp_path, "sample_cs.txt"), base_dir=GOLD)
    store = params['_cs_store']
    assert set(store) == {'CX'}                         # the missing SEE file is skipped
    t = store['CX']
    assert t['type'] == 'CX' and np.array_equal(t['energy'], np.sort(raw[:, 0]))
    # smoothing 0: the spline passes through every point of the table
    got = 10.0 ** t['spline'](np.log10(t['energy']))
    assert np.allclose(got, t['cs'], rtol=1e-9)
    # absolute path, no base_dir
    params, _ = cfg.parse_config(_write_cs_config(tmp_path, src))
    assert 'CX' in params['_cs_store']


def test_cross_section_comma_with_header_and_sorting(tmp_path):
    f = tmp_path / "t.csv"
    E = np.array([300.0, 1.0, 50.0, 700.0, 10.0, 2000.0])
    cs = 1e-19 * (1.0 + np.log(E))
    f.write_text("energy_eV,cs_m2\n" + "".join(f"{e},{c}\n" for e, c in zip(E, cs)))
    t = cfg.load_cross_section(str(f), 0.0, 'SEE')
    order = np.argsort(E)
    assert np.array_equal(t['energy'], E[order]) and np.array_equal(t['cs'], cs[order])
    assert np.allclose(10.0 ** t['spline'](np.log10(E)), cs, rtol=1e-9)
    bad = tmp_path / "one.txt"
    bad.write_text("1\n2\n3\n4\n5\n")
    with pytest.raises(ValueError):
        cfg.load_cross_section(str(bad))


class _StubSim:
    q = 1.602e-19

    def __init__(self):
        self.material = None

    def set_material(self, name=None, props=None):
        self.material = (name, props)


def test_apply_config_sets_species_material_tables(tmp_path):
    params, _ = cfg.parse_config(_write_cs_config(tmp_path, os.path.join(GOLD, "sample_cs.txt")))
    params['beam_mass_amu'], params['beam_charge_state'], params['m_e_ratio'] = 83.798, 2, 500.0
    sim = _StubSim()
    name = cfg.apply_config(sim, params)
    assert name == 'Molybdenum'
    assert sim.m_ion == 83.798 * 1.6605e-27 and sim.m_XE == sim.m_ion
    assert sim.Z_ion == 2 and sim.q_ion == 2 * sim.q and sim.m_e == sim.m_ion / 500.0
    assert sim.material == (None, cfg.MATERIAL_PRESETS['Molybdenum'])
    assert set(sim.user_cs) == {'CX'}
    for k in ('beam_mass_amu', 'beam_charge_state', '_material', '_cs_store'):
        assert k not in params                       # popped, as the reference's runner does


class _HistSim:
    dt = 5e-10

    def __init__(self):
        self.iteration = 0
        self.total_active_cells = 0
        self.low_ppc_cells = 0
        self._trans = 0.0

    def get_transparency(self):
        return self._trans


def test_history_csv_layout(tmp_path):
    sim, h = _HistSim(), tel.StepHistory()
    with pytest.raises(ValueError):
        h.export_csv(str(tmp_path / "empty.csv"))
    for it in range(1, 4):
        sim.iteration, sim.total_active_cells, sim.low_ppc_cells, sim._trans = it, 100 + it, it, 0.5 + 0.1 * it
        h.record(sim, (False, -3.0 * it, 1.5 * it if it != 2 else float('nan'), [300.0 + it, 310.0 + it], 0.9))
    path = h.export_csv(str(tmp_path / "h.csv"))
    rows = list(csv.reader(open(path, encoding='utf-8')))
    assert rows[0] == ['iteration', 't_sim_s', 'minpotential', 'beamdivergence°', 'transparency',
                       'total_active_cells', 'cells_less_than_3_macroparticles', 'grid_1_temp_K', 'grid_2_temp_K']
    assert len(rows) == 4
    assert rows[1][0] == '1' and float(rows[1][1]) == 5e-10 and float(rows[1][2]) == -3.0
    assert math.isnan(float(rows[2][3])) and float(rows[3][4]) == pytest.approx(0.8)
    assert rows[3][5:] == ['103', '3', '303.0', '313.0']
    # ragged optional column: empty field, as the GUI writes it
    h.transparency_history.pop()
    rows = list(csv.reader(open(h.export_csv(str(tmp_path / "h2.csv")), encoding='utf-8')))
    assert rows[3][4] == ''


class _KinSim:
    def __init__(self, chunks):
        self._chunks = list(chunks)

    def get_particle_kinematics(self):
        return self._chunks.pop(0)


def test_tracking_csv_layout(tmp_path):
    a = np.arange(16, dtype=np.float64).reshape(2, 8)
    b = np.arange(8, dtype=np.float64).reshape(1, 8) + 100
    buf = tel.TrackingBuffer()
    with pytest.raises(ValueError):
        buf.export_csv(str(tmp_path / "none.csv"))
    sim = _KinSim([a, np.empty((0, 8)), b])
    for _ in range(3):
        buf.append(sim)
    rows = list(csv.reader(open(buf.export_csv(str(tmp_path / "t.csv")), encoding='utf-8')))
    assert rows[0] == ['time_s', 'x_mm', 'y_mm', 'vx_ms', 'vy_ms', 'vz_ms', 'energy_eV', 'type']
    got = np.array(rows[1:], dtype=np.float64)
    assert np.array_equal(got, np.vstack([a, b]))
