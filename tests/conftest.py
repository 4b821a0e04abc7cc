import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU and the built HIP library")


def c1_params(nx=128, ny=48, ions=2000, electrons=2000, seed=0, geometry="half_hole",
              neutral_model="uniform", cx_model="loglinear", neut_rate=30, n0_plasma=1e17):
    """Two-grid Xe+ extraction (screen +1000 V, 0.38 mm, r 0.95 mm; accel
    -200 V, 0.5 mm, r 0.55 mm; 0.6 mm gap), Ti 0.2 eV, neutraliser Te 5 eV,
    uniform neutral Xe 1e18 m^-3 at 300 K, log-linear CEX cross-section."""
    return {
        "grids": [{"V": 1000.0, "t": 0.38, "gap": 0.6, "r": 0.95, "cham": 0.0},
                  {"V": -200.0, "t": 0.5, "gap": 1.0, "r": 0.55, "cham": 0.0}],
        "pitch_mm": 6.1447, "geometry": geometry, "V_plasma_offset": 20.0,
        "n0_plasma": n0_plasma, "Te_up": 3.0, "Ti": 0.2, "Te": 5.0, "Tn": 300.0,
        "n0": 1e18, "neutral_model": neutral_model, "cx_model": cx_model,
        "neut_rate": neut_rate, "sim_mode": "Trajectories", "mesh": {"nx": nx, "ny": ny},
        "preload": {"ions": ions, "electrons": electrons}, "seed": seed,
    }


@pytest.fixture
def params_small():
    return c1_params()


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def hip_lib():
    if not gpu_available():
        pytest.skip("no GPU")
    from thruster_cx_code_amd import _lib
    return _lib.load()
