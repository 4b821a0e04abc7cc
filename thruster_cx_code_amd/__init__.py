"""MI355X (gfx950) backend for the PY-BEMCS DigitalTwinSimulator time step.

Hand-written HIP kernels behind a C ABI (``include/bemcs_hip.h``, built into
``libbemcs_hip.so``), driven from Python through ctypes; torch-ROCm provides
device buffers and streams only.

    from thruster_cx_code_amd import DigitalTwinSimulator
    sim = DigitalTwinSimulator()
    sim.build_domain(params)
    remeshed, min_pot, div, T_grids, trans = sim.step(params)
"""
KERNEL_BACKEND = "hip"

__all__ = ["KERNEL_BACKEND", "DigitalTwinSimulator", "push_particles_boris_hip", "accumulate_rho_hip",
           "accumulate_rho_cic_hip", "thermal_conduction_hip"]


def __getattr__(name):
    if name == "DigitalTwinSimulator":
        from .physics_engine import DigitalTwinSimulator
        return DigitalTwinSimulator
    if name in ("push_particles_boris_hip", "accumulate_rho_hip", "accumulate_rho_cic_hip",
                "thermal_conduction_hip"):
        from . import kernels
        return getattr(kernels, name)
    raise AttributeError(name)
