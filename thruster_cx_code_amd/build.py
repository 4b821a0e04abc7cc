"""Build libbemcs_hip.so (gfx950) in-tree with hipcc.

Usage: python -m thruster_cx_code_amd.build
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB_NAME = "libbemcs_hip.so"
LIB_PATH = os.path.join(HERE, LIB_NAME)
SOURCES = ["bemcs_particles.hip", "bemcs_poisson.hip", "bemcs_diag.hip", "bemcs_surface.hip"]
ARCH = os.environ.get("BEMCS_OFFLOAD_ARCH", "gfx950")


def hipcc_path():
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (set HIPCC or install ROCm)")


def build(verbose=False):
    """Compile every HIP source into one shared library next to this file.

    -ffp-contract=off keeps every floating-point expression in the written
    order (bit-identical results with the NumPy formulation); packed FP32
    instructions are disabled for gfx950.
    """
    cmd = [hipcc_path(), "-O3", "--offload-arch=" + ARCH, "-fPIC", "-shared", "-std=c++17",
           "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-o", LIB_PATH + ".tmp"]
    cmd += [os.path.join(CSRC, s) for s in SOURCES]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + proc.stdout + proc.stderr)
    os.replace(LIB_PATH + ".tmp", LIB_PATH)
    if verbose:
        print("built", LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
