"""MI355X-native D3Q19 lattice-Boltzmann wind tunnel (HIP kernels for gfx950 behind a C ABI).

``lattice_fluid_dynamics_amd.lbm`` binds include/lbm.h; ``build()`` compiles
liblbm_amd.so in place with hipcc (kernels) and gcc (host library).
"""
import os
import subprocess

PACKAGE_DIR = os.path.dirname(os.path.abspath(__file__))
CSRC_DIR = os.path.join(PACKAGE_DIR, "csrc")


def build(jobs: int = 4) -> str:
    """Compiles liblbm_amd.so for gfx950 in the package directory; returns its path."""
    subprocess.run(["make", "-s", "-C", CSRC_DIR, f"-j{jobs}"], check=True)
    return os.path.join(PACKAGE_DIR, "liblbm_amd.so")


def rebuild(jobs: int = 4) -> str:
    """Forces a full recompilation."""
    subprocess.run(["make", "-s", "-C", CSRC_DIR, "clean"], check=True)
    return build(jobs)
