"""MI355X-native implicit-step hot path of OpenDeformer.

libodhip.so (HIP kernels for gfx950 + C-ABI, include/od.h) does the compute;
this package holds the host-side mirror of the reference's interfaces
(core.py), the ctypes binding (capi.py), the seeded box meshes (mesh.py) and
the benchmark workloads (workloads.py).
"""
import os
import subprocess

_HERE = os.path.dirname(os.path.abspath(__file__))


def build():
    """Compile libodhip.so for gfx950 (hipcc, in-tree)."""
    subprocess.run(["make", "-s", "-C", os.path.join(_HERE, "csrc")], check=True)
    return os.path.join(_HERE, "libodhip.so")


from . import capi, mesh, workloads  # noqa: E402,F401
from .core import (  # noqa: E402,F401
    BackwardEulerIntegrator, CGSolver, Context, DeviceArray, Matrix, Mesh, NewmarkIntegrator, Server,
    generate_matrix_and_virtual_works, mesh_autotune, mesh_spmv_variant, mesh_sweep_form, pcg_solve_noalloc, spmdv,
    spmdv_f32,
)
