"""mhs_amd -- MI355X-native assembler of mollified Mindlin half-space
triangular-dislocation Green's-function matrices.

A drop-in for the matrix path of brendanjmeade/mhs (``src/mhs/__init__.py``):
the same public functions, signatures, shapes and conventions, with the
per-source assembly loops replaced by hand-written gfx950 HIP kernels behind a
C-ABI (``include/mhs_amd.h``).

    import mhs_amd as mhs
    mat = mhs.Material(mu=30.0, lam=30.0)
    K = mhs.interaction_matrix(tris, mat, eps=0.1, receiver="strike", source="strike")
"""
from . import chunking, consumer, defaults, parallel, tdcs
from .api import (collocation_points, disp_matrix, eigenstress_matrix, elastic_strain_matrix,
                  interaction_matrix, stress_matrix, total_stress_matrix, traction_matrix)
from .materials import Material
from .runtime import get_precision, precision, set_precision
from .shared import SharedNodes, shared_nodes, shared_nodes_device

__all__ = [
    "Material", "disp_matrix", "stress_matrix", "traction_matrix", "interaction_matrix",
    "total_stress_matrix", "eigenstress_matrix", "elastic_strain_matrix", "collocation_points",
    "shared_nodes", "shared_nodes_device", "SharedNodes", "chunking", "consumer", "defaults", "parallel", "tdcs", "precision",
    "set_precision", "get_precision",
]

__version__ = "0.1.0"
