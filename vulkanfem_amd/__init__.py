"""MI355X-native matrix-free multigrid-preconditioned CG for voxel linear
elasticity -- a drop-in for the GPU solver path of hirademirkol/VulkanFEM.

The solver is the HIP library libvfem.so (C-ABI in include/vfem.h); this
package is its Python host interface plus the synthetic problem generators.
"""
from .api import (  # noqa: F401
    BF16, CYCLE_DAMPED, CYCLE_REFERENCE, FP32, FP64, L0_AUTO, L0_ELEMENT, L0_STENCIL, L0_STENCIL_DIRECT, L0_STENCIL_PLANES, MIXED, MatrixFreeSparse, MultigridPreconditioner, VfemError,
    apply_boundary_conditions, assemble_system_matrix, block_configs, header_functions, inspect_hierarchy, lib,
    options, solve_with_hip, UNet,
)
