"""nebula_amd — MI355X-native MPM substep with the reference MPMEngine surface.

The compute path is ``lib/libnebula_mpm.so`` (hand-written gfx950 HIP kernels
behind the C ABI in include/nebula_mpm.h); ``mpm`` mirrors the reference's
MPMEngine over that ABI and ``scenes`` builds the seeded workloads.
"""
from . import mpm  # noqa: F401
from .mpm import (ColliderSet, ConstantWindForce, GravityForce, MaterialModel, MaterialParams, MPMConfig,  # noqa: F401
                  MPMEngine, NebulaError, calcLambda, calcMu, presetJelly, presetMud, presetPowder, presetSand,
                  presetWater)

__all__ = [
    "mpm", "MPMEngine", "MPMConfig", "MaterialParams", "MaterialModel", "ColliderSet", "GravityForce",
    "ConstantWindForce", "NebulaError", "calcMu", "calcLambda", "presetJelly", "presetSand", "presetWater",
    "presetMud", "presetPowder",
]
