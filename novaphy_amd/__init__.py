"""novaphy_amd -- NovaPhy's PBF fluid step, MI355X-native.

The public names mirror the reference's ``novaphy`` fluid surface
(python/novaphy/__init__.py:97-107 and python/bindings/bind_fluid.cpp): a
C++20 host (``_core``, pybind11) with FluidWorld / PBFSolver / ModelBuilder
that computes through libnpx_hip.so, hand-written HIP kernels for gfx950
behind a C ABI (include/novaphy_hip.h).  Importing fails loudly when the
native modules have not been built -- there is no CPU fallback.
"""
from __future__ import annotations

import os as _os

_HERE = _os.path.dirname(_os.path.abspath(__file__))

try:
    from . import _core
except ImportError as e:  # pragma: no cover - exercised only on unbuilt trees
    raise ImportError(
        "novaphy_amd native modules are not built (libnpx_hip.so / _core); run "
        "`python -m novaphy_amd.build`: " + str(e)) from e

from ._core import (  # noqa: E402
    BoundaryParticle, CollisionShape, FluidBlockDef, FluidWorld, Model, ModelBuilder, ParticleState,
    PBFSettings, PBFSolver, PerformanceMetric, PerformanceMonitor, PerformancePhaseStat, RigidBody,
    ShapeType, SimState, SolverSettings, SPHKernels, SpatialHashGrid, Transform, World,
    generate_fluid_block, sample_model_boundaries, device_count, FLAG_FAST_MATH,
)

__all__ = [
    "BoundaryParticle", "CollisionShape", "FluidBlockDef", "FluidWorld", "Model", "ModelBuilder",
    "ParticleState", "PBFSettings", "PBFSolver", "PerformanceMetric", "PerformanceMonitor",
    "PerformancePhaseStat", "RigidBody", "ShapeType", "SimState", "SolverSettings", "SPHKernels",
    "SpatialHashGrid", "Transform", "World", "generate_fluid_block", "sample_model_boundaries",
    "device_count", "FLAG_FAST_MATH",
]
