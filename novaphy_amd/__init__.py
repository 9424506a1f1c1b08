"""novaphy_amd -- NovaPhy's PBF fluid step, MI355X-native.

The public names mirror the reference's ``novaphy`` fluid surface
(python/novaphy/__init__.py:97-107 and python/bindings/bind_fluid.cpp): a
C++20 host (``_core``, pybind11) with FluidWorld / PBFSolver / ModelBuilder
that computes through libnpx_hip.so, hand-written HIP kernels for gfx950
behind a C ABI (include/novaphy_hip.h).

The native module is imported on first use of a public name, so that
``novaphy_amd.build`` (which compiles it) runs in a tree where nothing is
built yet.  Using any public name without the native modules raises
ImportError -- there is no CPU fallback.
"""
from __future__ import annotations

import importlib as _importlib
import os as _os

_HERE = _os.path.dirname(_os.path.abspath(__file__))

__all__ = [
    "BoundaryParticle", "CollisionShape", "FluidBlockDef", "FluidWorld", "Model", "ModelBuilder",
    "ParticleState", "PBFSettings", "PBFSolver", "PerformanceMetric", "PerformanceMonitor",
    "PerformancePhaseStat", "RigidBody", "ShapeType", "SimState", "SolverSettings", "SPHKernels",
    "SpatialHashGrid", "Transform", "World", "generate_fluid_block", "sample_model_boundaries",
    "device_count", "FLAG_FAST_MATH", "AABB", "ContactPoint", "BroadPhasePair", "SweepAndPrune",
    "collide_shapes", "gather_poses7", "apply_forces_and_step",
]


def _native():
    try:
        return _importlib.import_module("._core", __name__)
    except ImportError as e:
        raise ImportError(
            "novaphy_amd native modules are not built (libnpx_hip.so / _core); run "
            "`python -m novaphy_amd.build`: " + str(e)) from e


def __getattr__(name):
    if name in __all__:
        value = getattr(_native(), name)
        globals()[name] = value
        return value
    if name == "_core":
        return _native()
    raise AttributeError(f"module {__name__!r} has no attribute {name!r}")


def __dir__():
    return sorted(set(globals()) | set(__all__))
