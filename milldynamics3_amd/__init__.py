"""MI355X-native engine for MillDynamics3's per-substep slurry-plus-media step.

``Simulation`` mirrors ``mill_core::Simulation`` (crates/mill-core/src/lib.rs:82-566) over the
C ABI declared in ``include/millgpu.h``; ``Batch`` advances many drums with one persistent
launch per substep. The compute path is ``libmillgpu.so`` (HIP, gfx950); there is no CPU path.
"""
from .sim import Batch, Simulation, DEFAULT_PARAMS, params_json, lib, drum_engine_enabled  # noqa: F401

__all__ = ["Simulation", "Batch", "DEFAULT_PARAMS", "params_json", "lib", "drum_engine_enabled"]
