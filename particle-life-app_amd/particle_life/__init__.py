"""particle_life -- MI355X (gfx950) physics step of particle-life.

Host-side mirror of the reference's backend API (physics.py), the C-ABI
binding (engine.py, _native.py), batches of worlds (batch.py), the app's
position/type/matrix presets (presets.py) and seeded benchmark worlds
(worlds.py).
"""
from .batch import Batch
from .engine import Engine
from .physics import (Accelerator, DefaultMatrix, DefaultMatrixGenerator, DefaultPositionSetter,
                      DefaultTypeSetter, GpuPhysics, Matrix, MatrixGenerator, Particle, Particles,
                      ParticleLifeAccelerator, PhysicsSettings, PhysicsSnapshot, PositionSetter,
                      TypeSetter)
from . import presets, worlds

__all__ = [
    "Accelerator", "Batch", "DefaultMatrix", "DefaultMatrixGenerator", "DefaultPositionSetter",
    "DefaultTypeSetter", "Engine", "GpuPhysics", "Matrix", "MatrixGenerator", "Particle",
    "Particles", "ParticleLifeAccelerator", "PhysicsSettings", "PhysicsSnapshot",
    "PositionSetter", "TypeSetter", "presets", "worlds",
]
