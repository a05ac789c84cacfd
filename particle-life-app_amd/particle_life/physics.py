"""Host-side mirror of the reference's physics API over the HIP library.

The reference's physics package (src/main/java/com/particle_life/backend/) is
mirrored class for class -- Physics, PhysicsSettings, Accelerator, Matrix /
DefaultMatrix, Particle, TypeSetter / PositionSetter / MatrixGenerator and
their defaults -- with the Java method names, so code written against the
reference reads the same.  ``GpuPhysics`` is the drop-in for ``Physics``: its
``update()`` runs the gfx950 kernels through the C ABI (include/plife.h); there
is no CPU code path.

Host state follows the reference's ownership (Physics.particles is the
API-visible truth): the device holds a cached copy that is re-uploaded when the
host view may have been mutated and downloaded lazily after steps.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass
from typing import Callable, Optional

import numpy as np

from . import _native
from ._native import PlifeArgumentError
from .engine import Engine


# ---------------------------------------------------------------------------
# Matrix (backend/Matrix.java:3-8, backend/DefaultMatrix.java)
# ---------------------------------------------------------------------------
class Matrix:
    """Interface of backend/Matrix.java:3-8."""

    def size(self) -> int:
        raise NotImplementedError

    def get(self, i: int, j: int) -> float:
        raise NotImplementedError

    def set(self, i: int, j: int, value: float) -> None:
        raise NotImplementedError

    def deepCopy(self) -> "Matrix":
        raise NotImplementedError

    def to_array(self) -> np.ndarray:
        n = self.size()
        return np.array([[self.get(i, j) for j in range(n)] for i in range(n)], dtype=np.float64)


class DefaultMatrix(Matrix):
    """backend/DefaultMatrix.java: a size x size table, row = acted-on type
    (get at DefaultMatrix.java:39-41)."""

    def __init__(self, size: int, values=None):
        self.values = np.zeros((size, size), dtype=np.float64) if values is None else \
            np.array(values, dtype=np.float64).reshape(size, size)

    def size(self) -> int:
        return int(self.values.shape[0])

    def get(self, i: int, j: int) -> float:
        return float(self.values[i, j])

    def set(self, i: int, j: int, value: float) -> None:
        self.values[i, j] = value

    def zero(self) -> None:
        self.values[:] = 0.0

    def randomize(self, rng: Optional[np.random.Generator] = None) -> None:
        """DefaultMatrix.java:25-31: entries 2*U - 1."""
        rng = rng or np.random.default_rng()
        self.values[:] = 2.0 * rng.random(self.values.shape) - 1.0

    def deepCopy(self) -> "DefaultMatrix":
        return DefaultMatrix(self.size(), self.values.copy())

    def to_array(self) -> np.ndarray:
        return self.values

    def __eq__(self, other) -> bool:
        return isinstance(other, Matrix) and other.size() == self.size() and \
            np.array_equal(other.to_array(), self.values)


# ---------------------------------------------------------------------------
# PhysicsSettings (backend/PhysicsSettings.java:8-38)
# ---------------------------------------------------------------------------
@dataclass(eq=False)
class PhysicsSettings:
    wrap: bool = True
    rmax: float = 0.02
    friction: float = 0.85
    force: float = 1.0
    dt: float = float(np.float32(0.02))   # the float literal 0.02f widened (PhysicsSettings.java:37)
    matrix: Matrix = None

    def __post_init__(self):
        if self.matrix is None:
            self.matrix = DefaultMatrix(6)

    def deepCopy(self) -> "PhysicsSettings":
        return PhysicsSettings(self.wrap, self.rmax, self.friction, self.force, self.dt,
                               self.matrix.deepCopy())

    def __eq__(self, o) -> bool:
        return isinstance(o, PhysicsSettings) and (o.wrap, o.rmax, o.friction, o.force, o.dt) == \
            (self.wrap, self.rmax, self.friction, self.force, self.dt) and o.matrix == self.matrix


# ---------------------------------------------------------------------------
# Accelerator (backend/Accelerator.java:7-16, app/Main.java:275-280)
# ---------------------------------------------------------------------------
class Accelerator:
    """backend/Accelerator.java: accelerate(a, pos) -> acceleration relative to rmax."""

    def accelerate(self, a: float, pos: np.ndarray) -> np.ndarray:
        raise NotImplementedError


class ParticleLifeAccelerator(Accelerator):
    """The app's accelerator (app/Main.java:275-280) as a named, parameterised
    type, so the device path can read beta.  ``accelerate`` is kept for API
    completeness (it is what the lambda computes); the device never calls it."""

    def __init__(self, beta: float = 0.3):
        self.beta = float(beta)

    def accelerate(self, a: float, pos: np.ndarray) -> np.ndarray:
        dist = math.sqrt(pos[0] * pos[0] + (pos[1] * pos[1] + pos[2] * pos[2]))
        b = self.beta
        f = (dist / b - 1) if dist < b else a * (1 - abs(1 + b - 2 * dist) / (1 - b))
        return np.asarray(pos, dtype=np.float64) * (f / dist)


# ---------------------------------------------------------------------------
# Setters and generators (backend/TypeSetter.java, PositionSetter.java,
# MatrixGenerator.java and the Default* classes)
# ---------------------------------------------------------------------------
class TypeSetter:
    def getType(self, position, velocity, type_: int, nTypes: int) -> int:
        raise NotImplementedError


class PositionSetter:
    def set(self, position: np.ndarray, type_: int, nTypes: int) -> None:
        raise NotImplementedError


class MatrixGenerator:
    def makeMatrix(self, size: int) -> Matrix:
        raise NotImplementedError


class DefaultTypeSetter(TypeSetter):
    """DefaultTypeSetter.java:9 -- uniform type (seedable here)."""

    def __init__(self, rng: Optional[np.random.Generator] = None):
        self.rng = rng or np.random.default_rng()

    def getType(self, position, velocity, type_: int, nTypes: int) -> int:
        return int(math.floor(self.rng.random() * nTypes))


class DefaultPositionSetter(PositionSetter):
    """DefaultPositionSetter.java:9-13 -- uniform in the unit square, z = 0."""

    def __init__(self, rng: Optional[np.random.Generator] = None):
        self.rng = rng or np.random.default_rng()

    def set(self, position: np.ndarray, type_: int, nTypes: int) -> None:
        position[0] = self.rng.random()
        position[1] = self.rng.random()
        position[2] = 0.0


class DefaultMatrixGenerator(MatrixGenerator):
    """DefaultMatrixGenerator.java -- a randomized DefaultMatrix."""

    def __init__(self, rng: Optional[np.random.Generator] = None):
        self.rng = rng or np.random.default_rng()

    def makeMatrix(self, size: int) -> Matrix:
        m = DefaultMatrix(size)
        m.randomize(self.rng)
        return m


# ---------------------------------------------------------------------------
# Particles (backend/Particle.java:5-9) as structure-of-arrays
# ---------------------------------------------------------------------------
@dataclass
class Particle:
    position: np.ndarray
    velocity: np.ndarray
    type: int = 0


class Particles:
    """Physics.particles as arrays: position (n,3), velocity (n,3), type (n,),
    plus the id each particle had when the state was last uploaded."""

    def __init__(self, position: np.ndarray, velocity: np.ndarray, type_: np.ndarray,
                 ids: Optional[np.ndarray] = None):
        self.position = np.ascontiguousarray(position, dtype=np.float64).reshape(-1, 3)
        self.velocity = np.ascontiguousarray(velocity, dtype=np.float64).reshape(-1, 3)
        self.type = np.ascontiguousarray(type_, dtype=np.int32).reshape(-1)
        self.ids = np.arange(len(self.type), dtype=np.int32) if ids is None else ids

    def __len__(self) -> int:
        return int(self.type.shape[0])

    def __getitem__(self, i: int) -> Particle:
        return Particle(self.position[i].copy(), self.velocity[i].copy(), int(self.type[i]))


class PhysicsSnapshot:
    """app/PhysicsSnapshot.java:11-62: double[3N] positions/velocities + types."""

    def __init__(self, positions=None, velocities=None, types=None, settings=None,
                 particleCount: int = 0, typeCount=None):
        self.positions = positions
        self.velocities = velocities
        self.types = types
        self.settings = settings
        self.particleCount = particleCount
        self.typeCount = typeCount
        self.snapshotTime = 0

    def take(self, physics: "GpuPhysics") -> None:
        """PhysicsSnapshot.take (app/PhysicsSnapshot.java:24-34): rewrites this
        snapshot's arrays in place, re-allocated only when the particle count
        changed (PhysicsSnapshot.java:39-44), straight from the device when it
        holds the newer state; then the settings, type counts and time stamp."""
        physics._write_snapshot(self)


# ---------------------------------------------------------------------------
# GpuPhysics: the drop-in for backend/Physics.java
# ---------------------------------------------------------------------------
class GpuPhysics:
    """Physics (backend/Physics.java:9-511) whose update() runs on the GPU.

    Only a ``ParticleLifeAccelerator`` is accepted (its beta is a device
    parameter); any other Accelerator raises, as the device has no way to run
    arbitrary host code per pair.  ``dim=3`` enables the 3D extension.
    """

    DEFAULT_MATRIX_SIZE = 7

    def __init__(self, accelerator: Accelerator, positionSetter: PositionSetter = None,
                 matrixGenerator: MatrixGenerator = None, typeSetter: TypeSetter = None,
                 device: int = 0, dim: int = 2, particleCount: int = 10000):
        if not isinstance(accelerator, ParticleLifeAccelerator):
            raise NotImplementedError("GpuPhysics needs a ParticleLifeAccelerator (beta-zone force law)")
        self.accelerator = accelerator
        self.positionSetter = positionSetter or DefaultPositionSetter()
        self.matrixGenerator = matrixGenerator or DefaultMatrixGenerator()
        self.typeSetter = typeSetter or DefaultTypeSetter()
        self.settings = PhysicsSettings()
        self.dim = int(dim)
        self.preferredNumberOfThreads = 12   # kept for API parity; the device ignores it
        self._engine = Engine(device)
        self._particles = Particles(np.zeros((0, 3)), np.zeros((0, 3)), np.zeros(0, np.int32))
        self._host_dirty = True     # host view may differ from the device copy
        self._device_newer = False  # device stepped since the host view was refreshed
        self.generateMatrix()
        self.setParticleCount(particleCount)

    # -- lifecycle ----------------------------------------------------------
    def kill(self) -> None:
        """Physics.kill (Physics.java:156-158): release the device state."""
        if getattr(self, "_engine", None) is not None:
            self._engine.close()
            self._engine = None

    def __del__(self):
        try:
            self.kill()
        except Exception:
            pass

    # -- the hot path -------------------------------------------------------
    def update(self, steps: int = 1) -> int:
        """Physics.update (Physics.java:112-134), `steps` times on the device."""
        self._push_settings()
        if self._host_dirty:
            self._upload()
        rc = self._engine.step(int(steps))
        self._device_newer = True
        return rc

    def forceUpdateStop(self) -> None:
        """Physics.forceUpdateStop (Physics.java:149-151), honoured between steps."""
        self._engine.stop()

    # -- particles ----------------------------------------------------------
    @property
    def particles(self) -> Particles:
        """Physics.particles.  Accessing it refreshes the host view and marks it
        possibly mutated, so the next update() re-uploads it."""
        self._download()
        self._host_dirty = True
        return self._particles

    @particles.setter
    def particles(self, p: Particles) -> None:
        self._particles = p
        self._device_newer = False
        self._host_dirty = True

    def snapshot(self) -> PhysicsSnapshot:
        """A new PhysicsSnapshot taken of this physics (PhysicsSnapshot.take)."""
        s = PhysicsSnapshot()
        s.take(self)
        return s

    def _write_snapshot(self, snap: PhysicsSnapshot) -> None:
        n = self._engine.n if self._device_newer else len(self._particles)
        if snap.types is None or len(snap.types) != n:
            snap.positions = np.empty(3 * n)
            snap.velocities = np.empty(3 * n)
            snap.types = np.empty(n, np.int32)
        if self._device_newer:
            # positions, velocities and types only: no ids, no host view refresh
            self._engine.download(snap.positions.reshape(n, 3), snap.velocities.reshape(n, 3),
                                  snap.types)
        else:
            p = self._particles
            snap.positions[:] = p.position.reshape(-1)
            snap.velocities[:] = p.velocity.reshape(-1)
            snap.types[:] = p.type
        snap.settings = self.settings.deepCopy()
        snap.particleCount = n
        snap.typeCount = np.bincount(snap.types, minlength=self.settings.matrix.size())
        snap.snapshotTime = int(time.time() * 1000)

    def editParticles(self, indices, position=None, velocity=None, type_=None) -> None:
        """In-place edits of some particles between steps (the app's cursor
        tools, app/Main.java:543-580): the given rows of positions, velocities
        and/or types change in the host view and only those rows go to the
        device (plife_set_particle_slots).  Positions are wrapped or clamped
        (ensurePosition) on the way."""
        self._download()
        p = self._particles
        idx = np.asarray(indices, dtype=np.int64).reshape(-1)
        if position is not None:
            p.position[idx] = np.asarray(position, dtype=np.float64).reshape(-1, 3)
            for i in idx:
                self.ensurePosition(p.position[i])
        if velocity is not None:
            p.velocity[idx] = np.asarray(velocity, dtype=np.float64).reshape(-1, 3)
        if type_ is not None:
            p.type[idx] = np.asarray(type_, dtype=np.int32).reshape(-1)
        if self._host_dirty:
            return                   # the whole view goes up at the next update anyway
        self._push_settings()
        self._engine.set_slots(idx, p.position[idx], p.velocity[idx], p.type[idx])

    def setPositions(self) -> None:
        """Physics.setPositions (Physics.java:166-168).  With one of the app's
        presets as the position setter (presets.PresetPositionSetter) it runs
        on the device for all particles at once; other setters run per
        particle on the host."""
        from .presets import PresetPositionSetter
        if isinstance(self.positionSetter, PresetPositionSetter):
            self._on_device(lambda e: e.set_positions(self.positionSetter.index,
                                                      self.positionSetter.next_seed()))
            return
        p = self.particles
        for i in range(len(p)):
            self._set_position(p, i)

    def generateMatrix(self) -> None:
        """Physics.generateMatrix (Physics.java:170-176)."""
        prev = self.settings.matrix.size() if self.settings.matrix is not None else self.DEFAULT_MATRIX_SIZE
        self.settings.matrix = self.matrixGenerator.makeMatrix(prev)

    def setParticleCount(self, n: int) -> None:
        """Physics.setParticleCount (Physics.java:190-223)."""
        if n < 0:
            raise PlifeArgumentError(_native.PLIFE_ERR_ARGUMENT, "particle count must be >= 0")
        p = self.particles if len(self._particles) else self._particles
        old = len(p)
        if n == old and old:
            return
        if n < old:
            order = np.random.default_rng().permutation(old)[:n]
            self.particles = Particles(p.position[order], p.velocity[order], p.type[order])
            return
        pos = np.zeros((n, 3))
        vel = np.zeros((n, 3))
        typ = np.zeros(n, np.int32)
        pos[:old], vel[:old], typ[:old] = p.position, p.velocity, p.type
        newp = Particles(pos, vel, typ)
        for i in range(old, n):
            self._set_type(newp, i)
            self._set_position(newp, i)
        self.particles = newp

    def setMatrixSize(self, newSize: int) -> None:
        """Physics.setMatrixSize (Physics.java:238-258)."""
        prev = self.settings.matrix
        if newSize == prev.size():
            return
        m = self.matrixGenerator.makeMatrix(newSize)
        c = min(prev.size(), newSize)
        for i in range(c):
            for j in range(c):
                m.set(i, j, prev.get(i, j))
        self.settings.matrix = m
        if newSize < prev.size():
            self.ensureTypes()

    def ensureTypes(self) -> None:
        """Physics.ensureTypes (Physics.java:266-272)."""
        p = self.particles
        for i in np.nonzero(p.type >= self.settings.matrix.size())[0]:
            self._set_type(p, int(i))

    def setTypes(self) -> None:
        """Physics.setTypes (Physics.java:509-511); on the device with a
        presets.PresetTypeSetter, as setPositions."""
        from .presets import PresetTypeSetter
        if isinstance(self.typeSetter, PresetTypeSetter):
            self._on_device(lambda e: e.set_types(self.typeSetter.index, self.typeSetter.next_seed()))
            return
        p = self.particles
        for i in range(len(p)):
            self._set_type(p, i)

    # -- helpers of the reference API ---------------------------------------
    def connection(self, pos1, pos2) -> np.ndarray:
        """Physics.connection (Physics.java:460-470) with Range.wrapConnection."""
        d = np.asarray(pos2, np.float64) - np.asarray(pos1, np.float64)
        if self.settings.wrap:
            d = np.where(d < -0.5, d + 1, np.where(d >= 0.5, d - 1, d))
            if self.dim == 2:
                d[2] = 0.0
        return d

    def distance(self, pos1, pos2) -> float:
        d = self.connection(pos1, pos2)
        return math.sqrt(d[0] * d[0] + (d[1] * d[1] + d[2] * d[2]))

    def ensurePosition(self, position: np.ndarray) -> None:
        """Physics.ensurePosition (Physics.java:499-505)."""
        k = 3 if self.dim == 3 else 2
        for a in range(k):
            v = position[a]
            if self.settings.wrap:
                v = v - math.floor(v)
            else:
                v = 0.0 if v < 0 else (1.0 if v > 1 else v)
            position[a] = v
        if self.dim == 2:
            position[2] = 0.0

    def _set_position(self, p: Particles, i: int) -> None:
        self.positionSetter.set(p.position[i], int(p.type[i]), self.settings.matrix.size())
        self.ensurePosition(p.position[i])
        p.velocity[i] = 0.0

    def _set_type(self, p: Particles, i: int) -> None:
        p.type[i] = self.typeSetter.getType(p.position[i].copy(), p.velocity[i].copy(),
                                            int(p.type[i]), self.settings.matrix.size())

    # -- device sync --------------------------------------------------------
    def _on_device(self, op) -> None:
        """Runs op(engine) on the device copy of the particles, bringing it up
        to date first; the host view is refreshed lazily afterwards."""
        self._push_settings()
        if self._host_dirty:
            self._upload()
        op(self._engine)
        self._device_newer = True

    def _push_settings(self) -> None:
        s = self.settings
        self._engine.configure(matrix=s.matrix.to_array(), rmax=s.rmax, friction=s.friction,
                               dt=s.dt, beta=self.accelerator.beta, force=s.force, wrap=s.wrap,
                               dim=self.dim)

    def _upload(self) -> None:
        p = self._particles
        self._push_settings()
        self._engine.set_state(p.position, p.velocity, p.type)
        self._host_dirty = False
        self._device_newer = False

    def _download(self) -> None:
        if not self._device_newer:
            return
        st = self._engine.state()
        self._particles = Particles(st["pos"], st["vel"], st["types"], st["ids"])
        self._device_newer = False
        self._host_dirty = False
