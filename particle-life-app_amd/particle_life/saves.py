"""Save states: the app's zip of particles.tsv, physics.toml and matrix.tsv.

Reader and writer of the reference's save format (app/Main.java:1258-1348):

* ``particles.tsv`` -- header ``x	y	vx	vy	color``, then one particle per line:
  position x, y, velocity x, y, type (app/io/ParticlesIO.java:8-46);
* ``matrix.tsv`` -- one matrix row per line, tab-separated (app/io/MatrixIO.java:10-42);
* ``physics.toml`` -- ``boundaries`` ("periodic" / "clamped"), ``radius``,
  ``friction``, ``force``; dt is not saved (app/PhysicsSettingsToml.java:9-34).

Doubles are written as Java's ``Double.toString`` writes them (shortest
round-trip digits, scientific notation outside [1e-3, 1e7)), so files are
lossless in both directions and, for normal doubles, byte-identical to the
app's own output for the same values.  The format is 2D: z and vz are not stored (the app's particles
live in the plane).  As in ``loadState``, entries missing from a zip leave that
part of the state unchanged, unknown entries are reported and skipped, and
unknown settings keys are an error.
"""
from __future__ import annotations

import math
import warnings
import zipfile

import numpy as np
import tomli

PARTICLES_HEADER = "x\ty\tvx\tvy\tcolor"
TOML_KEYS = ("boundaries", "radius", "friction", "force")


def java_double(x: float) -> str:
    """Double.toString (JDK 19 and later): the shortest digits that read back
    as x, in plain notation for 1e-3 <= |x| < 1e7 and as d.ddddE<exp>
    otherwise.  (Java breaks ties among one-digit candidates by also trying
    two digits; that changes the digits only for subnormals.)"""
    x = float(x)
    if math.isnan(x):
        return "NaN"
    if math.isinf(x):
        return "Infinity" if x > 0 else "-Infinity"
    if x == 0.0:
        return "-0.0" if math.copysign(1.0, x) < 0 else "0.0"
    sign = "-" if x < 0 else ""
    a = abs(x)
    r = repr(a)                       # shortest round-trip digits (Python's float repr)
    if "e" in r:
        m, e = r.split("e")
        digits, e10 = m.replace(".", ""), int(e)
    else:
        ip, _, fp = r.partition(".")
        if ip != "0":
            digits, e10 = ip + fp, len(ip) - 1
        else:
            nz = len(fp) - len(fp.lstrip("0"))
            digits, e10 = fp[nz:], -nz - 1
    digits = digits.strip("0") or "0"
    if 1e-3 <= a < 1e7:
        if e10 >= 0:
            ip = digits[: e10 + 1].ljust(e10 + 1, "0")
            fp = digits[e10 + 1:] or "0"
        else:
            ip, fp = "0", "0" * (-e10 - 1) + digits
        return f"{sign}{ip}.{fp}"
    return f"{sign}{digits[0]}.{digits[1:] or '0'}E{e10}"


# ---------------------------------------------------------------------------
# particles.tsv (app/io/ParticlesIO.java:8-46)
# ---------------------------------------------------------------------------
def write_particles(pos, vel, types) -> str:
    pos = np.asarray(pos, dtype=np.float64).reshape(-1, 3)
    vel = np.asarray(vel, dtype=np.float64).reshape(-1, 3)
    types = np.asarray(types).reshape(-1)
    lines = [PARTICLES_HEADER]
    for i in range(types.shape[0]):
        lines.append("\t".join((java_double(pos[i, 0]), java_double(pos[i, 1]),
                                java_double(vel[i, 0]), java_double(vel[i, 1]), str(int(types[i])))))
    return "\n".join(lines) + "\n"


def read_particles(text: str) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """(pos (n,3), vel (n,3), types (n,)); z and vz are 0, as loadParticles sets them."""
    rows = [ln for ln in text.splitlines()[1:] if ln != ""]
    n = len(rows)
    pos = np.zeros((n, 3))
    vel = np.zeros((n, 3))
    types = np.zeros(n, np.int32)
    for i, ln in enumerate(rows):
        f = ln.split("\t")
        pos[i, 0], pos[i, 1] = float(f[0]), float(f[1])
        vel[i, 0], vel[i, 1] = float(f[2]), float(f[3])
        types[i] = int(f[4])
    return pos, vel, types


# ---------------------------------------------------------------------------
# matrix.tsv (app/io/MatrixIO.java:10-42)
# ---------------------------------------------------------------------------
def write_matrix(m) -> str:
    m = np.asarray(m, dtype=np.float64)
    return "".join("\t".join(java_double(v) for v in row) + "\n" for row in m)


def read_matrix(text: str) -> np.ndarray:
    rows = [[float(v) for v in ln.split("\t")] for ln in text.splitlines() if ln != ""]
    n = len(rows)
    m = np.zeros((n, n))
    for i in range(n):
        m[i, :] = rows[i][:n]
    return m


# ---------------------------------------------------------------------------
# physics.toml (app/PhysicsSettingsToml.java:9-34)
# ---------------------------------------------------------------------------
def write_physics_toml(wrap: bool, rmax: float, friction: float, force: float) -> str:
    return (f'boundaries = "{"periodic" if wrap else "clamped"}"\n'
            f"radius = {java_double(rmax)}\n"
            f"friction = {java_double(friction)}\n"
            f"force = {java_double(force)}\n")


def read_physics_toml(text: str, wrap: bool = True, rmax: float = 0.02, friction: float = 0.85,
                      force: float = 1.0) -> dict:
    """Keys missing from the file keep the given values; unknown keys raise
    (the reference's TomlFile.load rejects them)."""
    t = tomli.loads(text)
    unknown = sorted(set(t) - set(TOML_KEYS))
    if unknown:
        raise ValueError(f"unknown key '{unknown[0]}' in physics.toml")
    return {
        "wrap": t.get("boundaries", "periodic" if wrap else "clamped") == "periodic",
        "rmax": float(t.get("radius", rmax)),
        "friction": float(t.get("friction", friction)),
        "force": float(t.get("force", force)),
    }


# ---------------------------------------------------------------------------
# the zip (app/Main.java:1258-1348)
# ---------------------------------------------------------------------------
def save_state(path_or_file, pos, vel, types, matrix, wrap: bool, rmax: float, friction: float,
               force: float) -> None:
    with zipfile.ZipFile(path_or_file, "w", zipfile.ZIP_DEFLATED) as z:
        z.writestr("particles.tsv", write_particles(pos, vel, types))
        z.writestr("physics.toml", write_physics_toml(wrap, rmax, friction, force))
        z.writestr("matrix.tsv", write_matrix(matrix))


def load_state(path_or_file) -> dict:
    """The parts present in the zip: any of "particles" (pos, vel, types),
    "settings" (dict) and "matrix"; img.png is ignored, other names are
    reported and skipped (as Main.loadState does)."""
    out: dict = {}
    with zipfile.ZipFile(path_or_file, "r") as z:
        for name in z.namelist():
            text = z.read(name)
            if name == "particles.tsv":
                out["particles"] = read_particles(text.decode("utf-8"))
            elif name == "physics.toml":
                out["settings"] = read_physics_toml(text.decode("utf-8"))
            elif name == "matrix.tsv":
                out["matrix"] = read_matrix(text.decode("utf-8"))
            elif name == "img.png":
                continue
            else:
                warnings.warn(f"unknown file in zip: {name}")
    return out


def save_physics(physics, path_or_file) -> None:
    """Main.saveState for a GpuPhysics: current particles, settings and matrix."""
    p = physics.particles
    s = physics.settings
    save_state(path_or_file, p.position, p.velocity, p.type, s.matrix.to_array(), s.wrap, s.rmax,
               s.friction, s.force)


def load_physics(physics, path_or_file) -> None:
    """Main.loadState for a GpuPhysics: replaces what the zip holds; dt is kept
    (the app re-sets it every frame); the matrix is followed by ensureTypes."""
    from .physics import DefaultMatrix, Particles

    st = load_state(path_or_file)
    if "particles" in st:
        pos, vel, types = st["particles"]
        physics.particles = Particles(pos, vel, types)
    if "settings" in st:
        s = st["settings"]
        physics.settings.wrap = s["wrap"]
        physics.settings.rmax = s["rmax"]
        physics.settings.friction = s["friction"]
        physics.settings.force = s["force"]
    if "matrix" in st:
        m = st["matrix"]
        physics.settings.matrix = DefaultMatrix(m.shape[0], m)
        physics.ensureTypes()
