"""Seeded synthetic worlds for the benchmark configurations.

The reference draws everything from unseeded ``Math.random`` (e.g.
backend/DefaultPositionSetter.java:10-11, backend/DefaultTypeSetter.java:9,
backend/DefaultMatrix.java:28), so reproducible inputs are this project's own:
counter-based SplitMix64 (Vigna's published generator: output k of a stream is
the SplitMix64 finaliser of base + (k+1) * 0x9E3779B97F4A7C15).

Worlds are generated directly in the reference's unit box.  A configuration
stated for the side-2 box [-1, 1)^D maps as x_ref = (x + 1) / 2,
rmax_ref = rmax / 2 (velocities halve as well); dt, force, beta and the matrix
are unchanged.  A friction half-life t maps to the reference's per-1/60 s
friction coefficient 0.5 ** (1 / (60 t)) (Physics.java:401 applies
friction ** (60 dt) per step).

Uniform draws are on a 2^-24 grid so FP32, the device's 32-bit fixed point and
FP64 all hold identical values.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)
_STREAM = np.uint64(0xD1B54A32D192ED03)


def _mix(z: np.ndarray) -> np.ndarray:
    z = (z ^ (z >> np.uint64(30))) * _M1
    z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def stream_u64(seed: int, stream: int, n: int) -> np.ndarray:
    """n outputs of SplitMix64 stream `stream` of `seed`."""
    with np.errstate(over="ignore"):
        base = _mix(np.array([np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + np.uint64(stream) * _STREAM],
                             dtype=np.uint64))[0]
        k = np.arange(1, n + 1, dtype=np.uint64)
        return _mix(base + k * GAMMA)


def uniform24(seed: int, stream: int, n: int) -> np.ndarray:
    """Uniform [0, 1) on the 2^-24 grid."""
    return (stream_u64(seed, stream, n) >> np.uint64(40)).astype(np.float64) * 2.0 ** -24


def friction_from_half_life(t_half: float) -> float:
    return 0.5 ** (1.0 / (60.0 * t_half))


def config_seed(config: int, world: int = 0) -> int:
    return 0x5EED0000 + 1000 * config + world


@dataclass
class World:
    """One world: state in the unit box plus its settings."""

    pos: np.ndarray            # (n, 3) float64
    vel: np.ndarray            # (n, 3) float64
    types: np.ndarray          # (n,) int32
    matrix: np.ndarray         # (T, T) float64, row = acted-on type
    rmax: float
    friction: float
    dt: float
    beta: float = 0.3
    force: float = 1.0
    wrap: bool = True
    dim: int = 2
    name: str = ""
    meta: dict = field(default_factory=dict)

    @property
    def n(self) -> int:
        return int(self.types.shape[0])

    @property
    def ntypes(self) -> int:
        return int(self.matrix.shape[0])


def uniform_world(seed: int, n: int, ntypes: int, rmax: float, dt: float, beta: float = 0.3,
                  friction: float | None = None, dim: int = 2, wrap: bool = True,
                  force: float = 1.0, name: str = "") -> World:
    """Positions uniform in the unit box, types uniform, matrix uniform [-1, 1), v = 0."""
    pos = np.zeros((n, 3))
    pos[:, 0] = uniform24(seed, 1, n)
    pos[:, 1] = uniform24(seed, 2, n)
    if dim == 3:
        pos[:, 2] = uniform24(seed, 3, n)
    types = np.minimum((uniform24(seed, 4, n) * ntypes).astype(np.int32), ntypes - 1)
    m = 2.0 * uniform24(seed, 5, ntypes * ntypes) - 1.0
    return World(pos=pos, vel=np.zeros((n, 3)), types=types, matrix=m.reshape(ntypes, ntypes),
                 rmax=rmax, dt=dt, beta=beta,
                 friction=friction_from_half_life(0.04) if friction is None else friction,
                 dim=dim, wrap=wrap, force=force, name=name)


def config1(n: int = 10_000) -> World:
    """Plain CPU-runnable world: N=10,000, 6 types, rmax 0.04 in [-1,1)^2
    (0.02 in the unit box), beta 0.3, half-life 0.04 s, dt 0.02 s."""
    return uniform_world(config_seed(1), n, 6, rmax=0.02, dt=0.02, name="c1_plain_10k_2d")


def config2(n: int = 1_000_000) -> World:
    """Single large 2D world: N=1M, 8 types, rmax 0.01 in [-1,1)^2 (0.005 in the
    unit box, ~78.5 interacting neighbours), beta 0.3, half-life 0.04 s, dt 0.01 s."""
    return uniform_world(config_seed(2), n, 8, rmax=0.005, dt=0.01, name="c2_large_1m_2d")


def gaussian_blob_world(seed: int, n: int, ntypes: int, rmax: float, dt: float, beta: float,
                        nblobs: int = 64, sigma: float = 0.015, dim: int = 3,
                        name: str = "") -> World:
    """Mixture of isotropic Gaussian blobs (centres uniform), wrapped into the
    unit box and snapped to the 2^-24 grid."""
    centres = np.stack([uniform24(seed, 10 + d, nblobs) for d in range(dim)], axis=1)
    blob = np.minimum((uniform24(seed, 20, n) * nblobs).astype(np.int64), nblobs - 1)
    pos = np.zeros((n, 3))
    for d in range(dim):
        # Box-Muller from two 53-bit uniforms in (0, 1]
        u1 = ((stream_u64(seed, 30 + 2 * d, n) >> np.uint64(11)).astype(np.float64) + 1.0) * 2.0 ** -53
        u2 = (stream_u64(seed, 31 + 2 * d, n) >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
        g = np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)
        x = centres[blob, d] + sigma * g
        x = np.round(x * 2.0 ** 24) * 2.0 ** -24
        x = x - np.floor(x)
        pos[:, d] = x
    types = np.minimum((uniform24(seed, 4, n) * ntypes).astype(np.int32), ntypes - 1)
    m = 2.0 * uniform24(seed, 5, ntypes * ntypes) - 1.0
    return World(pos=pos, vel=np.zeros((n, 3)), types=types, matrix=m.reshape(ntypes, ntypes),
                 rmax=rmax, dt=dt, beta=beta, friction=friction_from_half_life(0.04),
                 dim=dim, name=name)


def config3(n: int = 4_000_000) -> World:
    """Clustered 3D world: N=4M, 12 types, 64 Gaussian blobs (sigma 0.03 in
    [-1,1)^3, 0.015 in the unit box), rmax 0.02 (0.01), beta 0.25, dt 0.01 s."""
    return gaussian_blob_world(config_seed(3), n, 12, rmax=0.01, dt=0.01, beta=0.25,
                               name="c3_clustered_4m_3d")


def config4_specs(nworlds: int = 512) -> list[dict]:
    """Parameters of the batched workload: nworlds independent 2D worlds with
    N log-uniform in [1000, 65536], types uniform in {3..12}, rmax uniform in
    [0.01, 0.04] (unit box), beta uniform in [0.2, 0.4], half-life 0.04 s,
    dt 0.01 s, periodic.  Drawn from the batch's own seed, so a prefix of the
    list does not depend on nworlds."""
    s = config_seed(4)
    u_n, u_t, u_r, u_b = (uniform24(s, k, nworlds) for k in (101, 102, 103, 104))
    lo, hi = np.log(1000.0), np.log(65536.0)
    specs = []
    for w in range(nworlds):
        specs.append(dict(
            seed=config_seed(4, w),
            n=int(np.rint(np.exp(lo + (hi - lo) * u_n[w]))),
            ntypes=3 + min(int(u_t[w] * 10), 9),
            rmax=0.01 + 0.03 * float(u_r[w]),
            beta=0.2 + 0.2 * float(u_b[w]),
        ))
    return specs


def config4(nworlds: int = 512) -> list[World]:
    """The batched workload's worlds (see config4_specs); about 7.9M particles
    in total at 512 worlds."""
    return [uniform_world(p["seed"], p["n"], p["ntypes"], rmax=p["rmax"], dt=0.01, beta=p["beta"],
                          name=f"c4_world_{w}")
            for w, p in enumerate(config4_specs(nworlds))]


def config5_jobs(njobs: int = 200, rate: float = 40.0, max_steps: int = 2000) -> list[dict]:
    """The served workload: a Poisson stream of jobs (rate jobs/s), each one
    world with N log-uniform in [2000, 262144], dim 2 or 3 (even odds), 4-10
    types, rmax uniform in [0.01, 0.03] (unit box), beta 0.3, half-life
    0.04 s, dt 0.01 s, periodic, run for a step count uniform in
    [100, max_steps].  Returns dicts with job, arrival (s), steps and the
    world's parameters; job_world() builds the world itself."""
    s = config_seed(5)
    u = [uniform24(s, k, njobs) for k in (201, 202, 203, 204, 205, 206)]
    lo, hi = np.log(2000.0), np.log(262144.0)
    gaps = -np.log1p(-u[0]) / rate                   # exponential inter-arrival times
    arrival = np.cumsum(gaps) - gaps[0]
    jobs = []
    for j in range(njobs):
        jobs.append(dict(
            job=j,
            arrival=float(arrival[j]),
            steps=100 + min(int(u[5][j] * (max_steps - 99)), max_steps - 100),
            seed=config_seed(5, j),
            n=int(np.rint(np.exp(lo + (hi - lo) * u[1][j]))),
            dim=2 if u[2][j] < 0.5 else 3,
            ntypes=4 + min(int(u[3][j] * 7), 6),
            rmax=0.01 + 0.02 * float(u[4][j]),
        ))
    return jobs


def job_world(job: dict) -> World:
    """The world of one served job (config5_jobs)."""
    return uniform_world(job["seed"], job["n"], job["ntypes"], rmax=job["rmax"], dt=0.01, beta=0.3,
                         dim=job["dim"], name=f"c5_job_{job['job']}")
