"""Seeded synthetic scenes for the dam-break workloads.

The reference fills boxes with ``addRect`` (rtpslib/domain/IV.cpp:165-202: z
outer, y, x inner, ``w = 1``) and pushes them with zero velocity
(rtpslib/system/System.cpp:331-338); ``veleval`` starts at zero
(rtpslib/system/SPH.cpp:482-495).  The dam-break workloads use the same
lattice order with a seeded +-1 % jitter per component, drawn from SplitMix64
(uniform FP32 from the top 24 bits), so identical arrays feed the HIP path
and the CPU oracle.

Workloads (h = 0.04, spacing s = h/2 = 0.02, first site at s/2 from the
corner, rest density 1000, gas constant 3, viscosity 0.01, gravity -9.8 z,
dt = 0.003):

* ``dam16k``: 32x32x16 lattice in the unit box, 100 steps.
* ``dam1m``: 128x128x64 lattice, cube box of side L = (4 * 2.56 * 2.56 * 1.28)^(1/3)
  so the lattice fills a quarter of the volume, 1000 steps.
* ``adjoint64k``: 64x64x16 lattice, the same quarter-volume rule (L = 1.28),
  dt = 0.002, initial velocity N(0, 0.05) per component (Box-Muller).
* batched scenes: counts log-uniform in a range, each a jittered lattice
  block (2:2:1 aspect, truncated at the count like ``addRect``) at a random
  corner of its own quarter-volume box.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
M1 = np.uint64(0xBF58476D1CE4E5B9)
M2 = np.uint64(0x94D049BB133111EB)

H = 0.04
SPACING = 0.02
MASS = 0.008  # rho0 * s^3: the lattice sits at rest density (1009.8 by poly6)


def splitmix64(seed: int, n: int, offset: int = 0) -> np.ndarray:
    """n consecutive SplitMix64 outputs of the stream ``seed`` (uint64)."""
    with np.errstate(over="ignore"):
        ctr = np.arange(offset + 1, offset + n + 1, dtype=np.uint64)
        z = np.uint64(seed) + ctr * GOLDEN
        z = (z ^ (z >> np.uint64(30))) * M1
        z = (z ^ (z >> np.uint64(27))) * M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed: int, n: int, offset: int = 0) -> np.ndarray:
    """Uniform [0, 1) FP32 from the top 24 bits (exactly representable)."""
    z = splitmix64(seed, n, offset)
    return (z >> np.uint64(40)).astype(np.float32) * np.float32(1.0 / (1 << 24))


def normal(seed: int, n: int, sigma: float) -> np.ndarray:
    """N(0, sigma) FP32 by Box-Muller on pairs of uniform draws."""
    m = (n + 1) // 2
    u = uniform01(seed, 2 * m).astype(np.float64)
    u1 = 1.0 - u[0::2]  # (0, 1]
    u2 = u[1::2]
    r = np.sqrt(-2.0 * np.log(u1))
    z = np.concatenate([r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)])
    return (sigma * z[:n]).astype(np.float32)


def lattice(nx: int, ny: int, nz: int, seed: int, origin=(0.0, 0.0, 0.0), spacing: float = SPACING,
            jitter: float = 0.01, count: int | None = None) -> np.ndarray:
    """Jittered simple-cubic lattice, (n, 4) float32 with w = 1, addRect order."""
    n_full = nx * ny * nz
    n = n_full if count is None else min(count, n_full)
    k = np.arange(n, dtype=np.int64)
    ix, iy, iz = k % nx, (k // nx) % ny, k // (nx * ny)
    u = uniform01(seed, 3 * n).astype(np.float64).reshape(n, 3)
    jit = (2.0 * u - 1.0) * (jitter * spacing)
    pos = np.empty((n, 4), np.float32)
    x0 = np.asarray(origin, np.float64) + 0.5 * spacing
    pos[:, 0] = (x0[0] + ix * spacing + jit[:, 0]).astype(np.float32)
    pos[:, 1] = (x0[1] + iy * spacing + jit[:, 1]).astype(np.float32)
    pos[:, 2] = (x0[2] + iz * spacing + jit[:, 2]).astype(np.float32)
    pos[:, 3] = 1.0
    return pos


def quarter_box(nx: int, ny: int, nz: int, spacing: float = SPACING) -> float:
    """Cube side such that an nx*ny*nz lattice block fills a quarter of it."""
    return (4.0 * (nx * spacing) * (ny * spacing) * (nz * spacing)) ** (1.0 / 3.0)


def fmt(v: float) -> str:
    return f"{v:.9g}"


def dam_settings(box: float, max_num: int, gas_constant: float = 3.0, viscosity: float = 0.01,
                 dt: float = 0.003) -> dict:
    """RTPSSettings map of a dam-break scene (keys of rtpslib's SPH settings).

    ``mass``, ``rest_distance``, ``smoothing_distance`` and ``simulation_scale``
    are given explicitly so the workload has h = 0.04 and s = 0.02 (the
    reference would derive them from max_num_particles, SPHSettings.cpp:45-75).
    """
    return {
        "system": "sph",
        "max_num_particles": str(max_num),
        "domain_min": "0 0 0 0",
        "domain_max": f"{fmt(box)} {fmt(box)} {fmt(box)} 0",
        "rest_density": "1000",
        "gas_constant": fmt(gas_constant),
        "viscosity": fmt(viscosity),
        "velocity_limit": "600",
        "xsph_factor": "0.15",
        "friction_kinetic": "0.2",
        "boundary_stiffness": "20000",
        "boundary_dampening": "256",
        "gravity": "0 0 -9.8 0",
        "gravity_alphas": "0 0",
        "gravity_mass": "1 1.5",
        "gravity_sources": "2.5 2.5 2.5 1 7.5 7.5 7.5 1",
        "integrator": "leapfrog",
        "time_step": fmt(dt),
        "mass": fmt(MASS),
        "rest_distance": fmt(SPACING),
        "smoothing_distance": fmt(H),
        "simulation_scale": "1",
    }


@dataclass
class Scene:
    name: str
    settings: dict
    pos: np.ndarray
    vel: np.ndarray
    steps: int
    dt: float
    meta: dict = field(default_factory=dict)

    @property
    def n(self) -> int:
        return self.pos.shape[0]


def dam16k(seed: int = 1001) -> Scene:
    pos = lattice(32, 32, 16, seed)
    return Scene("dam16k", dam_settings(1.0, pos.shape[0]), pos, np.zeros_like(pos), 100, 0.003)


def dam1m(seed: int = 1002) -> Scene:
    box = quarter_box(128, 128, 64)
    pos = lattice(128, 128, 64, seed)
    return Scene("dam1m", dam_settings(box, pos.shape[0]), pos, np.zeros_like(pos), 1000, 0.003)


def dam_custom(nx: int, ny: int, nz: int, seed: int = 1000, box: float | None = None) -> Scene:
    """Dam-break block of any size in its quarter-volume box (test sizes)."""
    box = quarter_box(nx, ny, nz) if box is None else box
    pos = lattice(nx, ny, nz, seed)
    return Scene(f"dam{nx}x{ny}x{nz}", dam_settings(box, pos.shape[0]), pos, np.zeros_like(pos), 10, 0.003)


def dam4m(seed: int = 1006) -> Scene:
    """4,194,304 particles (256x128x128 lattice) in its quarter-volume box."""
    return dam_custom(256, 128, 128, seed=seed)


def batched_equal(num_scenes: int, nx: int, ny: int, nz: int, seed: int) -> list[Scene]:
    """num_scenes dam-break blocks of nx*ny*nz particles each (a uniform batch)."""
    return [dam_custom(nx, ny, nz, seed=seed * 1000 + k) for k in range(num_scenes)]


def adjoint64k(seed: int = 1004, viscosity: float = 0.01, gas_constant: float = 3.0) -> Scene:
    box = quarter_box(64, 64, 16)
    pos = lattice(64, 64, 16, seed)
    n = pos.shape[0]
    vel = np.zeros_like(pos)
    vel[:, :3] = normal(seed + 1, 3 * n, 0.05).reshape(n, 3)
    s = dam_settings(box, n, gas_constant=gas_constant, viscosity=viscosity, dt=0.002)
    return Scene("adjoint64k", s, pos, vel, 200, 0.002)


def batched(num_scenes: int, nmin: int, nmax: int, seed: int) -> list[Scene]:
    """Independent scenes with log-uniform counts, per-scene viscosity and gas constant."""
    u = uniform01(seed, 4 * num_scenes).astype(np.float64).reshape(num_scenes, 4)
    scenes = []
    for s in range(num_scenes):
        n = int(math.floor(math.exp(math.log(nmin) + u[s, 0] * (math.log(nmax) - math.log(nmin)))))
        n = max(nmin, min(nmax, n))
        nz = max(1, math.ceil((n / 4.0) ** (1.0 / 3.0)))
        nx = ny = 2 * nz
        while nx * ny * nz < n:
            nz += 1
        box = quarter_box(nx, ny, nz)
        corner = int(u[s, 3] * 8) & 7
        ox = box - nx * SPACING if corner & 1 else 0.0
        oy = box - ny * SPACING if corner & 2 else 0.0
        oz = box - nz * SPACING if corner & 4 else 0.0
        pos = lattice(nx, ny, nz, seed * 1000 + s, origin=(ox, oy, oz), count=n)
        visc = 0.005 + u[s, 1] * (0.05 - 0.005)
        gas = 1.0 + u[s, 2] * (5.0 - 1.0)
        st = dam_settings(box, n, gas_constant=gas, viscosity=visc)
        scenes.append(Scene(f"scene{s}", st, pos, np.zeros_like(pos), 500, 0.003,
                            meta={"viscosity": visc, "gas_constant": gas, "corner": corner}))
    return scenes
