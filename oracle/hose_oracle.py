"""CPU restatement of the particle hose, for tests only.

Test infrastructure: only tests/ may import this module, as the checker of
the library's hose (rtps_hose_*).  It restates, in pure Python on float32
scalars, one rounding per operation as the library computes them:

  Hose::Hose / update / refill / calc_vectors / calc_em / spray
      (rtpslib/system/common/Hose.cpp:30-128)
  addDiscRandom (rtpslib/domain/IV.cpp:313-365)
  normalize / magnitude (rtpslib/structs.cpp:53-57, :83-87)

The reference draws its perturbations from libc rand() / RAND_MAX; the
restatement, like the library, draws uniform [0, 1) numbers from a SplitMix64
stream (Steele, Lea and Flood, "Fast splittable pseudorandom number
generators", OOPSLA 2014), taking the top 24 bits of each output.
"""
from __future__ import annotations

import math

import numpy as np

f32 = np.float32
MASK = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed: int):
        self.s = seed & MASK

    def next(self) -> int:
        self.s = (self.s + 0x9E3779B97F4A7C15) & MASK
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
        return z ^ (z >> 31)

    def uniform01(self) -> float:
        return (self.next() >> 40) * (1.0 / 16777216.0)


def magnitude(v) -> np.float32:
    x, y, z = f32(v[0]), f32(v[1]), f32(v[2])
    return f32(np.sqrt(f32(f32(x * x) + f32(y * y)) + f32(z * z)))


def normalize(v):
    m = magnitude(v)
    return [f32(v[0]) / m, f32(v[1]) / m, f32(v[2]) / m, f32(v[3])]


class Hose:
    """Hose (Hose.cpp:30-128) with the library's random stream."""

    def __init__(self, dt, total_n, center, velocity, radius, spacing, seed):
        self.dt = f32(dt)
        self.total_n = self.n_count = int(total_n)
        self.em_count = 0
        self.em = 0
        self.center = [f32(c) for c in center]
        self.velocity = [f32(c) for c in velocity]
        self.radius = f32(radius)
        self.spacing = f32(spacing)
        self.rng = SplitMix64(seed)
        self.calc_vectors()

    def update(self, center, velocity, radius, spacing):   # Hose.cpp:48-56
        self.center = [f32(c) for c in center]
        self.velocity = [f32(c) for c in velocity]
        self.radius = f32(radius)
        self.spacing = f32(spacing)
        self.calc_vectors()

    def refill(self, num):   # Hose.cpp:58-62
        self.total_n = self.n_count = int(num)

    def calc_vectors(self):   # Hose.cpp:67-92
        v = self.velocity
        if v[1] == 0:
            v[1] = f32(1e-7)
        if v[2] == 0:
            v[2] = f32(1e-7)
        u = [f32(1), -(v[0] / v[1]), f32(0), f32(1)]
        b = f32(1)
        a = f32(b * v[0]) / v[1]
        c = f32(-b * f32(f32(v[0] * v[0]) + f32(v[1] * v[1]))) / f32(v[1] * v[2])
        self.u = normalize(u)
        self.w = normalize([a, b, c, f32(1)])

    def spray(self):   # Hose.cpp:94-128
        self.em_count += 1
        magv = magnitude(self.velocity)
        self.em = int(f32(1) + f32(f32(f32(self.spacing / self.dt) / magv) / f32(10)))
        if self.em_count >= self.em and self.n_count > 0:
            v = [c * self.dt for c in self.velocity]
            parts = add_disc_random(self.rng, self.n_count, self.center, v, self.u, self.w, self.radius,
                                    self.spacing)
            self.n_count -= len(parts)
            self.em_count = 0
            return parts
        return []


def add_disc_random(rng, num, center, v, u, w, radius, spacing):
    """addDiscRandom (IV.cpp:313-365): a perturbed spacing grid over the
    square [-radius, radius]^2 of the disc plane, kept inside the disc."""
    pert = f32(f32(0.05) * spacing)
    vpert = f32(10)
    umin = [-radius * c for c in u]
    wmin = [-radius * c for c in w]
    r2 = f32(radius * radius)
    out = []
    i = 0
    du = f32(0)
    while float(du) < 2.0 * float(radius):
        dw = f32(0)
        while float(dw) < 2.0 * float(radius):
            if i >= num:
                break
            rrv = f32(rng.uniform01() * float(vpert) * 2.0 - float(vpert))
            rru = f32(rng.uniform01() * float(pert) * 2.0 - float(pert))
            rrw = f32(rng.uniform01() * float(pert) * 2.0 - float(pert))
            du = f32(du + rru)
            dw = f32(dw + rrw)
            p = []
            for k in range(3):
                t = f32(center[k] + umin[k])
                t = f32(t + f32(u[k] * du))
                t = f32(t + wmin[k])
                t = f32(t + f32(w[k] * dw))
                t = f32(t + f32(rrv * v[k]))
                p.append(t)
            d = [f32(p[k] - center[k]) for k in range(3)]
            d2 = f32(f32(f32(d[0] * d[0]) + f32(d[1] * d[1])) + f32(d[2] * d[2]))
            if d2 < r2:
                out.append((p[0], p[1], p[2], f32(1)))
                i += 1
            dw = f32(dw + spacing)
        du = f32(du + spacing)
    return out


def em_of(spacing, dt, velocity) -> int:
    """Spray cadence of calc_em (Hose.cpp:94-109) in double precision, for
    checking the float32 result away from integer boundaries."""
    magv = math.sqrt(sum(float(c) ** 2 for c in velocity[:3]))
    return int(1 + float(spacing) / float(dt) / magv / 10)
