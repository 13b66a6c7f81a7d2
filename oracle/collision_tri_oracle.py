"""CPU restatement of the triangle collision, for tests only.

Test infrastructure: only tests/ may import this module, as the checker of
rtps_collision_triangles and of the system's triangle collisions.

  collision_triangles  collision_tri.cl:418-470 (collision_triangle) with
                       collisions_triangle (:175-231),
                       intersect_triangle_segment (:120-170, Ericson,
                       Real-Time Collision Detection pp. 191-192),
                       calculateRepulsionForce / calculateFrictionForce
                       (cl_sph_collision.h:31-73, zero friction
                       coefficients), float32 one rounding per operation,
                       every particle against every triangle
"""
from __future__ import annotations

import numpy as np

f32 = np.float32


def _dot(a, b):
    return f32(a[0] * b[0]) + f32(a[1] * b[1]) + f32(a[2] * b[2])


def _cross(a, b):
    return [a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]]


def collision_triangles(pos_s, vel_s, force, tris, scale, boundary_distance, stiffness, dampening):
    """force + triangle forces, float32 (n, 4).  tris: (m, 3, 3+) world
    vertex coordinates.  Particles with zero velocity get NaN directions and,
    as in the reference, no force."""
    with np.errstate(invalid="ignore", divide="ignore"):
        s = f32(scale)
        dist = f32(boundary_distance)
        k, damp = f32(stiffness), f32(dampening)
        pos = np.asarray(pos_s, np.float32)
        vel = np.asarray(vel_s, np.float32)
        p = [pos[:, c] * s for c in range(3)]
        v = [vel[:, c] for c in range(3)]
        mag = np.sqrt(f32(v[0] * v[0] + v[1] * v[1]) + v[2] * v[2])
        vn = [v[c] / mag for c in range(3)]
        P = [p[c] - vn[c] * dist for c in range(3)]
        Q = [p[c] + vn[c] * dist for c in range(3)]
        qp = [P[c] - Q[c] for c in range(3)]
        f = [np.zeros(len(pos), np.float32) for _ in range(3)]
        for tri in np.asarray(tris, np.float32):
            a = [tri[0, c] * s for c in range(3)]
            b = [tri[1, c] * s for c in range(3)]
            c_ = [tri[2, c] * s for c in range(3)]
            ab = [f32(b[i] - a[i]) for i in range(3)]
            ac = [f32(c_[i] - a[i]) for i in range(3)]
            n = [f32(x) for x in _cross(ab, ac)]
            nm = np.sqrt(f32(f32(n[0] * n[0]) + f32(n[1] * n[1])) + f32(n[2] * n[2]))
            u = [n[i] / nm for i in range(3)]
            d = _dot(qp, n)
            ap = [P[i] - a[i] for i in range(3)]
            t = _dot(ap, n)
            e = _cross(qp, ap)
            vv = _dot(ac, e)
            ww = -_dot(ab, e)
            hit = ~(d <= 0) & ~((t < 0) | (t > d)) & ~((vv < 0) | (vv > d)) & ~((ww < 0) | (vv + ww > d))
            ood = f32(1.0) / d
            vv = vv * ood
            ww = ww * ood
            uu = f32(1.0) - vv - ww
            tp = [uu * a[i] + vv * b[i] + ww * c_[i] for i in range(3)]
            dd = [p[i] - tp[i] for i in range(3)]
            pen = dist - np.sqrt(_dot(dd, dd))
            hit &= pen > f32(1e-6)
            ndv = _dot(u, v)
            m = k * pen - damp * ndv
            vt = [v[i] - v[i] * ndv for i in range(3)]
            fr = ((vt[0] + vt[1]) + vt[2]) / f32(3.0) > 0
            for i in range(3):
                add = m * u[i] + np.where(fr, -vt[i], f32(0.0))
                f[i] = np.where(hit, f[i] + add, f[i]).astype(np.float32)
        out = np.array(force, np.float32, copy=True)
        for i in range(3):
            out[:, i] = out[:, i] + f[i]
        return out
