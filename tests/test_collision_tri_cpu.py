"""Triangle collision restatement (collision_tri_oracle), pinned by closed
forms: a particle falling straight onto a horizontal triangle within the
boundary distance gets stiffness * penetration + dampening * speed along the
normal and nothing else (the friction term of cl_sph_collision.h vanishes
for a downward normal velocity, and is -v_t for an upward one);
outside the triangle, beyond the boundary distance, moving parallel to the
plane or at rest it gets nothing; two triangles forming one square act like
one; the result does not depend on vertex order up to the normal's sign.
"""
import numpy as np
import pytest

import collision_tri_oracle as T

K, DAMP, DIST, SCALE = 20000.0, 256.0, 0.01, 1.0
FLOOR = np.array([[[0.0, 0.0, 0.5], [1.0, 0.0, 0.5], [0.0, 1.0, 0.5]]], np.float32)   # normal +z


def _one(p, v, tris=FLOOR):
    pos = np.array([[*p, 1.0]], np.float32)
    vel = np.array([[*v, 0.0]], np.float32)
    return T.collision_triangles(pos, vel, np.zeros((1, 4), np.float32), tris, SCALE, DIST, K, DAMP)[0]


@pytest.mark.parametrize("h,w", [(0.004, 0.3), (0.009, 1.0), (0.0001, 0.05)])
def test_normal_impact_closed_form(h, w):
    f = _one((0.2, 0.2, 0.5 + h), (0.0, 0.0, -w))
    pen = DIST - h
    want = K * pen + DAMP * w
    assert f[0] == 0.0 and f[1] == 0.0
    assert f[2] == pytest.approx(want, rel=1e-4)


def test_no_force_cases():
    assert np.all(_one((0.2, 0.2, 0.5 + 0.02), (0, 0, -1)) == 0)    # beyond the boundary distance
    assert np.all(_one((0.9, 0.9, 0.504), (0, 0, -1)) == 0)        # outside the triangle
    assert np.all(_one((0.2, 0.2, 0.504), (1, 0, 0)) == 0)         # parallel to the plane
    assert np.all(_one((0.2, 0.2, 0.504), (0, 0, 0)) == 0)         # at rest: NaN direction, no hit


def test_square_of_two_triangles_and_orientation():
    sq = np.array([[[0, 0, 0.5], [1, 0, 0.5], [1, 1, 0.5]], [[0, 0, 0.5], [1, 1, 0.5], [0, 1, 0.5]]], np.float32)
    flipped = sq[:, ::-1, :].copy()
    for p in [(0.7, 0.2, 0.505), (0.2, 0.7, 0.505)]:
        f = _one(p, (0, 0, -0.5), sq)
        assert f[2] == pytest.approx(K * 0.005 + DAMP * 0.5, rel=1e-4)
        # reversed winding: the normal points down, so the test's segment
        # direction check (d > 0) fails for a downward velocity: no force
        assert np.all(_one(p, (0, 0, -0.5), flipped) == 0)
        # ... and the same particle moving up meets the flipped square; here
        # the tangential velocity of cl_sph_collision.h (v - v (n.v) =
        # (0, 0, w + w^2)) has a positive mean, so the friction -v_t adds too
        g = _one((p[0], p[1], 0.495), (0, 0, 0.5), flipped)
        assert g[2] == pytest.approx(-(K * 0.005 + DAMP * 0.5) - (0.5 + 0.5 * 0.5), rel=1e-4)
