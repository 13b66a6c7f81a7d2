"""Rigid-body coupling (SPH::interact, SPH.cpp:610-663) against the oracle's
restatement of staticrigidbody_force.cl / rigidbody_force.cl (itself pinned
by a two-particle closed form and an O(N M) brute force in test_oracle.py).

Bars: with a static boundary of particles holding a falling block (the box
walls far away), 10 steps within 1e-4 relative per position element and 1e-4
of the velocity field's max-norm of the FP32 oracle; the same with a moving
rigid body's particles pushing the fluid (5 steps); update() + interact() +
integrate() equals step() to 1e-6; removing the sets returns to the plain step.
"""
import numpy as np
import pytest

import oracle as O
from enjaparticles_amd import scenes
from enjaparticles_amd.rtps import RTPS, rigid_spring_dampening

pytestmark = pytest.mark.gpu

# test/test1.xml's rigid system: penetration_factor .02, restitution .15, velocity_limit 600
SPRING, DAMP = rigid_spring_dampening(0.02, 600.0, 0.04, 0.15)


def _rel_max(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-30)


def _scene():
    pos = scenes.lattice(16, 16, 8, 91, origin=(0.3, 0.3, 0.31))
    st = scenes.dam_settings(1.0, len(pos))
    return st, pos


def _boundary_cup():
    """A floor at z = 0.29 and four side walls up to z = 0.55 around the block,
    spacing 0.02 (world units)."""
    g = np.arange(0.26, 0.6401, 0.02)
    zs = np.arange(0.29, 0.5501, 0.02)
    pts = [(x, y, 0.29) for x in g for y in g]
    for z in zs[1:]:
        for t in g:
            pts += [(g[0], t, z), (g[-1], t, z), (t, g[0], z), (t, g[-1], z)]
    b = np.unique(np.round(np.array(pts), 6), axis=0)
    return np.c_[b, np.ones(len(b))].astype(np.float32)


def _body():
    """A 5x5x5 rigid body beside the block, moving into it."""
    b = scenes.lattice(5, 5, 5, 92, origin=(0.2, 0.38, 0.33), jitter=0.0)
    vel = np.zeros_like(b)
    vel[:, 0] = 0.8
    mass = np.full(len(b), 0.012, np.float32)
    return b, vel, mass


def _compare(s, st, n, k):
    pos, vel = s.read("pos"), s.read("vel")
    np.testing.assert_allclose(pos[:, :3], st.pos_u[:n, :3], rtol=1e-4, atol=0, err_msg=f"step {k}")
    assert _rel_max(vel[:, :3], st.vel_u[:n, :3]) <= 1e-4, k


def test_static_boundary_holds_the_fluid_like_the_oracle(cuda):
    st_, pos = _scene()
    b = _boundary_cup()
    s = RTPS.generateSystemInstance(st_)
    s.pushParticles(pos, np.zeros_like(pos))
    s.setStaticBoundary(b, SPRING, DAMP)
    sp, gp, gs = O.derive_params(st_)
    o = O.State(pos, np.zeros_like(pos), sp, gp, gs)
    O.set_rigid("static", b, SPRING, DAMP)
    try:
        for k in range(10):
            s.step(1)
            o.step(0.003, O.FLAG_COLLISION)
            _compare(s, o, len(pos), k)
    finally:
        O.set_rigid("static", None)
    # the boundary pushed: the bottom layer was decelerated against free fall
    free = O.State(pos, np.zeros_like(pos), sp, gp, gs)
    for _ in range(10):
        free.step(0.003, O.FLAG_COLLISION)
    assert s.read("vel")[:256, 2].mean() > free.vel_u[:256, 2].mean() + 0.01
    s.close()


def test_moving_rigid_body_pushes_the_fluid_like_the_oracle(cuda):
    st_, pos = _scene()
    b, bv, bm = _body()
    s = RTPS.generateSystemInstance(st_)
    s.pushParticles(pos, np.zeros_like(pos))
    s.setRigidParticles(b, bv, bm, SPRING, DAMP)
    sp, gp, gs = O.derive_params(st_)
    o = O.State(pos, np.zeros_like(pos), sp, gp, gs)
    O.set_rigid("dynamic", b, SPRING, DAMP, vel=bv, mass=bm)
    try:
        for k in range(5):
            s.step(1)
            o.step(0.003, O.FLAG_COLLISION)
            _compare(s, o, len(pos), k)
    finally:
        O.set_rigid("dynamic", None)
    assert s.read("vel")[:, 0].max() > 0.05    # pushed along +x
    s.close()


def test_interact_between_update_and_integrate_equals_step(cuda):
    st_, pos = _scene()
    b = _boundary_cup()
    body, bv, bm = _body()
    a = RTPS.generateSystemInstance(st_)
    c = RTPS.generateSystemInstance(st_)
    for s in (a, c):
        s.pushParticles(pos, np.zeros_like(pos))
        s.setStaticBoundary(b, SPRING, DAMP)
        s.setRigidParticles(body, bv, bm, SPRING, DAMP)
    for _ in range(4):
        a.step(1)
        c.update()
        c.interact()
        c.integrate()
    np.testing.assert_allclose(c.read("pos"), a.read("pos"), rtol=1e-6, atol=1e-7)
    np.testing.assert_allclose(c.read("vel"), a.read("vel"), rtol=1e-5, atol=1e-5)
    # without the sets, a step is the plain one again
    for s in (a, c):
        s.setStaticBoundary(np.zeros((0, 4), np.float32), SPRING, DAMP)
        s.setRigidParticles(np.zeros((0, 4), np.float32), np.zeros((0, 4), np.float32), np.zeros(0, np.float32),
                            SPRING, DAMP)
    sp, gp, gs = O.derive_params(st_)
    o = O.State(a.read("pos"), a.read("vel"), sp, gp, gs, veleval=a.read("veleval"))
    a.step(1)
    o.step(0.003, O.FLAG_COLLISION)
    _compare(a, o, len(pos), 0)
    a.close()
    c.close()
