"""Triangle collision on the GPU (rtps_collision_triangles, and the system's
collision stage with loaded triangles) against collision_tri_oracle.

Bars: forces within 1e-4 of the largest triangle force (FP32 with fused
multiply-adds on the GPU, one rounding per operation in the restatement);
hit / no-hit decisions agree except for particles whose penetration is
within 1e-5 of the 1e-6 threshold (none in these cases).
"""
import numpy as np
import pytest

import collision_tri_oracle as T
from enjaparticles_amd import ops, scenes
from enjaparticles_amd.rtps import RTPS

pytestmark = pytest.mark.gpu


def _terrain(n_side, z0, amp, seed):
    """A wavy triangulated sheet over [0.05, 0.95]^2 at height ~z0 (normals up)."""
    rng = np.random.default_rng(seed)
    xs = np.linspace(0.05, 0.95, n_side + 1)
    zz = z0 + amp * rng.uniform(-1, 1, size=(n_side + 1, n_side + 1))
    tris = []
    for i in range(n_side):
        for j in range(n_side):
            a = (xs[i], xs[j], zz[i, j])
            b = (xs[i + 1], xs[j], zz[i + 1, j])
            c = (xs[i + 1], xs[j + 1], zz[i + 1, j + 1])
            d = (xs[i], xs[j + 1], zz[i, j + 1])
            tris += [(a, b, c), (a, c, d)]
    return np.array(tris, np.float32)


@pytest.mark.parametrize("ntri", [648, 1, 2, 256, 257])
def test_stage_matches_restatement(cuda, ntri):
    """648 triangles fill three 256-triangle LDS chunks, the last one partly;
    256 and 257 sit on a chunk boundary; 1 and 2 leave most threads of the
    chunk without a triangle."""
    import torch
    rng = np.random.default_rng(11)
    tris = _terrain(18, 0.5, 0.01, 3)[:ntri]
    n = 20000
    pos = np.zeros((n, 4), np.float32)
    pos[:, :2] = rng.uniform(0.0, 1.0, (n, 2))
    pos[:, 2] = 0.5 + rng.uniform(-0.03, 0.03, n)
    pos[:, 3] = 1.0
    vel = np.zeros((n, 4), np.float32)
    vel[:, :3] = rng.normal(0, 0.5, (n, 3))
    vel[::97] = 0.0                             # at rest: no hit
    force = rng.normal(0, 1, (n, 4)).astype(np.float32)
    sc = scenes.dam_custom(16, 16, 8, seed=1)
    s = RTPS.generateSystemInstance(sc.settings)
    sp, _, _ = s.params()
    s.close()
    dev = torch.device("cuda:0")
    got = ops.collision_triangles(torch.from_numpy(pos).to(dev), torch.from_numpy(vel).to(dev),
                                  torch.from_numpy(force).to(dev), tris, sp).cpu().numpy()
    want = T.collision_triangles(pos, vel, force, tris, sp.simulation_scale, sp.boundary_distance,
                                 sp.boundary_stiffness, sp.boundary_dampening)
    d_got, d_want = got - force, want - force
    hit = np.abs(d_want[:, :3]).max(1) > 0
    print(f"{hit.sum()} of {n} particles hit, max triangle force {np.abs(d_want).max():.3e}")
    assert (500 if ntri > 250 else 0) < hit.sum() < n
    assert np.array_equal(np.abs(d_got[:, :3]).max(1) > 0, hit)
    scale = np.abs(d_want).max()
    np.testing.assert_allclose(d_got[:, :3], d_want[:, :3], rtol=0, atol=1e-4 * scale)
    np.testing.assert_array_equal(got[:, 3], force[:, 3])


def test_system_collision_stage_adds_the_triangle_forces(cuda):
    """One step with and without triangles from the same state: the sorted
    forces differ by exactly the restated triangle forces, and the fused
    step path equals update() + integrate()."""
    sc = scenes.dam_custom(16, 16, 8, seed=13)
    box = float(sc.settings["domain_max"].split()[0])
    top = float(sc.pos[:, 2].max())
    vel = sc.vel.copy()
    vel[:, 2] = 0.8                             # everything moving up into the sheet
    # a sheet just above the block, wound so its normal points down at the
    # fluid (the segment test only sees a triangle's front face)
    tris = (_terrain(10, top + 0.004, 0.002, 5) * np.float32([box, box, 1.0]))[:, ::-1, :].copy()
    a = RTPS.generateSystemInstance(sc.settings)
    b = RTPS.generateSystemInstance(sc.settings)
    c = RTPS.generateSystemInstance(sc.settings)
    for x in (a, b, c):
        x.pushParticles(sc.pos, vel)
    a.loadTriangles(tris)
    c.loadTriangles(tris)
    a.update()                                  # with triangles
    b.update()                                  # without
    fa, fb = a.read("force"), b.read("force")
    sp, _, _ = a.params()
    want = T.collision_triangles(a.read("pos_s"), a.read("vel_s"), np.zeros_like(fa), tris, sp.simulation_scale,
                                 sp.boundary_distance, sp.boundary_stiffness, sp.boundary_dampening)
    hit = np.abs(want[:, :3]).max(1) > 0
    print(f"{hit.sum()} particles touch the sheet")
    assert hit.sum() > 50
    np.testing.assert_allclose(fa[:, :3] - fb[:, :3], want[:, :3], rtol=0, atol=1e-4 * np.abs(want).max())
    a.integrate()
    c.step(1)                                   # graph path with triangles
    np.testing.assert_allclose(c.read("pos"), a.read("pos"), rtol=1e-6, atol=1e-7)
    for x in (a, b, c):
        x.close()
