"""Adjoint (rtps_adj_*): gradients of L = MSE(x_T, x_T^target) with respect to
the initial velocities, the viscosity and the gas constant, checked against
FP64 central finite differences of the CPU oracle.

The target is the same simulation with viscosity 0.02 and gas constant 3.5
from the identical initial state.  The oracle's parameter block is FP32, so
the scalar steps are +-1e-4 relative rounded to FP32 and the difference
quotient divides by the FP32 values actually used; velocities are perturbed
along one random unit-RMS direction by +-1e-4.
"""
import numpy as np
import pytest

import oracle as O
from enjaparticles_amd import scenes

pytestmark = pytest.mark.gpu

MU, K = 0.01, 3.0
MU_T, K_T = 0.02, 3.5


def _scene(nx, ny, nz, origin, box, seed, dt=0.002):
    pos = scenes.lattice(nx, ny, nz, seed, origin=origin)
    n = len(pos)
    vel = np.zeros_like(pos)
    vel[:, :3] = scenes.normal(seed + 1, 3 * n, 0.05).reshape(n, 3)
    st = scenes.dam_settings(box, n, gas_constant=K, viscosity=MU, dt=dt)
    return st, pos, vel


def _oracle_final(st, pos, vel, steps, mu, gas, collision):
    sp, gp, gs = O.derive_params(st)
    sp.viscosity = float(np.float32(mu))
    sp.K = float(np.float32(gas))
    s = O.State(pos.astype(np.float64), vel.astype(np.float64), sp, gp, gs, dtype=np.float64)
    dt = float(np.float32(st["time_step"]))
    for _ in range(steps):
        s.step(dt, O.FLAG_COLLISION if collision else 0)
    assert s.num == len(pos)
    return s.pos_u[:, :3].copy()


def _loss(x, t):
    return float(np.mean((x - t) ** 2))


def _check(st, pos, vel, steps, ck, collision, tol):
    import torch
    from enjaparticles_amd.adjoint import AdjointSPH, sph_rollout
    n = len(pos)
    target = _oracle_final(st, pos, vel, steps, MU_T, K_T, collision)
    dev = torch.device("cuda:0")
    sim = AdjointSPH(st, n, checkpoint_every=ck, max_steps=steps)
    p0 = torch.from_numpy(pos).to(dev)
    v0 = torch.from_numpy(vel).to(dev).requires_grad_(True)
    mu = torch.tensor(MU, dtype=torch.float64, requires_grad=True)
    gas = torch.tensor(K, dtype=torch.float64, requires_grad=True)
    xT = sph_rollout(sim, p0, v0, mu, gas, steps, collision)
    loss = torch.nn.functional.mse_loss(xT[:, :3], torch.from_numpy(target.astype(np.float32)).to(dev))
    loss.backward()
    g_v0 = v0.grad.cpu().numpy()[:, :3].astype(np.float64)
    g_mu, g_K = float(mu.grad), float(gas.grad)
    sim.close()

    def fd_scalar(which):
        base = MU if which == "mu" else K
        hi, lo = np.float32(base * (1 + 1e-4)), np.float32(base * (1 - 1e-4))
        args = lambda val: (MU if which != "mu" else val, K if which != "K" else val)  # noqa: E731
        lp = _loss(_oracle_final(st, pos, vel, steps, *args(hi), collision), target)
        lm = _loss(_oracle_final(st, pos, vel, steps, *args(lo), collision), target)
        return (lp - lm) / (float(hi) - float(lo))

    rng = np.random.default_rng(3)
    u = rng.normal(size=(n, 3))
    u /= np.sqrt(np.mean(u ** 2))
    d = 1e-4
    vp, vm = vel.astype(np.float64).copy(), vel.astype(np.float64).copy()
    vp[:, :3] += d * u
    vm[:, :3] -= d * u
    fd_v = (_loss(_oracle_final(st, pos, vp, steps, MU, K, collision), target)
            - _loss(_oracle_final(st, pos, vm, steps, MU, K, collision), target)) / (2 * d)
    ad_v = float(np.sum(g_v0 * u))
    fd_mu, fd_K = fd_scalar("mu"), fd_scalar("K")
    print(f"dL/dmu adj {g_mu:.6e} fd {fd_mu:.6e} | dL/dK adj {g_K:.6e} fd {fd_K:.6e} | "
          f"dL/dv0.u adj {ad_v:.6e} fd {fd_v:.6e}")
    for a, f in ((g_mu, fd_mu), (g_K, fd_K), (ad_v, fd_v)):
        assert abs(a - f) <= tol * abs(f), (a, f)


def test_adjoint_matches_fd_free_block(cuda):
    """A block falling freely in the middle of the box (no wall contact): 40 steps."""
    st, pos, vel = _scene(16, 16, 8, origin=(0.34, 0.34, 0.45), box=1.0, seed=21)
    _check(st, pos, vel, steps=40, ck=10, collision=True, tol=1e-3)


def test_adjoint_matches_fd_with_walls(cuda):
    """A small dam break resting on the floor and two walls (contact from step 0)."""
    st, pos, vel = _scene(10, 10, 6, origin=(0.0, 0.0, 0.0), box=0.5, seed=23)
    _check(st, pos, vel, steps=20, ck=5, collision=True, tol=1e-2)


def test_adjoint_forward_equals_plain_system(cuda):
    import torch
    from enjaparticles_amd.adjoint import AdjointSPH
    from enjaparticles_amd.rtps import SPH
    st, pos, vel = _scene(12, 12, 6, origin=(0.0, 0.0, 0.0), box=0.6, seed=25)
    n = len(pos)
    sim = AdjointSPH(st, n, checkpoint_every=4, max_steps=12)
    dev = torch.device("cuda:0")
    xT, vT = sim.forward(torch.from_numpy(pos).to(dev), torch.from_numpy(vel).to(dev), 12, MU, K)
    p = SPH(st)
    p.pushParticles(pos, vel)
    for _ in range(12):
        p.update()
        p.integrate()
    np.testing.assert_allclose(xT.cpu().numpy(), p.read("pos"), rtol=1e-6, atol=1e-7)
    np.testing.assert_allclose(vT.cpu().numpy(), p.read("vel"), rtol=1e-5, atol=1e-6)
    p.close()
    sim.close()
