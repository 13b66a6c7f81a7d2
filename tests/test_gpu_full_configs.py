"""The workloads at their full stated sizes.

* dam break, 1,048,576 particles: 10 steps against the FP32 oracle at the
  1e-4 bar of the 16k dam (positions per element, velocities and densities
  against their fields' max-norm).
* batched, 256 scenes of 4,096-65,536 particles (log-uniform, seed 1003,
  ~5.6 M particles): every scene's segmented keys and idx bit-exact with a
  plain system of the same inputs, positions within 1e-6, densities 1e-5;
  scenes 0, 128, 255 and the largest against the FP32 oracle at 1e-4.
* differentiable, 65,536 particles x 200 steps (dt 0.002, checkpoints every
  20), loss MSE(x_T, target of viscosity 0.02 / gas constant 3.5): the
  adjoint's directional derivatives <dL/dv0, u>, dL/dmu, dL/dK against the
  FP64 forward-mode oracle (oracle.TLState, the exact derivative of the same
  step with branches fixed) taken along the GPU's own trajectory, at 1e-3,
  and against the FP64 trajectory's forward-mode derivative over 20 steps
  (dL/dmu, dL/dK at 1e-3; <dL/dv0, u>, whose terms cancel 2700-fold, at
  1e-5 of the terms' magnitudes).  Over 200 steps the FP32 and FP64 rollouts part: the wall penalty's
  damping term switches on with a jump (collision_wall.cl, diff > eps) and
  the friction term flips sign with a jump, so the two rollouts cross
  different switches; central differences of the loss grow like 1 / step for
  the same reason (asserted).
* batched + differentiable, 64 scenes of 16,384-131,072 particles x 100
  steps: every scene's final state and gradients bit-identical to the scene
  run alone through the batched kernels, and over one 20-step segment within
  1e-4 (gradients) and 5e-4 (dL/dmu, dL/dK) of the single-system adjoint
  (rtps_adj).
"""
import numpy as np
import pytest

import oracle as O
from enjaparticles_amd import scenes

pytestmark = pytest.mark.gpu


def _rel_max(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-30)


def test_dam1m_ten_steps_match_fp32_oracle(cuda):
    from enjaparticles_amd.rtps import RTPS
    sc = scenes.dam1m()
    print(f"dam1m: {sc.n} particles", flush=True)
    s = RTPS.generateSystemInstance(sc.settings)
    s.pushParticles(sc.pos, sc.vel)
    sp, gp, gs = O.derive_params(sc.settings)
    st = O.State(sc.pos, sc.vel, sp, gp, gs)
    for k in range(10):
        s.step(1)
        st.step(sc.dt, O.FLAG_COLLISION)
        assert s.getNum() == st.num == sc.n
        np.testing.assert_array_equal(s.read("keys", sc.n), st.keys)   # sorted cell keys, bit-exact
        dens = np.empty(sc.n, np.float32)
        dens[s.read("idx")] = s.read("density")
        ref = np.empty(sc.n, np.float32)
        ref[st.idx] = st.density
        np.testing.assert_allclose(s.read("pos")[:, :3], st.pos_u[:, :3], rtol=1e-4, atol=0)
        assert _rel_max(s.read("vel")[:, :3], st.vel_u[:, :3]) <= 1e-4, k
        np.testing.assert_allclose(dens, ref, rtol=1e-4)
    s.close()


def test_batched_256_scenes_equal_plain_systems(cuda):
    from enjaparticles_amd.rtps import SPH, BatchedSPH
    scs = scenes.batched(256, 4096, 65536, seed=1003)
    total = sum(sc.n for sc in scs)
    big = int(np.argmax([sc.n for sc in scs]))
    print(f"256 scenes, {total} particles, largest scene {big} ({scs[big].n})", flush=True)
    assert 5.0e6 < total < 6.5e6 and max(sc.n for sc in scs) > 60000
    b = BatchedSPH([sc.settings for sc in scs])
    for s, sc in enumerate(scs):
        b.pushParticles(s, sc.pos, sc.vel)
    steps = 5
    b.step(steps)
    oracle_scenes = {0, 128, 255, big}
    for s, sc in enumerate(scs):
        p = SPH(sc.settings)
        p.pushParticles(sc.pos, sc.vel)
        p.step(steps)
        assert b.getNum(s) == p.getNum() == sc.n
        kp = p.read("keys")
        np.testing.assert_array_equal(b.read(s, "keys"), (np.uint32(s) << np.uint32(24)) | kp)
        np.testing.assert_array_equal(b.read(s, "idx"), p.read("idx"))
        pb = b.read(s, "pos")
        np.testing.assert_allclose(pb[:, :3], p.read("pos")[:, :3], rtol=1e-6, atol=1e-7)
        # the batched and plain kernels are compiled separately: single roundings
        # differ and compound over the steps (3.9e-6 seen on a 62,859-particle scene)
        np.testing.assert_allclose(b.read(s, "density"), p.read("density"), rtol=1e-5)
        p.close()
        if s in oracle_scenes:
            sp, gp, gs = O.derive_params(sc.settings)
            st = O.State(sc.pos, sc.vel, sp, gp, gs)
            for _ in range(steps):
                st.step(sc.dt, O.FLAG_COLLISION)
            np.testing.assert_allclose(pb[:, :3], st.pos_u[:, :3], rtol=1e-4, atol=0)
            assert _rel_max(b.read(s, "vel")[:, :3], st.vel_u[:, :3]) <= 1e-4
            print(f"scene {s}: {sc.n} particles match the oracle", flush=True)
    b.close()


MU_T, K_T = 0.02, 3.5


def _cfg4_reference(sc, steps, u):
    """FP64 target (viscosity 0.02, gas constant 3.5) and the forward-mode
    run along (u, d mu, d K): returns target, x_T (FP64) and the three
    tangents of x_T."""
    sp, gp, gs = O.derive_params(sc.settings)
    dt = float(np.float32(sc.dt))
    spt, _, _ = O.derive_params(sc.settings)
    spt.viscosity, spt.K = MU_T, K_T
    t = O.State(sc.pos.astype(np.float64), sc.vel.astype(np.float64), spt, gp, gs, dtype=np.float64)
    for _ in range(steps):
        t.step(dt, O.FLAG_COLLISION)
    assert t.num == sc.n
    tl = O.TLState(sc.pos.astype(np.float64), sc.vel.astype(np.float64), sp, gp, gs, dvel0=u, dmu=[0, 1, 0],
                   dK=[0, 0, 1])
    for _ in range(steps):
        tl.step(dt, O.FLAG_COLLISION)
    assert tl.num == sc.n
    return t.pos_u[:, :3].copy(), tl.pos_u[:, :3, 0].copy(), [tl.pos_u[:, :3, 1 + k].copy() for k in range(3)]


def _adjoint_dot_products(sc, steps, ck, w, u):
    """GPU adjoint seeded with w = dL/dx_T: <dL/dv0, u>, dL/dmu, dL/dK."""
    import torch
    from enjaparticles_amd.adjoint import AdjointSPH
    dev = torch.device("cuda:0")
    sp, _, _ = O.derive_params(sc.settings)
    sim = AdjointSPH(sc.settings, sc.n, checkpoint_every=ck, max_steps=steps)
    sim.forward(torch.from_numpy(sc.pos).to(dev), torch.from_numpy(sc.vel).to(dev), steps, sp.viscosity, sp.K)
    g = np.zeros((sc.n, 4), np.float32)
    g[:, :3] = w
    dv0, _, dmu, dK = sim.backward(torch.from_numpy(g).to(dev))
    sim.close()
    dv0 = dv0.cpu().numpy()[:, :3].astype(np.float64)
    return [float(np.sum(dv0 * u[0][:, :3])), dmu, dK]


def _gpu_trajectory(sim, p0, v0, steps, mu, gas):
    """States (pos, vel, veleval) after k = 0..steps steps of the adjoint
    handle's own forward (each k replayed from the start: the same kernels
    on the same inputs give the same states), in original order, FP32."""
    P, V = [p0.cpu().numpy()], [v0.cpu().numpy()]
    for k in range(1, steps + 1):
        pk, vk = sim.forward(p0, v0, k, mu, gas)
        P.append(pk.cpu().numpy())
        V.append(vk.cpu().numpy())
    E = [np.zeros_like(P[0])] + [np.float32(0.5) * (V[k - 1] + V[k]) for k in range(1, steps + 1)]
    return P, V, E


def _shadow_tangents(sc, traj, steps, u):
    """Forward-mode derivative along the GPU's own trajectory: before every
    step the tangent state's values are set to the GPU state, so each step's
    Jacobian is taken where the GPU linearises (the same wall contacts,
    friction signs and neighbours), and only the tangents are carried."""
    P, V, E = traj
    sp, gp, gs = O.derive_params(sc.settings)
    dt = float(np.float32(sc.dt))
    tl = O.TLState(P[0].astype(np.float64), V[0].astype(np.float64), sp, gp, gs, dvel0=u, dmu=[0, 1, 0],
                   dK=[0, 0, 1])
    for k in range(steps):
        tl.pos_u[:, :, 0] = P[k]
        tl.vel_u[:, :, 0] = V[k]
        tl.veleval_u[:, :, 0] = E[k]
        tl.step(dt, O.FLAG_COLLISION)
        assert tl.num == sc.n
    return [tl.pos_u[:, :3, 1 + k].copy() for k in range(3)]


def test_adjoint_65536_particles_200_steps_matches_forward_mode(cuda):
    """The adjoint's <dL/dv0, u>, dL/dmu, dL/dK on the full rollout against
    the forward-mode derivative taken along the same (GPU) trajectory, at
    1e-3; the FP64 trajectory's own forward-mode derivative agrees at 20
    steps and drifts away by 200 (printed): the FP32 and FP64 rollouts take
    different wall-contact and friction branches, across which the loss
    jumps (the central differences below grow like 1 / step)."""
    import torch
    from enjaparticles_amd.adjoint import AdjointSPH
    sc = scenes.adjoint64k()
    steps, ck = 200, 20
    dev = torch.device("cuda:0")
    print(f"adjoint: {sc.n} particles, {steps} steps, checkpoints every {ck}", flush=True)
    rng = np.random.default_rng(3)
    u = np.zeros((3, sc.n, 4))
    d = rng.normal(size=(sc.n, 3))
    u[0, :, :3] = d / np.sqrt(np.mean(d ** 2))
    sp, gp, gs = O.derive_params(sc.settings)
    p0 = torch.from_numpy(sc.pos).to(dev)
    v0 = torch.from_numpy(sc.vel).to(dev)
    sim = AdjointSPH(sc.settings, sc.n, checkpoint_every=ck, max_steps=steps)
    traj = _gpu_trajectory(sim, p0, v0, steps, sp.viscosity, sp.K)
    print("gpu trajectory recorded", flush=True)
    target, xT64, J64 = _cfg4_reference(sc, steps, u)
    # one seed for both sides: dL/dx_T of L = mean((x_T - target)^2) at the GPU's x_T
    w = 2.0 * (traj[0][steps][:, :3].astype(np.float64) - target) / (3 * sc.n)
    sim.forward(p0, v0, steps, sp.viscosity, sp.K)
    g = np.zeros((sc.n, 4), np.float32)
    g[:, :3] = w
    dv0, _, dmu, dK = sim.backward(torch.from_numpy(g).to(dev))
    sim.close()
    dv0 = dv0.cpu().numpy()[:, :3].astype(np.float64)
    ad = [float(np.sum(dv0 * u[0][:, :3])), dmu, dK]
    Js = _shadow_tangents(sc, traj, steps, u)
    shadow = [float(np.sum(w * Js[k])) for k in range(3)]
    fp64 = [float(np.sum(w * J64[k])) for k in range(3)]
    for name, a, t, f in zip(("dL/dv0 . u", "dL/dmu", "dL/dK"), ad, shadow, fp64):
        print(f"{name}: adjoint {a:.8e}  forward-mode on the GPU trajectory {t:.8e} (rel {abs(a - t) / abs(t):.2e})"
              f"  forward-mode on the FP64 trajectory {f:.8e}", flush=True)
    for a, t in zip(ad, shadow):
        assert abs(a - t) <= 1e-3 * abs(t)
    # Central differences along u do not converge: the quotient grows like 1 / step.
    dt = float(np.float32(sc.dt))

    def loss(v):
        s = O.State(sc.pos.astype(np.float64), v, sp, gp, gs, dtype=np.float64)
        for _ in range(steps):
            s.step(dt, O.FLAG_COLLISION)
        return float(np.mean((s.pos_u[:, :3] - target) ** 2))

    v = sc.vel.astype(np.float64)
    fd = {h: (loss(v + h * u[0]) - loss(v - h * u[0])) / (2 * h) for h in (1e-4, 1e-5)}
    tl64 = float(np.sum(2.0 * (xT64 - target) / (3 * sc.n) * J64[0]))
    print(f"central differences along u (FP64 loss): step 1e-4 {fd[1e-4]:.6e}, step 1e-5 {fd[1e-5]:.6e}, "
          f"forward-mode {tl64:.6e}", flush=True)
    assert abs(fd[1e-5] - tl64) > 3.0 * abs(fd[1e-4] - tl64)


def test_adjoint_65536_particles_20_steps_matches_fp64_forward_mode(cuda):
    """The same workload over 20 steps (one checkpoint segment), before the
    FP32 and FP64 rollouts part: the adjoint against the forward-mode
    derivative of the FP64 trajectory itself.  dL/dmu and dL/dK at 1e-3
    relative.  <dL/dv0, u> is a sum of per-particle terms that cancel about
    2700-fold (sum |w J u| = 1.07e-6 against the sum, -3.94e-10), so the
    FP32 rollout's own rounding moves it by far more than 1e-3 of itself
    (2.1e-3 measured); it is held to 1e-5 of the sum of the terms' magnitudes
    (7.7e-7 measured), which a dropped or mis-signed term would break."""
    import torch
    from enjaparticles_amd.adjoint import AdjointSPH
    sc = scenes.adjoint64k()
    steps, ck = 20, 20
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(4)
    u = np.zeros((3, sc.n, 4))
    d = rng.normal(size=(sc.n, 3))
    u[0, :, :3] = d / np.sqrt(np.mean(d ** 2))
    target, xT, J = _cfg4_reference(sc, steps, u)
    w = 2.0 * (xT - target) / (3 * sc.n)
    tl = [float(np.sum(w * J[k])) for k in range(3)]
    mag = [float(np.sum(np.abs(w * J[k]))) for k in range(3)]
    ad = _adjoint_dot_products(sc, steps, ck, w, u)
    for name, a, t, m in zip(("dL/dv0 . u", "dL/dmu", "dL/dK"), ad, tl, mag):
        print(f"{name}: adjoint {a:.8e}  forward-mode {t:.8e}  rel {abs(a - t) / abs(t):.2e}  "
              f"(terms' magnitudes {m:.3e}, difference {abs(a - t) / m:.2e} of them)", flush=True)
    assert abs(ad[0] - tl[0]) <= 1e-5 * mag[0]
    assert mag[0] > 1e3 * abs(tl[0])   # the cancellation that the bar above accounts for
    for a, t in zip(ad[1:], tl[1:]):
        assert abs(a - t) <= 1e-3 * abs(t)


def test_adjoint_with_walls_matches_fd_and_forward_mode(cuda):
    """Contact from step 0 (a block on the floor and two walls, 20 steps):
    with the same FP64 seed w, the adjoint matches FP64 central differences
    (step 1e-4, which converge here) and the forward-mode oracle at 1e-3."""
    pos = scenes.lattice(10, 10, 6, 23, origin=(0.0, 0.0, 0.0))
    n = len(pos)
    vel = np.zeros_like(pos)
    vel[:, :3] = scenes.normal(24, 3 * n, 0.05).reshape(n, 3)
    st = scenes.dam_settings(0.5, n, dt=0.002)
    sc = scenes.Scene("walls", st, pos, vel, 20, 0.002)
    steps, ck = 20, 5
    rng = np.random.default_rng(3)
    u = np.zeros((3, n, 4))
    d = rng.normal(size=(n, 3))
    u[0, :, :3] = d / np.sqrt(np.mean(d ** 2))
    target, xT, J = _cfg4_reference(sc, steps, u)
    w = 2.0 * (xT - target) / (3 * n)
    tl = [float(np.sum(w * J[k])) for k in range(3)]
    ad = _adjoint_dot_products(sc, steps, ck, w, u)
    sp, gp, gs = O.derive_params(st)
    dt = float(np.float32(sc.dt))

    def loss(v, mu, K):
        p2, _, _ = O.derive_params(st)
        p2.viscosity, p2.K = mu, K
        s = O.State(pos.astype(np.float64), v, p2, gp, gs, dtype=np.float64)
        for _ in range(steps):
            s.step(dt, O.FLAG_COLLISION)
        return float(np.mean((s.pos_u[:, :3] - target) ** 2))

    v = vel.astype(np.float64)
    h = 1e-4
    fd = [(loss(v + h * u[0], sp.viscosity, sp.K) - loss(v - h * u[0], sp.viscosity, sp.K)) / (2 * h)]
    for which in ("mu", "K"):
        base = sp.viscosity if which == "mu" else sp.K
        hi, lo = float(np.float32(base * (1 + h))), float(np.float32(base * (1 - h)))
        if which == "mu":
            fd.append((loss(v, hi, sp.K) - loss(v, lo, sp.K)) / (hi - lo))
        else:
            fd.append((loss(v, sp.viscosity, hi) - loss(v, sp.viscosity, lo)) / (hi - lo))
    for name, a, t, f in zip(("dL/dv0 . u", "dL/dmu", "dL/dK"), ad, tl, fd):
        print(f"{name}: adjoint {a:.8e}  forward-mode {t:.8e}  fd {f:.8e}", flush=True)
        assert abs(a - t) <= 1e-3 * abs(t)
        assert abs(a - f) <= 1e-3 * abs(f)


def test_batched_adjoint_64_scenes_100_steps_equal_single_adjoints(cuda):
    """One batched forward + backward of all 64 scenes equals each scene run
    alone through the same batched machinery (one-scene batches: the same
    kernels, so the per-scene arithmetic must not depend on the other
    scenes, the segmented sort or the shared work queues): bit for bit."""
    import torch
    from enjaparticles_amd.adjoint import BatchedAdjointSPH
    dev = torch.device("cuda:0")
    steps, ck = 100, 20
    scs = scenes.batched(64, 16384, 131072, seed=1005)
    total = sum(sc.n for sc in scs)
    print(f"batched adjoint: 64 scenes, {total} particles, {steps} steps", flush=True)
    vel = np.zeros((total, 4), np.float32)
    vel[:, :3] = scenes.normal(1006, 3 * total, 0.05).reshape(total, 3)
    mus = [sc.meta["viscosity"] for sc in scs]
    gases = [sc.meta["gas_constant"] for sc in scs]
    b = BatchedAdjointSPH([sc.settings for sc in scs], checkpoint_every=ck, max_steps=steps)
    assert b.num == total
    pos0 = torch.from_numpy(np.concatenate([sc.pos for sc in scs])).to(dev)
    xT, vT = b.forward(pos0, torch.from_numpy(vel).to(dev), steps, mus, gases)
    g = torch.from_numpy(np.random.default_rng(8).normal(size=(total, 4)).astype(np.float32)).to(dev)
    g[:, 3] = 0
    dv0, dx0, dmu, dK = b.backward(g)
    xT, vT, dv0, gh = (t.cpu().numpy() for t in (xT, vT, dv0, g))
    offsets = b.offsets
    b.close()
    for s, sc in enumerate(scs):
        lo, hi = offsets[s], offsets[s + 1]
        a = BatchedAdjointSPH([sc.settings], checkpoint_every=ck, max_steps=steps)
        x1, v1 = a.forward(torch.from_numpy(sc.pos).to(dev), torch.from_numpy(vel[lo:hi]).to(dev), steps, [mus[s]],
                           [gases[s]])
        d1, _, m1, k1 = a.backward(torch.from_numpy(gh[lo:hi]).to(dev))
        a.close()
        x1, v1, d1 = (t.cpu().numpy() for t in (x1, v1, d1))
        np.testing.assert_array_equal(xT[lo:hi], x1, err_msg=f"scene {s}")
        np.testing.assert_array_equal(vT[lo:hi], v1, err_msg=f"scene {s}")
        np.testing.assert_array_equal(dv0[lo:hi], d1, err_msg=f"scene {s}")
        assert dmu[s] == m1[0] and dK[s] == k1[0], s
    print("all 64 scenes bit-identical to their one-scene runs", flush=True)


def test_batched_adjoint_64_scenes_equal_plain_adjoints_over_a_segment(cuda):
    """The 64-scene batch against the single-system adjoint (rtps_adj, its own
    plain kernels) scene by scene over one checkpoint segment (20 steps):
    final state within 1e-6, per-particle gradients within 1e-4 of the
    field's max, dL/dmu and dL/dK within 5e-4 relative (the two compile the
    same pair terms separately; their roundings differ and grow with the
    steps; the largest measured over the 64 scenes is 9.5e-5 for dL/dmu,
    a sum over all pairs whose terms partly cancel)."""
    import torch
    from enjaparticles_amd.adjoint import AdjointSPH, BatchedAdjointSPH
    dev = torch.device("cuda:0")
    steps, ck = 20, 20
    scs = scenes.batched(64, 16384, 131072, seed=1005)
    total = sum(sc.n for sc in scs)
    vel = np.zeros((total, 4), np.float32)
    vel[:, :3] = scenes.normal(1006, 3 * total, 0.05).reshape(total, 3)
    mus = [sc.meta["viscosity"] for sc in scs]
    gases = [sc.meta["gas_constant"] for sc in scs]
    b = BatchedAdjointSPH([sc.settings for sc in scs], checkpoint_every=ck, max_steps=steps)
    pos0 = torch.from_numpy(np.concatenate([sc.pos for sc in scs])).to(dev)
    xT, vT = b.forward(pos0, torch.from_numpy(vel).to(dev), steps, mus, gases)
    g = torch.from_numpy(np.random.default_rng(9).normal(size=(total, 4)).astype(np.float32)).to(dev)
    g[:, 3] = 0
    dv0, _, dmu, dK = b.backward(g)
    xT, vT, dv0, gh = (t.cpu().numpy() for t in (xT, vT, dv0, g))
    offsets = b.offsets
    b.close()
    worst = [0.0, 0.0, 0.0]
    for s, sc in enumerate(scs):
        lo, hi = offsets[s], offsets[s + 1]
        a = AdjointSPH(sc.settings, sc.n, checkpoint_every=ck, max_steps=steps)
        x1, v1 = a.forward(torch.from_numpy(sc.pos).to(dev), torch.from_numpy(vel[lo:hi]).to(dev), steps, mus[s],
                           gases[s])
        d1, _, m1, k1 = a.backward(torch.from_numpy(gh[lo:hi]).to(dev))
        a.close()
        x1, v1, d1 = (t.cpu().numpy() for t in (x1, v1, d1))
        np.testing.assert_allclose(xT[lo:hi], x1, rtol=1e-6, atol=1e-6, err_msg=f"scene {s}")
        r = [_rel_max(dv0[lo:hi, :3], d1[:, :3]), abs(dmu[s] - m1) / abs(m1), abs(dK[s] - k1) / abs(k1)]
        worst = [max(a_, b_) for a_, b_ in zip(worst, r)]
        assert r[0] <= 1e-4 and r[1] <= 5e-4 and r[2] <= 5e-4, (s, r)
    print(f"largest differences over the 64 scenes: dv0 {worst[0]:.2e}, dmu {worst[1]:.2e}, dK {worst[2]:.2e}",
          flush=True)
