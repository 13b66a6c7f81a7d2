"""Whole-step parity of the SPH system (rtps_sph_* C ABI) with the CPU oracle.

Bars (written per test): over the first 10 steps of the 16k dam-break the FP32
positions, velocities and densities stay within 1e-4 relative of the FP32
oracle; the FP32 HIP run stays within a stated tolerance of the FP64 oracle;
the fused/graph path equals the update()+integrate() path; the out-of-grid
deletion path shrinks the count exactly as the oracle does.
"""
import numpy as np
import pytest

import oracle as O
from enjaparticles_amd import scenes
from enjaparticles_amd._native import DEFAULT_FORCE_LOOP, RTPSError
from enjaparticles_amd.rtps import RTPS, SPH

pytestmark = pytest.mark.gpu


def _system(sc, collision=True, graph=True):
    s = RTPS.generateSystemInstance(sc.settings)
    s.setOption(0, int(collision))
    s.setOption(1, int(graph))
    s.pushParticles(sc.pos, sc.vel)
    return s


def _oracle(sc, dtype=np.float32):
    sp, gp, gs = O.derive_params(sc.settings)
    return O.State(sc.pos, sc.vel, sp, gp, gs, dtype=dtype)


def _rel_max(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-30)


def test_unknown_system_returns_none(cuda):
    assert RTPS.generateSystemInstance({"system": "flock"}) is None


def test_ten_steps_match_fp32_oracle(cuda):
    sc = scenes.dam16k()
    s = _system(sc)
    st = _oracle(sc)
    for k in range(10):
        s.step(1)
        st.step(sc.dt, O.FLAG_COLLISION)
        pos, vel = s.read("pos"), s.read("vel")
        # densities are in sorted order and the within-cell order of the two
        # sorts differs: compare by original index
        dens = np.empty(sc.n, np.float32)
        dens[s.read("idx")] = s.read("density")
        ref = np.empty(sc.n, np.float32)
        ref[st.idx[: st.num]] = st.density[: st.num]
        # per-element 1e-4 relative on positions; velocities and densities
        # max-norm 1e-4 relative of their fields
        np.testing.assert_allclose(pos[:, :3], st.pos_u[:, :3], rtol=1e-4, atol=0)
        assert _rel_max(vel[:, :3], st.vel_u[:, :3]) <= 1e-4, k
        np.testing.assert_allclose(dens, ref, rtol=1e-4)
    assert s.getNum() == st.num == sc.n
    s.close()


@pytest.mark.parametrize("n", [1, 2, 63, 64, 65, 127])
def test_tiny_systems_match_fp32_oracle(cuda, n):
    """A single isolated particle (free fall, self-term density), a pair, and
    counts around one 64-lane wave (last wave partly empty or one lane
    alone): 10 steps within the same 1e-4 bars as the 16k dam."""
    full = scenes.dam16k(seed=1009)
    sc = scenes.Scene("tiny", full.settings, full.pos[:n].copy(), np.zeros((n, 4), np.float32), 10, 0.003)
    s = _system(sc)
    st = _oracle(sc)
    for k in range(10):
        s.step(1)
        st.step(sc.dt, O.FLAG_COLLISION)
        assert s.getNum() == st.num == n
        dens = np.empty(n, np.float32)
        dens[s.read("idx")] = s.read("density")
        ref = np.empty(n, np.float32)
        ref[st.idx[: st.num]] = st.density[: st.num]
        np.testing.assert_allclose(s.read("pos")[:, :3], st.pos_u[:, :3], rtol=1e-4, atol=0)
        assert _rel_max(s.read("vel")[:, :3], st.vel_u[:, :3]) <= 1e-4, k
        np.testing.assert_allclose(dens, ref, rtol=1e-4)
    if n == 1:   # alone: the density is the self term m * poly6(0) = m * 315 / (64 pi h^3)
        sp, _, _ = O.derive_params(sc.settings)
        h = np.float64(sp.smoothing_distance)
        assert dens[0] == pytest.approx(sp.mass * 315.0 / (64.0 * np.pi * h ** 3), rel=1e-5)
    s.close()


def test_empty_system_steps(cuda):
    """No particles pushed: steps, reads and a later push all work."""
    sc = scenes.dam16k(seed=1011)
    s = RTPS.generateSystemInstance(sc.settings)
    s.step(3)
    assert s.getNum() == 0
    assert s.read("pos").shape[0] == 0
    s.pushParticles(sc.pos[:100], sc.vel[:100])
    s.step(2)
    assert s.getNum() == 100
    assert np.isfinite(s.read("pos")).all()
    s.close()


def test_reference_derived_parameters_match_fp32_oracle(cuda):
    """Settings without the workload's overrides: mass, rest distance, h and
    the simulation scale are derived as SPHSettings.cpp:45-75 does (scale
    0.31, not 1), so every neighbour kernel takes its scaled-position path;
    a lattice at the derived rest spacing (in world units) steps like the FP32
    oracle for 5 steps."""
    st = scenes.dam_settings(1.0, 16384)
    for k in ("mass", "rest_distance", "smoothing_distance", "simulation_scale"):
        del st[k]
    sp, gp, gs = O.derive_params(st)
    assert abs(sp.simulation_scale - 1.0) > 0.5
    pos = scenes.lattice(24, 24, 12, seed=86, spacing=sp.rest_distance / sp.simulation_scale)
    sc = scenes.Scene("derived", st, pos, np.zeros_like(pos), 5, 0.003)
    s = _system(sc)
    ref = _oracle(sc)
    for k in range(5):
        s.step(1)
        ref.step(sc.dt, O.FLAG_COLLISION)
        dens = np.empty(sc.n, np.float32)
        dens[s.read("idx")] = s.read("density")
        rd = np.empty(sc.n, np.float32)
        rd[ref.idx[: ref.num]] = ref.density[: ref.num]
        np.testing.assert_allclose(s.read("pos")[:, :3], ref.pos_u[:, :3], rtol=1e-4, atol=0)
        assert _rel_max(s.read("vel")[:, :3], ref.vel_u[:, :3]) <= 1e-4, k
        np.testing.assert_allclose(dens, rd, rtol=1e-4)
    assert s.getNum() == ref.num == sc.n
    s.close()


def test_hundred_steps_vs_fp64_oracle(cuda):
    """FP32 HIP vs FP64 oracle after 100 steps.  Pair sums are chaotic, so FP32
    rounding grows into visible position differences; the bar is the FP32
    oracle's own distance from FP64 (the reference's FP32 arithmetic): the HIP
    run may be at most 3x as far (max and RMS over particles) plus 1e-4."""
    sc = scenes.dam16k()
    s = _system(sc)
    st64 = _oracle(sc, np.float64)
    st32 = _oracle(sc, np.float32)
    s.step(100)
    for _ in range(100):
        st64.step(sc.dt, O.FLAG_COLLISION)
        st32.step(sc.dt, O.FLAG_COLLISION)
    pos = s.read("pos")
    e_gpu = np.linalg.norm(pos[:, :3] - st64.pos_u[:, :3], axis=1)
    e_ref = np.linalg.norm(st32.pos_u[:, :3] - st64.pos_u[:, :3], axis=1)
    print(f"100 steps: max/rms |HIP - FP64| = {e_gpu.max():.3e}/{np.sqrt((e_gpu**2).mean()):.3e}, "
          f"|FP32 oracle - FP64| = {e_ref.max():.3e}/{np.sqrt((e_ref**2).mean()):.3e}")
    assert e_gpu.max() <= 3 * e_ref.max() + 1e-4
    assert np.sqrt((e_gpu ** 2).mean()) <= 3 * np.sqrt((e_ref ** 2).mean()) + 1e-4
    s.close()


def test_fused_graph_equals_update_integrate(cuda):
    sc = scenes.dam_custom(24, 24, 12, seed=77)
    a = _system(sc, graph=True)
    b = _system(sc, graph=False)
    c = _system(sc)
    a.step(5)
    b.step(5)
    for _ in range(5):
        c.update()
        c.interact()
        c.integrate()
    pa, pb, pc = a.read("pos"), b.read("pos"), c.read("pos")
    np.testing.assert_array_equal(pa, pb)
    np.testing.assert_allclose(pc, pa, rtol=1e-6, atol=1e-7)
    for s in (a, b, c):
        s.close()


def test_deletion_path_matches_oracle(cuda):
    """Particles pushed out of the grid are deleted (SPH.cpp:220-250) on the
    device, with the same survivors as the oracle."""
    sc = scenes.dam16k()
    pos = sc.pos.copy()
    pos[5::211, 2] = 70.0    # far above the grid: key nb_cells
    sc2 = scenes.Scene("del", sc.settings, pos, sc.vel, 3, sc.dt)
    s = _system(sc2)
    st = _oracle(sc2)
    for _ in range(3):
        s.step(1)
        st.step(sc.dt, O.FLAG_COLLISION)
    n = s.getNum()
    assert n == st.num == sc.n - len(pos[5::211])
    # The deletion path renumbers the survivors in sorted order, and the
    # within-cell order of the two sorts differs: match particles by position
    # (a bijection, every pair within 1e-5).
    from scipy.spatial import cKDTree
    a = s.read("pos")[:, :3].astype(np.float64)
    b = st.pos_u[:n, :3].astype(np.float64)
    d, j = cKDTree(b).query(a)
    assert d.max() <= 1e-5
    assert len(np.unique(j)) == n
    s.close()


def test_shipped_order_without_walls(cuda):
    """Collision off (the shipped SPH.cpp:297): still matches the oracle."""
    sc = scenes.dam16k()
    s = _system(sc, collision=False)
    st = _oracle(sc)
    for _ in range(5):
        s.step(1)
        st.step(sc.dt, 0)
    np.testing.assert_allclose(s.read("pos")[:, :3], st.pos_u[:, :3], rtol=1e-4)
    s.close()


def test_push_capacity_and_settings_change(cuda):
    sc = scenes.dam_custom(16, 16, 8, seed=3)
    s = RTPS.generateSystemInstance(sc.settings)
    s.pushParticles(sc.pos[:1000])
    s.pushParticles(sc.pos[1000:], sc.vel[1000:])
    assert s.getNum() == sc.n
    from enjaparticles_amd._native import RTPSError
    with pytest.raises(RTPSError):
        s.pushParticles(sc.pos[:1])
    s.step(2)
    s.setSetting("viscosity", "0.05")
    st = _oracle(sc)
    st.step(sc.dt, 1)
    st.step(sc.dt, 1)
    st.sphp, _, _ = O.derive_params({**sc.settings, "viscosity": "0.05"})
    s.step(2)
    st.step(sc.dt, 1)
    st.step(sc.dt, 1)
    np.testing.assert_allclose(s.read("pos")[:, :3], st.pos_u[:, :3], rtol=1e-4)
    s.close()


@pytest.mark.parametrize("nn,lo,hi", [(1000, (0.1, 0.1, 0.1), (0.3, 0.3, 0.3)),
                                      (100000, (0.05, 0.2, 0.3), (0.33, 0.41, 0.5)),
                                      (37, (0.1, 0.2, 0.3), (0.33, 0.41, 0.5))])
def test_addbox_matches_addrect(cuda, nn, lo, hi):
    """System::addBox (System.cpp:331-338) pushes exactly addRect's lattice
    (IV.cpp:165-202, restated in the oracle): every particle bit for bit, the
    count capped at nn, zero velocity and veleval, default colour."""
    sc = scenes.dam16k()
    s = SPH(sc.settings)
    n = s.addBox(nn, lo, hi, spacing=0.02)
    ref = O.add_rect(nn, lo, hi, 0.02)
    p = s.read("pos")
    assert n == len(p) == len(ref) == s.getNum()
    np.testing.assert_array_equal(p, ref)
    assert not np.any(s.read("vel")) and not np.any(s.read("veleval"))
    np.testing.assert_array_equal(s.read("color"), np.tile(np.float32([1, 0, 0, 1]), (n, 1)))
    s.close()


def test_stage_timers_do_not_change_the_step(cuda):
    """Per-stage timers (events inside the step graph, two alternating sets)
    time every step and leave the results bit-identical; the force-only mode
    (RTPS_OPT_TIMERS 2) times just the force stage, and mode 3 the force stage
    on every fourth step (graph and eager launches).  The hash runs inside the
    sort's first kernel and the permute inside the cellindices kernel, so
    those two stages are empty."""
    sc = scenes.dam_custom(24, 24, 12, seed=78)
    a = _system(sc, graph=True)
    b = _system(sc, graph=True)
    b.setOption(3, 1)
    c = _system(sc, graph=False)
    c.setOption(3, 1)
    d = _system(sc, graph=True)
    d.setOption(3, 2)
    e = _system(sc, graph=True)
    e.setOption(3, 3)
    f = _system(sc, graph=False)
    f.setOption(3, 3)
    for s in (a, b, c, d, e, f):
        s.step(1)          # the first step tunes the launches (not timed)
        s.resetTimers()
        s.step(7)
    for s in (b, c, d, e, f):
        np.testing.assert_array_equal(s.read("pos"), a.read("pos"))
    for s in (e, f):       # sampled force timers: steps 0 and 4 of the 7
        st = s.stageTimes()
        assert st["force"][1] == 2 and st["force"][0] > 0.0, st
        assert all(cnt == 0 for k, (_, cnt) in st.items() if k != "force"), st
    for s in (b, c):
        st = s.stageTimes()
        assert all(cnt == 7 for _, cnt in st.values()), st
        assert all(ms > 0.0 for k, (ms, _) in st.items() if k not in ("hash", "permute")), st
        assert st["hash"][0] >= 0.0 and st["permute"][0] >= 0.0, st
    st = d.stageTimes()
    assert st["force"][1] == 7 and st["force"][0] > 0.0, st
    assert all(cnt == 0 for k, (_, cnt) in st.items() if k != "force"), st
    with pytest.raises(RTPSError):
        d.setOption(3, 4)
    for s in (a, b, c, d, e, f):
        s.close()


def test_autotune_choice_does_not_change_the_step(cuda):
    """The autotuned step (launch configuration timed on the live state) gives
    the same state as the fixed default configuration, and reports a valid
    choice; the removed force loops and autotune modes are refused."""
    sc = scenes.dam_custom(24, 24, 12, seed=79)
    a = _system(sc)
    a.setOption(7, 0)          # autotune off: 256 threads, the default force loop
    b = _system(sc)
    tune = b.autotune(measure=True)
    assert tune["block_density"] in (64, 128, 256) and tune["block_force"] in (64, 128, 256)
    assert tune["force_variant"] in (0, 4, 5)
    a.step(1)
    np.testing.assert_array_equal(b.read("pos"), a.read("pos"))
    a.step(4)
    b.step(4)
    np.testing.assert_array_equal(b.read("pos"), a.read("pos"))
    np.testing.assert_array_equal(b.read("vel"), a.read("vel"))
    for opt, bad in ((7, 2), (4, 1), (4, 2), (4, 3), (4, 6)):
        with pytest.raises(RTPSError):
            b.setOption(opt, bad)
    for s in (a, b):
        s.close()


@pytest.mark.parametrize("scene", ["dam", "clump"])
def test_force_loop_variants_agree(cuda, scene):
    """The three force loops (direct, direct with the rows staged through LDS,
    staged rows with the accepted candidates of each 32-candidate chunk marked
    in a bit mask first) visit the same pairs in the same order: the same
    bits.  The clump packs 300 particles into one cell, so its rows exceed one
    LDS copy and the staged loops fall back to global memory there."""
    sc = scenes.dam_custom(20, 20, 10, seed=80)
    if scene == "clump":
        rng = np.random.default_rng(81)
        pos = sc.pos.copy()
        pos[:300, :3] = 0.3 + 0.012 * rng.random((300, 3), dtype=np.float32)
        sc = scenes.Scene("clump", sc.settings, pos, sc.vel, 2, sc.dt)
    out = {}
    for v in (0, 4, 5):
        s = _system(sc)
        s.setOption(7, 0)      # autotune off: the process-wide variant applies
        s.setOption(4, v)
        s.step(2)
        out[v] = (s.read("pos"), s.read("vel"))
        s.close()
    s = _system(sc)            # the variant option is process-wide: restore the default
    s.setOption(4, DEFAULT_FORCE_LOOP)
    s.close()
    for v in (4, 5):
        np.testing.assert_array_equal(out[v][0], out[0][0], err_msg=f"loop {v}")
        np.testing.assert_array_equal(out[v][1], out[0][1], err_msg=f"loop {v}")


def test_tile_queues_give_the_same_step(cuda):
    """RTPS_OPT_TILE_QUEUES: the density and force kernels run as resident
    grids over eight per-XCD tile queues and give the one-tile-per-block
    step's bits for each force loop and block size, over several steps (each
    launch's last block clears the counters for the next), through a deletion
    step (device-side count), with fewer tiles than resident blocks, on the
    split (triangle-free euler) path, and when autotuned."""
    from enjaparticles_amd._native import OPT_TILE_QUEUES
    cases = [("dam", scenes.dam16k(), {}), ("small", scenes.dam_custom(10, 8, 6, seed=91), {}),
             ("euler", scenes.dam16k(), {"integrator": "euler"})]
    for name, sc, extra in cases:
        settings = dict(sc.settings, **extra)
        for bs, loop in ((64, 5), (128, 4), (256, 0)):
            out = []
            for q in (0, 1):
                s = RTPS.generateSystemInstance(settings)
                s.pushParticles(sc.pos, sc.vel)
                s.setOption(4, loop)
                s.setOption(11, bs)
                s.setOption(12, bs)
                s.setOption(OPT_TILE_QUEUES, q)
                if name == "dam" and bs == 64:
                    s.testDelete()   # the next update deletes 40 particles
                s.step(4)
                out.append((s.getNum(), s.read("pos"), s.read("vel"), s.read("density")))
                s.close()
            assert out[0][0] == out[1][0], (name, bs)
            for a, b in zip(out[0][1:], out[1][1:]):
                np.testing.assert_array_equal(b, a, err_msg=f"{name} block {bs} loop {loop}")
    s = _system(scenes.dam16k())   # the variant option is process-wide: restore the default
    s.setOption(4, DEFAULT_FORCE_LOOP)
    s.close()
    sc = scenes.dam16k()
    a = _system(sc)
    a.setOption(OPT_TILE_QUEUES, 1)
    a.autotune(measure=True)
    a.step(3)
    b = _system(sc)
    b.step(4)
    np.testing.assert_array_equal(a.read("pos"), b.read("pos"))
    with pytest.raises(RTPSError):
        a.setOption(OPT_TILE_QUEUES, 2)
    a.close()
    b.close()


def test_fixed_launch_options(cuda):
    """RTPS_OPT_BLOCK_DENSITY / RTPS_OPT_BLOCK_FORCE fix the neighbour launches
    (no autotuning): any block size gives the autotuned step's bits; other
    values are rejected."""
    sc = scenes.dam_custom(24, 24, 12, seed=85)
    a = _system(sc)
    a.autotune(measure=True)
    a.step(3)
    for bd, bf in ((64, 256), (128, 64)):
        b = _system(sc)
        b.setOption(11, bd)
        b.setOption(12, bf)
        b.step(4)
        np.testing.assert_array_equal(b.read("pos"), a.read("pos"))
        np.testing.assert_array_equal(b.read("vel"), a.read("vel"))
        with pytest.raises(RuntimeError):
            b.setOption(12, 96)
        b.close()
    a.close()


def test_sort_tile_sizes_give_the_same_step(cuda):
    """Every radix-sort tile (RTPS_OPT_SORT_TILE = threads * 100 + keys per
    thread) sorts identically (stable), so the steps are bit-identical."""
    sc = scenes.dam_custom(32, 32, 16, seed=82)
    out = {}
    for items in (102408, 102404, 102402, 102401, 51204, 51208, 25604, 25608, 25616):
        s = _system(sc)
        s.setOption(8, items)
        s.step(3)
        out[items] = (s.read("pos"), s.read("keys"), s.read("idx"))
        s.close()
    s = _system(sc)
    s.setOption(8, 0)           # process-wide: restore the default (by size)
    s.close()
    for items in out:
        for a, b in zip(out[items], out[102402]):
            np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("scene", ["dam", "clump"])
def test_density_stage_gives_the_same_step(cuda, scene):
    """Density rows copied once per wave into LDS (RTPS_OPT_DENSITY_STAGE 1)
    or read by every lane from global memory (0): the same candidates in the
    same order, so the steps are bit-identical.  The clump's rows hold more
    candidates than one LDS copy, so those waves take the global path."""
    sc = scenes.dam_custom(30, 30, 15, seed=83)   # 13,500 particles: the last wave is partly empty
    if scene == "clump":
        rng = np.random.default_rng(84)
        pos = sc.pos.copy()
        pos[:600, :3] = 0.3 + 0.012 * rng.random((600, 3), dtype=np.float32)
        sc = scenes.Scene("clump", sc.settings, pos, sc.vel, 2, sc.dt)
    out = {}
    for stage in (1, 0):
        s = _system(sc)
        s.setOption(10, stage)
        s.step(3)
        out[stage] = (s.read("pos"), s.read("vel"), s.read("density"))
        s.close()
    s = _system(sc)
    s.setOption(10, 1)          # process-wide: restore the default
    s.close()
    for a, b in zip(out[1], out[0]):
        np.testing.assert_array_equal(a, b)


def _append(st, pos, vel):
    """Oracle state with particles pushed after the live ones (veleval 0)."""
    n = st.num
    return O.State(np.concatenate([st.pos_u[:n], pos]), np.concatenate([st.vel_u[:n], vel]), st.sphp, st.gp,
                   st.gps, veleval=np.concatenate([st.veleval_u[:n], np.zeros_like(pos)]))


def test_hose_sprays_match_oracle(cuda):
    """Hoses (SPH::addHose / sprayHoses, SPH.cpp:166, 523-557) spray before
    every update: after each step the system holds the particles the hose
    restatement sprays (same seed), moved as the oracle moves them (positions
    1e-4 relative per element, as over the first steps of the dam-break)."""
    import hose_oracle as H
    sc = scenes.dam_custom(16, 16, 8, seed=41)
    settings = {**sc.settings, "max_num_particles": str(sc.n + 600)}
    box = float(sc.settings["domain_max"].split()[0])
    center, vel, radius, total = (0.5 * box, 0.5 * box, 0.8 * box, 1.0), (0.1, 0.05, -1.0, 0.0), 3.0, 500
    s = RTPS.generateSystemInstance(settings)
    s.pushParticles(sc.pos, sc.vel)
    assert s.addHose(total, center, vel, radius, color=(0.0, 0.0, 1.0, 1.0)) == 0
    spacing = np.float32(s.getSetting("spacing"))
    dt = np.float32(s.getSetting("time_step"))
    hose = H.Hose(dt, total, center, vel, np.float32(radius) * spacing, spacing, 0x5EED)
    sp, gp, gs = O.derive_params(settings)
    st = O.State(sc.pos, sc.vel, sp, gp, gs)
    sprays = 0
    for k in range(24):
        p = np.array(hose.spray(), np.float32).reshape(-1, 4)
        if len(p):
            sprays += 1
            st = _append(st, p, np.tile(np.array(hose.velocity, np.float32), (len(p), 1)))
        st.step(sc.dt, O.FLAG_COLLISION)
        s.step(1)
        assert s.getNum() == st.num, k
        np.testing.assert_allclose(s.read("pos")[:, :3], st.pos_u[: st.num, :3], rtol=1e-4, atol=1e-6)
    assert sprays >= 3 and st.num > sc.n + 100
    # sprayed particles carry the hose colour, the pushed ones the default
    col = s.read("color")
    np.testing.assert_array_equal(col[: sc.n], np.tile([1.0, 0.0, 0.0, 1.0], (sc.n, 1)).astype(np.float32))
    np.testing.assert_array_equal(col[sc.n:], np.tile([0.0, 0.0, 1.0, 1.0], (st.num - sc.n, 1)).astype(np.float32))
    # the split update()/integrate() path sprays at the start of update()
    b = RTPS.generateSystemInstance(settings)
    b.pushParticles(sc.pos, sc.vel)
    b.addHose(total, center, vel, radius)
    for _ in range(24):
        b.update()
        b.integrate()
    assert b.getNum() == st.num
    np.testing.assert_allclose(b.read("pos")[:, :3], st.pos_u[: st.num, :3], rtol=1e-4, atol=1e-6)
    # refill and move
    b.refillHose(0, 10)
    b.updateHose(0, (0.4 * box, 0.5 * box, 0.8 * box, 1.0), vel, 2.0)
    n0 = b.getNum()
    for _ in range(10):
        b.step(1)
    assert b.getNum() == n0 + 10
    with pytest.raises(Exception):
        b.refillHose(3, 10)
    for x in (s, b):
        x.close()


def test_test_delete_removes_forty(cuda):
    """System::testDelete (System.cpp:366-376) moves the first 40 particles
    out of the grid; the next update deletes them (SPH.cpp:220-250)."""
    sc = scenes.dam16k()
    s = _system(sc)
    s.testDelete()
    pos = sc.pos.copy()
    pos[:40] = (100.0, 100.0, 100.0, 1.0)
    st = _oracle(scenes.Scene("del40", sc.settings, pos, sc.vel, 2, sc.dt))
    for _ in range(2):
        s.step(1)
        st.step(sc.dt, O.FLAG_COLLISION)
    n = s.getNum()
    assert n == st.num == sc.n - 40
    from scipy.spatial import cKDTree
    d, j = cKDTree(st.pos_u[:n, :3].astype(np.float64)).query(s.read("pos")[:, :3].astype(np.float64))
    assert d.max() <= 1e-5 and len(np.unique(j)) == n
    s.close()


@pytest.mark.parametrize("fused", [True, False])
def test_colors_follow_their_particles(cuda, fused):
    """Colours are permuted with the particles (permute.cl:41) and renumbered
    with the survivors of a deletion (leapfrog.cl:107, SPH.cpp:450-457): every
    particle is labelled by its colour and must keep its label."""
    sc = scenes.dam16k()
    pos = sc.pos.copy()
    gone = np.arange(5, sc.n, 211)
    pos[gone, 2] = 70.0    # far above the grid: deleted by the first step
    s = RTPS.generateSystemInstance(sc.settings)
    s.pushParticles(pos, sc.vel)
    label = np.zeros((sc.n, 4), np.float32)
    label[:, 0] = np.arange(sc.n)
    label[:, 3] = 1.0
    s.write("color", label)   # written colours are kept with their particles
    p0 = s.read("pos")

    def one_step():
        if fused:
            s.step(1)
        else:
            s.update()
            s.integrate()

    one_step()
    n = s.getNum()
    assert n == sc.n - len(gone)
    idx = s.read("idx", sc.n)
    cs = s.read("color_s", sc.n)
    np.testing.assert_array_equal(cs, label[idx])             # sorted colours: colour_u[idx]
    cu = s.read("color")
    np.testing.assert_array_equal(cu, cs[:n])                  # survivors renumbered in sorted order
    ids = cu[:, 0].astype(np.int64)
    assert not np.isin(ids, gone).any() and len(np.unique(ids)) == n
    # each surviving slot holds the particle its label names, moved by one step
    p1 = s.read("pos")
    assert np.abs(p1[:, :3] - p0[ids, :3]).max() < 0.01
    # without a deletion colours stay in their slots
    one_step()
    np.testing.assert_array_equal(s.read("color"), cu)
    np.testing.assert_array_equal(s.read("color_s", n), cu[s.read("idx", n)])
    m = s.read("mass")
    assert m.shape == (n,) and np.all(m == np.float32(s.getSetting("mass")))
    s.close()


def test_read_async_into_page_locked_memory(cuda):
    """rtps_sph_read_async: the copy lands in a page-locked buffer at the next
    synchronisation, holding the state at the point it was enqueued."""
    from enjaparticles_amd._native import RTPSError
    from enjaparticles_amd.rtps import HostBuffer
    sc = scenes.dam_custom(16, 16, 8, seed=9)
    s = _system(sc)
    s.step(2)
    buf, ib = HostBuffer(sc.n), HostBuffer(sc.n, np.uint32)
    before, idx_before = s.read("pos"), s.read("idx")
    s.readAsync("pos", buf)
    s.readAsync("idx", ib)
    s.step(3)                  # enqueued after the copies
    s.synchronize()
    np.testing.assert_array_equal(buf.array, before)
    np.testing.assert_array_equal(ib.array, idx_before)
    after = s.read("pos")
    assert not np.array_equal(after, before)
    s.readAsync("pos", buf)
    s.synchronize()
    np.testing.assert_array_equal(buf.array, after)
    with pytest.raises(RTPSError):
        s.readAsync("density", buf)
    buf.close()
    ib.close()
    s.close()


@pytest.mark.parametrize("scene", ["dam", "clump_deletion"])
def test_window_table_equals_key_search(cuda, scene):
    """The x windows of the neighbour rows looked up in the sub-cell start
    table (RTPS_OPT_WINDOW 1) are exactly the binary-search windows (2): the
    steps are bit-identical, also with empty cells, a 300-particle cell and
    particles leaving the grid."""
    sc = scenes.dam_custom(20, 20, 10, seed=90)
    if scene == "clump_deletion":
        rng = np.random.default_rng(91)
        pos = sc.pos.copy()
        pos[:300, :3] = 0.3 + 0.012 * rng.random((300, 3), dtype=np.float32)
        pos[400::97, 2] = 70.0
        sc = scenes.Scene("clump", sc.settings, pos, sc.vel, 2, sc.dt)
    out = []
    for w in (1, 2):
        s = _system(sc)
        s.setOption(5, w)
        s.step(4)
        out.append((s.read("pos"), s.read("vel"), s.getNum(), s.countPairs()))
        s.close()
    np.testing.assert_array_equal(out[0][0], out[1][0])
    np.testing.assert_array_equal(out[0][1], out[1][1])
    assert out[0][2] == out[1][2] and out[0][3] == out[1][3]


def test_window_table_is_used_only_when_written(cuda):
    """After steps with the searched windows (RTPS_OPT_WINDOW 2), switching to
    the table (1) takes effect at the next step: queries on the current
    sorted state (pair counts, colour field) keep the searched windows, since
    the table was not written for it."""
    sc = scenes.dam_custom(16, 16, 8, seed=92)
    s = _system(sc)
    s.setOption(5, 2)
    s.step(3)
    k2, f2 = s.countPairs(), s.colorField(12)
    s.setOption(5, 1)
    assert s.countPairs() == k2
    np.testing.assert_array_equal(s.colorField(12), f2)
    s.step(1)                  # writes the table; queries now read it
    k1 = s.countPairs()
    s.setOption(5, 2)
    assert s.countPairs() == k1
    s.close()


def test_full_size_dam_invariants(cuda):
    """The 1,048,576-particle dam break (the benchmark workload) after a few
    steps, through properties that do not depend on size: the sorted keys
    ascend and idx is a permutation; every occupied cell's [start, end)
    holds exactly its keys; the stored SPH forces and XSPH corrections sum to
    zero over all particles (equal masses: every pair term has a mirror term
    of opposite sign, m / (rho_i rho_j) being symmetric), to FP32 rounding."""
    sc = scenes.dam1m()
    s = _system(sc)
    s.setOption(2, 1)           # keep force / xsph
    s.step(3)
    n = s.getNum()
    assert n == sc.n
    keys = s.read("keys").astype(np.int64)
    idx = s.read("idx")
    assert np.all(np.diff(keys) >= 0)
    assert np.array_equal(np.sort(idx), np.arange(n, dtype=np.uint32))
    start, end = s.read("cell_start"), s.read("cell_end")
    occ = np.unique(keys)
    occ = occ[occ < len(start) - 1]
    first = np.searchsorted(keys, occ, side="left")
    last = np.searchsorted(keys, occ, side="right")
    np.testing.assert_array_equal(start[occ], first)
    np.testing.assert_array_equal(end[occ], last)
    F = s.read("force")[:, :3].astype(np.float64)
    X = s.read("xsph")[:, :3].astype(np.float64)
    print(f"|sum F| / sum |F| = {np.abs(F.sum(0)).max() / np.abs(F).sum():.2e}, "
          f"|sum X| / sum |X| = {np.abs(X.sum(0)).max() / np.abs(X).sum():.2e}")
    assert np.abs(F.sum(0)).max() <= 1e-6 * np.abs(F).sum()     # measured ~2e-10
    assert np.abs(X.sum(0)).max() <= 1e-6 * np.abs(X).sum()     # measured ~3e-11
    s.close()
