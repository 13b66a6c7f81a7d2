"""Batched scenes (rtps_batch_*): every scene of a batch must compute what a
plain single-scene system computes from the same settings and particles
(segmented keys bit-exact, positions/densities within 1e-6 relative), and
match the CPU oracle at the plain system's 1e-4 bar."""
import numpy as np
import pytest

import oracle as O
from enjaparticles_amd import scenes
from enjaparticles_amd._native import RTPSError
from enjaparticles_amd.rtps import SPH, BatchedSPH

pytestmark = pytest.mark.gpu


def _batch(scs):
    b = BatchedSPH([sc.settings for sc in scs])
    for s, sc in enumerate(scs):
        b.pushParticles(s, sc.pos, sc.vel)
    return b


def test_each_scene_equals_plain_system(cuda):
    scs = scenes.batched(6, 600, 6000, seed=77)
    b = _batch(scs)
    b.step(5)
    for s, sc in enumerate(scs):
        p = SPH(sc.settings)
        p.pushParticles(sc.pos, sc.vel)
        p.step(5)
        assert b.getNum(s) == p.getNum() == sc.n
        kb = b.read(s, "keys")
        kp = p.read("keys")
        np.testing.assert_array_equal(kb, (np.uint32(s) << np.uint32(24)) | kp)
        np.testing.assert_array_equal(b.read(s, "idx"), p.read("idx"))
        np.testing.assert_allclose(b.read(s, "pos")[:, :3], p.read("pos")[:, :3], rtol=1e-6, atol=1e-7)
        np.testing.assert_allclose(b.read(s, "density"), p.read("density"), rtol=1e-6)
        p.close()
    b.close()


def test_ragged_tiny_and_empty_scenes_equal_plain_systems(cuda):
    """Scenes of 1, 65, 0 and 2 particles beside a full-size one: every
    scene, the empty one included, steps as a plain system does."""
    scs = scenes.batched(5, 600, 6000, seed=79)
    for s, n in [(0, 1), (1, 65), (2, 0), (3, 2)]:
        sc = scs[s]
        scs[s] = scenes.Scene(sc.name, sc.settings, sc.pos[:n].copy(), sc.vel[:n].copy(), sc.steps, sc.dt)
    b = _batch(scs)
    b.step(5)
    for s, sc in enumerate(scs):
        p = SPH(sc.settings)
        p.pushParticles(sc.pos, sc.vel)
        p.step(5)
        assert b.getNum(s) == p.getNum() == sc.n
        np.testing.assert_array_equal(b.read(s, "keys"), (np.uint32(s) << np.uint32(24)) | p.read("keys"))
        np.testing.assert_array_equal(b.read(s, "idx"), p.read("idx"))
        np.testing.assert_allclose(b.read(s, "pos")[:, :3], p.read("pos")[:, :3], rtol=1e-6, atol=1e-7)
        np.testing.assert_allclose(b.read(s, "density"), p.read("density"), rtol=1e-6)
        p.close()
    b.close()


def test_derived_parameter_scenes_equal_plain_systems(cuda):
    """Scenes whose mass, h and simulation scale are derived as in
    SPHSettings.cpp:45-75 (scale != 1, a different one per scene): the batched
    kernels' scaled positions give what plain systems give."""
    scs = []
    for k, (max_num, nx) in enumerate([(4096, 12), (16384, 16), (65536, 20)]):
        st = scenes.dam_settings(1.0, max_num)
        for key in ("mass", "rest_distance", "smoothing_distance", "simulation_scale"):
            del st[key]
        sp, _, _ = O.derive_params(st)
        assert abs(sp.simulation_scale - 1.0) > 0.3
        pos = scenes.lattice(nx, nx, nx // 2, seed=90 + k, spacing=sp.rest_distance / sp.simulation_scale)
        scs.append(scenes.Scene(f"derived{k}", st, pos, np.zeros_like(pos), 4, 0.003))
    b = _batch(scs)
    b.step(4)
    for s, sc in enumerate(scs):
        p = SPH(sc.settings)
        p.pushParticles(sc.pos, sc.vel)
        p.step(4)
        np.testing.assert_array_equal(b.read(s, "idx"), p.read("idx"))
        np.testing.assert_allclose(b.read(s, "pos")[:, :3], p.read("pos")[:, :3], rtol=1e-6, atol=1e-7)
        np.testing.assert_allclose(b.read(s, "density"), p.read("density"), rtol=1e-6)
        p.close()
    b.close()


def test_scene_matches_oracle(cuda):
    scs = scenes.batched(3, 2000, 4000, seed=5)
    b = _batch(scs)
    for _ in range(5):
        b.step(1)
    for s, sc in enumerate(scs):
        sp, gp, gs = O.derive_params(sc.settings)
        st = O.State(sc.pos, sc.vel, sp, gp, gs)
        for _ in range(5):
            st.step(sc.dt, O.FLAG_COLLISION)
        np.testing.assert_allclose(b.read(s, "pos")[:, :3], st.pos_u[:, :3], rtol=1e-4)
        # densities are stored in sorted order; compare them by original index
        # (the within-cell order of the two sorts differs)
        d_gpu = np.empty(sc.n, np.float32)
        d_gpu[b.read(s, "idx")] = b.read(s, "density")
        d_ref = np.empty(sc.n, np.float32)
        d_ref[st.idx[: st.num]] = st.density[: st.num]
        np.testing.assert_allclose(d_gpu, d_ref, rtol=1e-4)
    b.close()


def test_graph_equals_eager_and_deletion(cuda):
    scs = scenes.batched(4, 800, 3000, seed=9)
    pos = scs[2].pos.copy()
    pos[::50, 2] = 60.0                       # out of scene 2's grid: deleted
    scs[2] = scenes.Scene(scs[2].name, scs[2].settings, pos, scs[2].vel, 3, scs[2].dt)
    a, e = _batch(scs), _batch(scs)
    e.setOption(1, 0)                         # eager launches
    a.step(3)
    e.step(3)
    for s in range(4):
        assert a.getNum(s) == e.getNum(s)
        np.testing.assert_array_equal(a.read(s, "pos"), e.read(s, "pos"))
    assert a.getNum(2) == len(pos) - len(pos[::50])
    assert a.getNum(1) == scs[1].n
    a.close()
    e.close()


def test_capacity_and_arguments(cuda):
    scs = scenes.batched(2, 500, 600, seed=1)
    b = _batch(scs)
    with pytest.raises(RTPSError):
        b.pushParticles(0, scs[0].pos[:1])      # scene full
    with pytest.raises(RTPSError):
        b.pushParticles(5, scs[0].pos[:1])      # no such scene
    b.close()
    with pytest.raises(RTPSError):
        BatchedSPH([{"system": "flock", "max_num_particles": "10"}])


def test_pair_kernel_configurations_give_the_same_step(cuda):
    """Every launch configuration the batched autotuner chooses from (block
    of 64, 128 or 256 slots; 4, 5 or 6 waves per SIMD for density, 3, 4 or 5 for
    force) computes the same step, bit for bit; the autotuner's own step
    too.  Force at 6 waves is refused."""
    scs = scenes.batched(5, 600, 6000, seed=83)
    ref = None
    for block in (256, 128, 64):
        for wd, wf in ((5, 4), (4, 3), (6, 4), (5, 3), (5, 5)):
            b = _batch(scs)
            b.setOption(11, block * 10 + wd)   # RTPS_OPT_BLOCK_DENSITY
            b.setOption(12, block * 10 + wf)   # RTPS_OPT_BLOCK_FORCE
            b.step(4)
            out = [(b.read(s, "pos"), b.read(s, "density")) for s in range(len(scs))]
            b.close()
            if ref is None:
                ref = out
                continue
            for (p, d), (pr, dr) in zip(out, ref):
                np.testing.assert_array_equal(p, pr, err_msg=f"block {block} waves {wd}/{wf}")
                np.testing.assert_array_equal(d, dr, err_msg=f"block {block} waves {wd}/{wf}")
    b = _batch(scs)
    with pytest.raises(RTPSError):
        b.setOption(12, 2566)
    with pytest.raises(RTPSError):
        b.setOption(11, 2563)
    cfg = b.autotune(measure=True)
    assert cfg["density"]["block"] in (64, 128, 256) and cfg["density"]["waves"] in (4, 5, 6)
    assert cfg["force"]["block"] in (64, 128, 256) and cfg["force"]["waves"] in (3, 4, 5)
    b.step(3)
    for s in range(len(scs)):
        np.testing.assert_array_equal(b.read(s, "pos"), ref[s][0])
    b.close()
