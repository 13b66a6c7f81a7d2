"""Force loop 5 reading the neighbour rows the density kernel recorded
(RTPS_OPT_SHARED_ROWS 1, the default) against the same loop looking the rows
up and testing every candidate itself (0), and against loop 4 (no bits): the
same pairs in the same order, so every state after the steps is
bit-identical.

The scenes cover the recorded rows' paths: windows of at most 32 candidates
(the dam lattice: the whole window's bits come from the record), windows
above 32 (uniform dense positions and clumps: the first chunks are tested,
the last one is recorded), rows whose wave-wide union exceeds the LDS copy
(a dense clump: global-memory walk), the scaled-position path (a simulation
scale of 0.31), a deletion step (device-side count), the unfused force
kernel (euler integrator), tile queues, every block size, and systems of 1,
2 and 65 particles.
"""
import numpy as np
import pytest

import oracle as O
from enjaparticles_amd import scenes
from enjaparticles_amd._native import (DEFAULT_FORCE_LOOP, OPT_BLOCK_DENSITY, OPT_BLOCK_FORCE, OPT_FORCE_VARIANT,
                                       OPT_SHARED_ROWS, OPT_TILE_QUEUES, RTPSError)
from enjaparticles_amd.rtps import RTPS, BatchedSPH

pytestmark = pytest.mark.gpu


def _uniform(n, lo, hi, seed):
    pos = np.ones((n, 4), np.float32)
    pos[:, :3] = scenes.uniform01(seed, 3 * n).reshape(n, 3) * np.float32(hi - lo) + np.float32(lo)
    return pos


def _clump(n, side_cells, seed):
    u = scenes.uniform01(seed, 3 * n).reshape(n, 3).astype(np.float64)
    pos = np.ones((n, 4), np.float32)
    pos[:, :3] = (0.5 + (u - 0.5) * side_cells * scenes.H).astype(np.float32)
    return pos


def _scaled():
    st = scenes.dam_settings(1.0, 16384)
    for k in ("mass", "rest_distance", "smoothing_distance", "simulation_scale"):
        del st[k]
    sp, _, _ = O.derive_params(st)
    assert abs(sp.simulation_scale - 1.0) > 0.5
    return st, scenes.lattice(24, 24, 12, seed=87, spacing=sp.rest_distance / sp.simulation_scale)


def _cases():
    dam = scenes.dam16k()
    box = dict(dam.settings, max_num_particles="20000")
    scaled_st, scaled_pos = _scaled()
    # name, settings, positions, steps, deletion, block, tile queues
    return [
        ("dam16k", dam.settings, dam.pos, 4, False, 64, 0),
        ("dam16k-delete", dam.settings, dam.pos, 3, True, 64, 0),
        ("dam16k-b128-queues", dam.settings, dam.pos, 3, False, 128, 1),
        ("dam16k-b256", dam.settings, dam.pos, 2, False, 256, 0),
        ("euler", dict(dam.settings, integrator="euler"), dam.pos, 3, False, 64, 0),
        ("scaled", scaled_st, scaled_pos, 3, False, 64, 0),
        ("random-mid", box, _uniform(20000, 0.3, 0.7, 9), 2, False, 64, 0),      # windows of 32-64, copied
        ("random-dense", box, _uniform(20000, 0.4, 0.6, 8), 2, False, 64, 0),    # windows above 64
        ("clump-1500", box, _clump(1500, 1.5, 12), 2, False, 64, 0),             # unions above the copy
        ("n1", box, dam.pos[:1], 2, False, 64, 0),
        ("n2", box, dam.pos[:2], 2, False, 64, 0),
        ("n65", box, dam.pos[:65], 2, False, 64, 0),
    ]


def _run(settings, pos, rows, steps, delete, block, queues, loop=None):
    s = RTPS.generateSystemInstance(settings)
    s.pushParticles(pos, np.zeros_like(pos))
    s.setOption(OPT_BLOCK_DENSITY, block)   # fixed launches: loop 5 unless asked otherwise
    s.setOption(OPT_BLOCK_FORCE, block)
    s.setOption(OPT_TILE_QUEUES, queues)
    s.setOption(OPT_SHARED_ROWS, rows)
    if loop is not None:
        s.setOption(OPT_FORCE_VARIANT, loop)
    if delete:
        s.testDelete()   # the next update deletes 40 particles
    s.step(steps)
    out = (s.getNum(), s.read("pos"), s.read("vel"), s.read("density"), s.read("idx"))
    s.close()
    return out


def _restore_loop():
    s = RTPS.generateSystemInstance(scenes.dam16k().settings)
    s.setOption(OPT_FORCE_VARIANT, DEFAULT_FORCE_LOOP)   # process-wide option
    s.close()


@pytest.mark.parametrize("case", _cases(), ids=lambda c: c[0])
def test_recorded_rows_give_the_looked_up_step(cuda, case):
    name, settings, pos, steps, delete, block, queues = case
    a = _run(settings, pos, 1, steps, delete, block, queues)
    b = _run(settings, pos, 0, steps, delete, block, queues)
    try:
        c = _run(settings, pos, 1, steps, delete, block, queues, loop=4)
    finally:
        _restore_loop()
    assert a[0] == b[0] == c[0], name
    for x, y, z, what in zip(a[1:], b[1:], c[1:], ("pos", "vel", "density", "idx")):
        np.testing.assert_array_equal(x, y, err_msg=f"{name}: {what}, recorded vs looked-up rows")
        np.testing.assert_array_equal(x, z, err_msg=f"{name}: {what}, loop 5 vs loop 4")
    assert np.isfinite(a[1]).all(), name


def test_full_size_dam_recorded_rows(cuda):
    """The 1,048,576-particle dam (the benchmark workload), autotuned as the
    benchmark runs it: three steps with and without recorded rows give the
    same positions, velocities and densities."""
    sc = scenes.dam1m()
    out = []
    for rows in (1, 0):
        s = RTPS.generateSystemInstance(sc.settings)
        s.pushParticles(sc.pos, sc.vel)
        s.setOption(OPT_SHARED_ROWS, rows)
        s.step(3)
        out.append((s.read("pos"), s.read("vel"), s.read("density")))
        s.close()
    for x, y in zip(*out):
        np.testing.assert_array_equal(x, y)


def _batched_run(scs, rows, steps):
    b = BatchedSPH([sc.settings for sc in scs])
    for s, sc in enumerate(scs):
        b.pushParticles(s, sc.pos, sc.vel)
    b.setOption(OPT_SHARED_ROWS, rows)
    b.step(steps)
    out = [(b.getNum(s), b.read(s, "pos"), b.read(s, "vel"), b.read(s, "density")) for s in range(len(scs))]
    b.close()
    return out


def test_batched_recorded_rows_give_the_looked_up_step(cuda):
    """The batched pair kernels with recorded rows (the default) against the
    same kernels looking the rows up: every scene bit for bit, including a
    scene at a simulation scale of 0.31, a scene losing particles out of its
    grid (deletion), scenes of 0, 1, 2 and 65 particles, and a dense scene
    whose windows exceed 32 candidates."""
    scs = scenes.batched(6, 600, 6000, seed=88)
    for s, n in [(0, 1), (1, 65), (2, 0), (3, 2)]:
        sc = scs[s]
        scs[s] = scenes.Scene(sc.name, sc.settings, sc.pos[:n].copy(), sc.vel[:n].copy(), sc.steps, sc.dt)
    st, pos = _scaled()
    scs.append(scenes.Scene("scaled", st, pos, np.zeros_like(pos), 3, 0.003))
    big = scenes.batched(1, 6000, 6000, seed=89)[0]
    pos = big.pos.copy()
    pos[::50, 2] = 60.0                       # out of the scene's grid: deleted
    scs.append(scenes.Scene("delete", big.settings, pos, big.vel, 3, big.dt))
    box = dict(scenes.dam16k().settings, max_num_particles="20000")
    dense = _uniform(20000, 0.3, 0.7, 10)
    scs.append(scenes.Scene("dense", box, dense, np.zeros_like(dense), 2, 0.003))
    a = _batched_run(scs, 1, 3)
    b = _batched_run(scs, 0, 3)
    for s, (x, y) in enumerate(zip(a, b)):
        assert x[0] == y[0], s
        for u, v, what in zip(x[1:], y[1:], ("pos", "vel", "density")):
            np.testing.assert_array_equal(u, v, err_msg=f"scene {s}: {what}")
    assert a[-2][0] == len(pos) - len(pos[::50])


def test_shared_rows_option_is_checked(cuda):
    s = RTPS.generateSystemInstance(scenes.dam16k().settings)
    with pytest.raises(RTPSError):
        s.setOption(OPT_SHARED_ROWS, 2)
    s.close()
    b = BatchedSPH([scenes.dam16k().settings])
    with pytest.raises(RTPSError):
        b.setOption(OPT_SHARED_ROWS, 2)
    b.close()
