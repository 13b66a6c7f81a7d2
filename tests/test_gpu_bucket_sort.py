"""The step's bucket sort of the cell keys (RTPS_OPT_BUCKET_SORT 1, the
default) against the LSD radix sort (0), which the stage tests hold bit-exact
to a stable sort: the sorted keys and particle indices, and therefore every
state after the steps, must be bit-identical.

The scenes cover the bucket sort's paths: buckets of up to 256 and 512 keys
sorted by one wave (the dam lattice; uniform positions), buckets above 512
keys sorted by a block (clumps in one cell: 4, 8 and 16 slots per thread),
runs of equal keys ordered by particle index (every scene; up to ~190 keys
long in the 3000-particle clump), the global-memory passes (a bucket above
4096 keys; a run of more than 256 equal keys in a small bucket: particles
stacked on one point), out-of-grid particles (all in the clamped cell), a
deletion step (pad keys past the device count), and systems of 1, 2 and 65
particles.
"""
import numpy as np
import pytest

from enjaparticles_amd import scenes
from enjaparticles_amd._native import OPT_BUCKET_SORT, RTPSError
from enjaparticles_amd.rtps import RTPS

pytestmark = pytest.mark.gpu


def _clump(n, side_cells, seed, offset=(0.5, 0.5, 0.5)):
    """n particles uniform in a cube of side_cells cells (h = 0.04) around
    offset in the 1.0 box of the 16k dam settings."""
    u = scenes.uniform01(seed, 3 * n).reshape(n, 3).astype(np.float64)
    pos = np.ones((n, 4), np.float32)
    pos[:, :3] = (np.asarray(offset) + (u - 0.5) * side_cells * scenes.H).astype(np.float32)
    return pos


def _run(settings, pos, vel, bucket, steps, delete=False):
    s = RTPS.generateSystemInstance(settings)
    s.pushParticles(pos, vel)
    s.setOption(OPT_BUCKET_SORT, bucket)
    if delete:
        s.testDelete()   # the next update deletes 40 particles
    s.step(1)
    first = (s.read("keys"), s.read("idx"))
    if steps > 1:
        s.step(steps - 1)
    out = first + (s.getNum(), s.read("keys"), s.read("idx"), s.read("pos"), s.read("vel"), s.read("density"))
    s.close()
    return out


def _uniform(n, lo, hi, seed):
    pos = np.ones((n, 4), np.float32)
    pos[:, :3] = scenes.uniform01(seed, 3 * n).reshape(n, 3) * np.float32(hi - lo) + np.float32(lo)
    return pos


def _cases():
    # the 16k dam's grid: 30^3 cells of 0.04 from -0.08, 19-bit keys, buckets
    # of 16 cells along x; 0.5 is a cell centre, so a clump of half a cell
    # around it is one cell (one bucket)
    dam = scenes.dam16k()
    box = dict(dam.settings, max_num_particles="20000")
    out_of_grid = dam.pos.copy()
    out_of_grid[::97, 0] = np.float32(5.0)
    out_of_grid[1::89, 2] = np.float32(-3.0)
    return [
        ("dam16k", dam.settings, dam.pos, 4, False),
        ("dam16k-delete", dam.settings, dam.pos, 3, True),
        ("random-wide", box, _uniform(20000, 0.01, 0.99, 7), 2, False),     # ~20 keys per bucket
        ("random-dense", box, _uniform(20000, 0.4, 0.6, 8), 2, False),      # ~400-800 keys per bucket
        ("clump-700", box, _clump(700, 0.5, 11), 2, False),                 # block, 4 slots
        ("clump-1500", box, _clump(1500, 0.5, 12), 2, False),               # block, 8 slots
        ("clump-3000", box, _clump(3000, 0.5, 13), 1, False),               # block, 16 slots, long runs
        ("clump-9000", box, _clump(9000, 0.5, 14), 1, False),               # above 4096: global memory
        ("stack-400", box, _clump(400, 0.0, 15), 1, False),                 # 400 equal keys: global memory
        ("out-of-grid", dam.settings, out_of_grid, 2, False),
        ("n1", box, dam.pos[:1], 2, False),
        ("n2", box, dam.pos[:2], 2, False),
        ("n65", box, dam.pos[:65], 2, False),
    ]


@pytest.mark.parametrize("case", _cases(), ids=lambda c: c[0])
def test_bucket_sort_gives_the_radix_sort_step(cuda, case):
    name, settings, pos, steps, delete = case
    vel = np.zeros_like(pos)
    a = _run(settings, pos, vel, 1, steps, delete)
    b = _run(settings, pos, vel, 0, steps, delete)
    np.testing.assert_array_equal(a[0], b[0], err_msg=f"{name}: keys after the first step")
    np.testing.assert_array_equal(a[1], b[1], err_msg=f"{name}: idx after the first step")
    assert a[2] == b[2], name
    for x, y, what in zip(a[3:], b[3:], ("keys", "idx", "pos", "vel", "density")):
        np.testing.assert_array_equal(x, y, err_msg=f"{name}: {what}")
    if len(a[0]) > 1:
        assert np.all(np.diff(a[0].astype(np.int64)) >= 0), name


def test_full_size_dam_bucket_sort_equals_radix_sort(cuda):
    """The 1,048,576-particle dam (the benchmark workload): two steps with
    each sort give the same keys, indices and positions."""
    sc = scenes.dam1m()
    out = []
    for bucket in (1, 0):
        s = RTPS.generateSystemInstance(sc.settings)
        s.pushParticles(sc.pos, sc.vel)
        s.setOption(OPT_BUCKET_SORT, bucket)
        s.step(2)
        out.append((s.read("keys"), s.read("idx"), s.read("pos")))
        s.close()
    for x, y in zip(*out):
        np.testing.assert_array_equal(x, y)


def test_bucket_sort_option_is_checked(cuda):
    s = RTPS.generateSystemInstance(scenes.dam16k().settings)
    with pytest.raises(RTPSError):
        s.setOption(OPT_BUCKET_SORT, 2)
    s.close()
