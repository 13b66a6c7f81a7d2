"""Stage-by-stage parity of the HIP kernels (through the C ABI) with the CPU oracle.

Bars: hash keys/idx, sorted keys, cell start (and end on occupied cells) and
the permutation are bit-exact; density/force/XSPH/collision/leapfrog are
FP32 within the tolerances written in each test (1e-4 relative or tighter).
"""
import numpy as np
import pytest

import oracle as O
from enjaparticles_amd import scenes

pytestmark = pytest.mark.gpu

EMPTY = np.uint32(0xFFFFFFFF)


def _t(a, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def _np(t):
    return t.cpu().numpy()


def _u(t):
    return _np(t).view(np.uint32)


def _dam(nx=32, ny=32, nz=16, seed=1001, box=1.0):
    sc = scenes.dam_custom(nx, ny, nz, seed=seed, box=box)
    sp, gp, gs = O.derive_params(sc.settings)
    return sc, sp, gp, gs


def _oracle_front(pos, sp, gp, gs, vel=None, veleval=None):
    n = pos.shape[0]
    vel = np.zeros_like(pos) if vel is None else vel
    veleval = np.zeros_like(pos) if veleval is None else veleval
    k, i = O.hash(pos, gp)
    k, i = O.sort_stable(k, i)
    start, end, nc = O.cellindices(k, gp.nb_cells)
    ps, vs, es = O.permute(i, pos, vel, veleval)
    return dict(keys=k, idx=i, start=start, end=end, nc=nc, pos_s=ps, vel_s=vs, veleval_s=es)


def test_params_match_oracle():
    from enjaparticles_amd import _native as N
    for sc in (scenes.dam16k(), scenes.dam_custom(20, 20, 10)):
        a = N.derive_params(sc.settings)
        b = O.derive_params(sc.settings)
        for x, y in zip(a, b):
            assert bytes(x) == bytes(y)


@pytest.mark.parametrize("case", ["dam16k", "random_out_of_grid", "empty", "one"])
def test_hash_bitexact(cuda, case):
    from enjaparticles_amd import ops
    sc, sp, gp, gs = _dam()
    pos = sc.pos
    if case == "random_out_of_grid":
        rng = np.random.default_rng(7)
        pos = rng.uniform(-0.3, 1.3, size=(5003, 4)).astype(np.float32)
        pos[:, 3] = 1
        pos[0] = [np.nan, 0.5, 0.5, 1]
        pos[1] = [1e30, 0.5, 0.5, 1]
        pos[2] = [-1e30, -1e30, -1e30, 1]
        pos[3] = [-0.07999, -0.0799, -0.0799, 1]   # just inside the padded grid
    elif case == "empty":
        pos = pos[:0]
    elif case == "one":
        pos = pos[:1]
    keys, idx = ops.hash(_t(pos, cuda), gp)
    ok, oi = O.hash(np.ascontiguousarray(pos), gp)
    np.testing.assert_array_equal(_u(keys), ok)
    np.testing.assert_array_equal(_u(idx), oi)


def test_hash_pad_keys(cuda):
    from enjaparticles_amd import ops
    sc, sp, gp, gs = _dam()
    keys, idx = ops.hash(_t(sc.pos, cuda), gp, num=1000, n_pad=1500)
    bits = ops.key_bits(gp.nb_cells)
    k = _u(keys)
    assert np.all(k[1000:] == (1 << bits) - 1)
    assert (1 << bits) - 1 > gp.nb_cells
    np.testing.assert_array_equal(_u(idx), np.arange(1500))


@pytest.mark.parametrize("n,bits", [(0, 15), (1, 15), (2, 1), (4095, 8), (4097, 15), (16384, 15), (70001, 20),
                                    (1 << 20, 20), ((1 << 20) + 4097, 24), (333333, 24), (5000, 32)])
def test_sort_matches_stable_sort(cuda, n, bits):
    from enjaparticles_amd import ops
    rng = np.random.default_rng(n + bits)
    hi = (1 << bits) - 1 if bits < 32 else 0xFFFFFFFF
    k = rng.integers(0, hi, size=n, dtype=np.uint64, endpoint=True).astype(np.uint32)
    if n > 10:
        k[: n // 3] = k[0]       # long runs of equal keys exercise stability
    v = np.arange(n, dtype=np.uint32)
    kt, vt = _t(k.view(np.int32), cuda), _t(v.view(np.int32), cuda)
    ops.sort_pairs(kt, vt, bits)
    order = np.argsort(k, kind="stable")
    np.testing.assert_array_equal(_u(kt), k[order])
    np.testing.assert_array_equal(_u(vt), v[order])


@pytest.mark.parametrize("direction,in_place", [(1, False), (0, False), (1, True), (0, True)])
def test_sort_segments_bitonic_contract(cuda, direction, in_place):
    """Bitonic<T>::Sort(batch, arrayLength, dir, dst, src) (BitonicSort.h:26-28):
    three independent arrays of 5,000 pairs (not a power of two) on all 32
    key bits, ascending (dir 1) or descending (dir 0), src -> dst or in place;
    sources untouched when dst differs.  Keys bit-exact; values in the order
    a stable sort gives."""
    from enjaparticles_amd import ops
    import torch
    batch, L = 3, 5000
    rng = np.random.default_rng(40 + direction)
    k = rng.integers(0, 0xFFFFFFFF, size=batch * L, dtype=np.uint64, endpoint=True).astype(np.uint32)
    k[100:900] = k[50]          # runs of equal keys
    v = np.arange(batch * L, dtype=np.uint32)
    sk, sv = _t(k.view(np.int32), cuda), _t(v.view(np.int32), cuda)
    dk, dv = (sk, sv) if in_place else (torch.empty_like(sk), torch.empty_like(sv))
    ops.sort_segments(dk, dv, sk, sv, batch, L, direction)
    for b in range(batch):
        kb, vb = k[b * L:(b + 1) * L], v[b * L:(b + 1) * L]
        order = np.argsort(kb if direction else ~kb, kind="stable")
        np.testing.assert_array_equal(_u(dk)[b * L:(b + 1) * L], kb[order])
        np.testing.assert_array_equal(_u(dv)[b * L:(b + 1) * L], vb[order])
    if not in_place:
        np.testing.assert_array_equal(_u(sk), k)
        np.testing.assert_array_equal(_u(sv), v)


def test_sort_dam_keys_per_cell_sets(cuda):
    """Sorted keys bit-exact; per-cell index sets equal (the reference's bitonic
    network is unstable, so only sets are comparable with it)."""
    from enjaparticles_amd import ops
    sc, sp, gp, gs = _dam()
    keys, idx = ops.hash(_t(sc.pos, cuda), gp)
    ops.sort_pairs(keys, idx, ops.key_bits(gp.nb_cells))
    ok, oi = O.hash(sc.pos, gp)
    bk, bi = O.sort_bitonic(ok, oi)
    k = _u(keys)
    np.testing.assert_array_equal(k, bk)
    i = _u(idx)
    bounds = np.flatnonzero(np.diff(k)) + 1
    for a, b in zip(np.r_[0, bounds], np.r_[bounds, len(k)]):
        assert set(i[a:b]) == set(bi[a:b])


@pytest.mark.parametrize("case", ["dam16k", "with_out_of_grid", "single_cell"])
def test_cellindices_bitexact(cuda, case):
    from enjaparticles_amd import ops
    sc, sp, gp, gs = _dam()
    pos = sc.pos.copy()
    if case == "with_out_of_grid":
        pos[::97, 2] = 50.0         # far above: linear index > nb_cells -> key nb_cells (deletion sentinel)
    elif case == "single_cell":
        pos[:, :3] = 0.5
    ok, oi = O.hash(pos, gp)
    ok, oi = O.sort_stable(ok, oi)
    start, end, nc = ops.cellindices(_t(ok.view(np.int32), cuda), gp.nb_cells)
    os_, oe, onc = O.cellindices(ok, gp.nb_cells)
    s, e = _u(start), _u(end)
    np.testing.assert_array_equal(s, os_)
    occ = os_ != EMPTY
    np.testing.assert_array_equal(e[occ], oe[occ])
    assert int(_np(nc)[0]) == onc
    if case == "with_out_of_grid":
        assert onc == len(pos) - len(pos[::97])


def test_permute_bitexact(cuda):
    from enjaparticles_amd import ops
    sc, sp, gp, gs = _dam()
    rng = np.random.default_rng(3)
    vel = rng.normal(size=sc.pos.shape).astype(np.float32)
    ve = rng.normal(size=sc.pos.shape).astype(np.float32)
    f = _oracle_front(sc.pos, sp, gp, gs, vel, ve)
    ps, vs, es = ops.permute(_t(f["idx"].view(np.int32), cuda), _t(sc.pos, cuda), _t(vel, cuda), _t(ve, cuda))
    np.testing.assert_array_equal(_np(ps), f["pos_s"])
    np.testing.assert_array_equal(_np(vs), f["vel_s"])
    np.testing.assert_array_equal(_np(es), f["veleval_s"])


def _front_dev(f, cuda):
    return (_t(f["pos_s"], cuda), _t(f["start"].view(np.int32), cuda), _t(f["end"].view(np.int32), cuda))


@pytest.mark.parametrize("shape", [(32, 32, 16), (20, 13, 7)])
def test_density_matches_oracle(cuda, shape):
    from enjaparticles_amd import ops
    sc, sp, gp, gs = _dam(*shape)
    f = _oracle_front(sc.pos, sp, gp, gs)
    ps, st, en = _front_dev(f, cuda)
    d = _np(ops.density(ps, st, en, sp, gs))
    od = O.density(f["pos_s"], f["start"], f["end"], sp, gs)
    # tolerance: 1e-5 relative per particle (FP32 sums of <= ~40 positive terms)
    np.testing.assert_allclose(d, od, rtol=1e-5, atol=0)


def _perturbed_state(seed=5, shape=(32, 32, 16)):
    sc, sp, gp, gs = _dam(*shape)
    rng = np.random.default_rng(seed)
    pos = sc.pos.copy()
    pos[:, :3] += rng.normal(scale=0.003, size=(len(pos), 3)).astype(np.float32)
    ve = np.zeros_like(pos)
    ve[:, :3] = rng.normal(scale=0.1, size=(len(pos), 3)).astype(np.float32)
    return sc, sp, gp, gs, pos, ve


def test_force_matches_oracle(cuda):
    from enjaparticles_amd import ops
    sc, sp, gp, gs, pos, ve = _perturbed_state()
    f = _oracle_front(pos, sp, gp, gs, ve, ve)
    od = O.density(f["pos_s"], f["start"], f["end"], sp, gs)
    of, ox = O.force(f["pos_s"], od, f["veleval_s"], f["start"], f["end"], sp, gs)
    ps, st, en = _front_dev(f, cuda)
    F, X = ops.force(ps, _t(od, cuda), _t(f["veleval_s"], cuda), st, en, sp, gs)
    F, X = _np(F), _np(X)
    # tolerance: max-norm relative 1e-4 of each field (per component)
    for a, b in ((F, of), (X, ox)):
        scale = np.abs(b[:, :3]).max()
        assert np.abs(a[:, :3] - b[:, :3]).max() <= 1e-4 * scale
        assert np.all(a[:, 3] == 0)
    # and per particle 1e-4 relative of the particle's own force magnitude where it is not tiny
    mag = np.linalg.norm(of[:, :3], axis=1)
    big = mag > 1e-3 * mag.max()
    err = np.linalg.norm(F[:, :3] - of[:, :3], axis=1)
    assert np.all(err[big] <= 1e-4 * mag[big])


def test_collision_and_leapfrog_match_oracle(cuda):
    from enjaparticles_amd import ops
    sc, sp, gp, gs, pos, ve = _perturbed_state(seed=9)
    rng = np.random.default_rng(11)
    pos[:200, 2] = rng.uniform(-0.02, 0.012, 200).astype(np.float32)   # penetrating the floor
    pos[200:300, 0] = rng.uniform(0.99, 1.02, 100).astype(np.float32)  # the x max wall
    vel = np.zeros_like(pos)
    vel[:, :3] = rng.normal(scale=0.5, size=(len(pos), 3)).astype(np.float32)
    force = np.zeros_like(pos)
    force[:, :3] = rng.normal(scale=50, size=(len(pos), 3)).astype(np.float32)
    force[:5, :3] = 1e4                                                  # velocity-limit clamp
    xs = np.zeros_like(pos)
    xs[:, :3] = rng.normal(scale=0.01, size=(len(pos), 3)).astype(np.float32)
    fc = O.collision_wall(pos, vel, force, sp, gs)
    Fc = _np(ops.collision_wall(_t(pos, cuda), _t(vel, cuda), _t(force, cuda), sp, gs))
    scale = np.abs(fc[:, :3]).max()
    assert np.abs(Fc[:, :3] - fc[:, :3]).max() <= 1e-6 * scale
    idx = rng.permutation(len(pos)).astype(np.uint32)
    pu = np.zeros_like(pos)
    o = O.leapfrog(pos, vel, fc, xs, idx, sp, 0.003, pu, pu, pu)
    g = ops.leapfrog(_t(pos, cuda), _t(vel, cuda), _t(fc, cuda), _t(xs, cuda), _t(idx.view(np.int32), cuda), sp,
                     0.003, _t(pu, cuda), _t(pu, cuda), _t(pu, cuda))
    for a, b in zip(g, o):
        np.testing.assert_allclose(_np(a), b, rtol=1e-6, atol=1e-7)


def test_gravity_sources_match_oracle(cuda):
    from enjaparticles_amd import ops
    sc, sp, gp, gs, pos, ve = _perturbed_state(seed=13)
    force = np.zeros_like(pos)
    src = np.array([[0.2, 0.3, 0.1, 1], [0.5, 0.5, 0.2, 1]], np.float32)
    src[1, :3] = pos[17, :3]   # one particle sits on a source (inside the cutoff)
    m = np.array([1.0, 1.5], np.float32)
    a = np.array([0.3, -0.2], np.float32)
    og = O.gravity(pos, force, src, m, a, 1.0)
    g = _np(ops.gravity(_t(pos, cuda), _t(force, cuda), src, m, a, 1.0))
    scale = np.abs(og[:, :3]).max()
    # fast_normalize: rsqrt approximation; 1e-5 of the field's scale
    assert np.abs(g[:, :3] - og[:, :3]).max() <= 1e-5 * scale
