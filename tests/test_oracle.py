"""Pins the CPU oracle independently of its own formulas.

The reference ships no tests or golden vectors for this path, so every check
here comes from mathematics the oracle's code does not share:
* closed forms: isolated-particle and simple-cubic-lattice poly6 densities,
  two-particle forces from Mueller et al. 2003's kernels (spiky gradient,
  viscosity Laplacian) and Monaghan's XSPH, free fall under leapfrog, the
  velocity clamp, single-wall penalty forces;
* brute force: O(N^2) neighbour sums with no cell grid at all;
* invariants: pairwise antisymmetry (sum of pressure/viscosity/XSPH terms is
  zero), sortedness / stability, searchsorted identities;
* FP32 against FP64.
A plausible mistake (a dropped term, a wrong sign, i/j swapped, a missed
neighbour cell) fails at least one of them.
"""
import math

import numpy as np
import pytest

import oracle as O
from enjaparticles_amd import scenes

H = 0.04
M = 0.008


def params(**over):
    st = scenes.dam_settings(1.0, 4096)
    st.update({k: str(v) for k, v in over.items()})
    return O.derive_params(st)


def one_step_front(pos, sp, gp, gs, vel=None, ve=None):
    vel = np.zeros_like(pos) if vel is None else vel
    ve = np.zeros_like(pos) if ve is None else ve
    k, i = O.hash(pos, gp)
    k, i = O.sort_stable(k, i)
    start, end, nc = O.cellindices(k, gp.nb_cells)
    ps, vs, es = O.permute(i, pos, vel, ve)
    return k, i, start, end, nc, ps, vs, es


# ---------------------------------------------------------------- parameters
def test_params_closed_forms():
    sp, gp, gs = params()
    assert sp.smoothing_distance == np.float32(0.04)
    pi = math.pi
    for got, want in ((sp.wpoly6_coef, 315 / (64 * pi * H ** 9)), (sp.wspiky_d_coef, -45 / (pi * H ** 6)),
                      (sp.wvisc_dd_coef, 45 / (pi * H ** 6))):
        assert abs(got - want) <= 1e-5 * abs(want)       # 6-significant-digit settings round trip
    # Domain::calculateCells: the unit box padded by 2 cells each side is 1.16 wide;
    # 1.16f / 0.04f rounds above 29 in FP32, so ceil gives 30 cells per axis.
    assert (gp.grid_res.x, gp.grid_res.y, gp.grid_res.z) == (30.0, 30.0, 30.0)
    assert gp.nb_cells == 27000
    assert gp.grid_min.x == np.float32(np.float32(0) - np.float32(0.08))
    assert abs(gp.grid_delta.x - 25.0) < 1e-5
    sc = scenes.dam1m()
    sp2, gp2, _ = O.derive_params(sc.settings)
    assert gp2.nb_cells == 85 ** 3


def test_params_reference_derivation():
    """Without overrides the mass follows 0.0256 / (int)log2(max_num) and h = 2 * 0.87 (m / rho0)^(1/3)."""
    st = scenes.dam_settings(10.0, 262144)
    for k in ("mass", "rest_distance", "smoothing_distance", "simulation_scale"):
        del st[k]
    sp, gp, gs = O.derive_params(st)
    m = 0.0256 / 18
    assert abs(sp.mass - m) <= 5e-6 * m        # 6 significant digits
    rd = 0.87 * (m / 1000) ** (1 / 3)
    assert abs(sp.rest_distance - rd) <= 1e-5 * rd
    assert abs(sp.smoothing_distance - 2 * rd) <= 1e-5 * rd
    vol = 10 ** 3 / 2
    scale = (0.5 * (m / 1000) * 262144 / vol) ** (1 / 3)
    assert abs(sp.simulation_scale - scale) <= 1e-5 * scale
    # the scaled grid is the unscaled one times the scale, delta divided by it
    assert abs(gs.grid_min.x - gp.grid_min.x * sp.simulation_scale) <= 1e-6 * abs(gp.grid_min.x * sp.simulation_scale)
    assert abs(gs.grid_delta.x - gp.grid_delta.x / sp.simulation_scale) <= 1e-6 * gs.grid_delta.x


# ---------------------------------------------------------------- hash / sort / cells
def test_hash_known_cells():
    sp, gp, gs = params()
    c = 1.0 / gp.grid_delta.x
    g0 = np.array([gp.grid_min.x, gp.grid_min.y, gp.grid_min.z])
    pts = np.array([g0 + np.array([3.5, 4.5, 5.5]) * c, g0 + np.array([0.2, 0.2, 0.2]) * c,
                    g0 - np.array([0.5, 0.5, 0.5]) * c,      # below the grid: truncation toward zero -> cell 0
                    g0 + np.array([1, 1, 1e3]) * c,            # far above: linear index > nb_cells -> nb_cells
                    g0 + np.array([-50.5, 2.5, 2.5]) * c])   # negative x index (truncates to -50)
    pos = np.c_[pts, np.ones(len(pts))].astype(np.float32)
    keys, idx = O.hash(pos, gp)
    rx, ry, C = 30, 30, gp.nb_cells
    assert keys[0] == (5 * ry + 4) * rx + 3
    assert keys[1] == 0 and keys[2] == 0
    assert keys[3] == C
    # gx = -50, gy = gz = 2 -> linear index (2*30+2)*30-50 = 1810: aliases into a valid cell (no edge wrap)
    assert keys[4] == (2 * ry + 2) * rx - 50
    np.testing.assert_array_equal(idx, np.arange(len(pts)))


def test_sorts():
    rng = np.random.default_rng(1)
    k = rng.integers(0, 300, 5000).astype(np.uint32)
    v = np.arange(5000, dtype=np.uint32)
    sk, sv = O.sort_stable(k, v)
    order = np.argsort(k, kind="stable")
    np.testing.assert_array_equal(sk, k[order])
    np.testing.assert_array_equal(sv, v[order])
    bk, bv = O.sort_bitonic(k, v)
    np.testing.assert_array_equal(bk, np.sort(k))
    np.testing.assert_array_equal(k[bv], bk)                 # a permutation carrying its keys
    assert len(set(bv.tolist())) == len(bv)
    for x in np.unique(k):                                   # per-key index sets agree
        assert set(sv[sk == x]) == set(bv[bk == x])
    for n in (0, 1, 2, 3, 1000):
        kk = rng.integers(0, 7, n).astype(np.uint32)
        assert np.array_equal(O.sort_bitonic(kk, np.arange(n, dtype=np.uint32))[0], np.sort(kk))


def test_cellindices_searchsorted():
    rng = np.random.default_rng(2)
    C = 1000
    k = np.sort(rng.integers(0, C + 1, 3000)).astype(np.uint32)   # includes the sentinel C
    start, end, nc = O.cellindices(k, C)
    for c in range(C + 1):
        a, b = np.searchsorted(k, c, "left"), np.searchsorted(k, c, "right")
        if a == b:
            assert start[c] == 0xFFFFFFFF
            assert end[c] == 666                                   # never written (stale in the reference)
        else:
            assert (start[c], end[c]) == (a, b)
    assert nc == (np.searchsorted(k, C) if np.any(k == C) else -1)


# ---------------------------------------------------------------- density
def test_density_isolated_particle():
    sp, gp, gs = params()
    pos = np.array([[0.5, 0.5, 0.5, 1.0]], np.float32)
    k, i, start, end, nc, ps, vs, es = one_step_front(pos, sp, gp, gs)
    d = O.density(ps, start, end, sp, gs)
    want = 315 * M / (64 * math.pi * H ** 3)
    assert abs(d[0] - want) <= 1e-5 * want


def test_density_simple_cubic_lattice():
    """Interior site of a lattice with spacing h/2: 33 neighbours within h,
    sum of (4 - n)^3 over |(i,j,k)|^2 = n <= 4 is 330."""
    sp, gp, gs = params()
    g = np.arange(-4, 5) * (H / 2) + 0.5
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    pos = np.c_[X.ravel(), Y.ravel(), Z.ravel(), np.ones(X.size)].astype(np.float32)
    k, i, start, end, nc, ps, vs, es = one_step_front(pos, sp, gp, gs)
    d = O.density(ps, start, end, sp, gs)
    centre = np.flatnonzero(np.all(np.abs(ps[:, :3] - 0.5) < 1e-6, axis=1))[0]
    want = 315 * 330 * M / (4096 * math.pi * H ** 3)
    assert abs(d[centre] - want) <= 2e-5 * want
    assert 1009 < want < 1011


def _brute(pos64, vel64, sp):
    """O(N^2) density, force and XSPH with no cell grid (FP64)."""
    h = float(sp.smoothing_distance)
    m = float(sp.mass)
    p = pos64[:, :3]
    r = p[:, None, :] - p[None, :, :]
    d2 = (r ** 2).sum(-1)
    nb = d2 <= h * h
    w = np.where(nb, (h * h - d2) ** 3, 0.0)
    rho = float(sp.wpoly6_coef) * m * w.sum(1)
    return rho, r, d2, nb


def test_brute_force_density_and_force():
    sc = scenes.dam_custom(10, 9, 8, seed=11)
    sp, gp, gs = O.derive_params(sc.settings)
    rng = np.random.default_rng(4)
    pos = sc.pos.copy()
    pos[:, :3] += rng.normal(scale=0.004, size=(len(pos), 3)).astype(np.float32)
    ve = np.zeros_like(pos)
    ve[:, :3] = rng.normal(scale=0.2, size=(len(pos), 3)).astype(np.float32)
    p64, v64 = pos.astype(np.float64), ve.astype(np.float64)
    k, i, start, end, nc, ps, vs, es = one_step_front(p64, sp, gp, gs, v64, v64)
    rho = O.density(ps, start, end, sp, gs)
    F, X = O.force(ps, rho, es, start, end, sp, gs)
    # brute force in the sorted order
    brho, r, d2, nb = _brute(ps, es, sp)
    np.testing.assert_allclose(rho, brho, rtol=1e-12)
    # Mueller et al. 2003 with a_i = f_i / rho_i:
    #   pressure  f_i = sum_j m (P_i + P_j) / (2 rho_j) * 45/(pi h^6) (h - r)^2 rhat
    #   viscosity f_i = mu sum_j m (v_j - v_i) / rho_j * 45/(pi h^6) (h - r)
    # XSPH (Monaghan): x_i = sum_j 2 m (v_j - v_i) / (rho_i + rho_j) * 315/(64 pi h^9) (h^2 - r^2)^3
    h, m, mu, Kg = H, float(sp.mass), float(sp.viscosity), float(sp.K)
    P = Kg * (brho - 1000.0)
    rl = np.sqrt(d2)
    off = nb & ~np.eye(len(ps), dtype=bool)
    rhat = np.where(off[..., None], r / np.where(rl > 0, rl, 1)[..., None], 0)
    c_sp = 45 / (math.pi * h ** 6)
    fp = (m * (P[:, None] + P[None, :]) / (2 * brho[None, :]) * c_sp * (h - rl) ** 2)[..., None] * rhat
    dv = es[None, :, :3] - es[:, None, :3]
    fv = (mu * m / brho[None, :] * c_sp * (h - rl))[..., None] * dv * off[..., None]
    acc = (fp.sum(1) + fv.sum(1)) / brho[:, None]
    c6 = 315 / (64 * math.pi * h ** 9)
    xs = ((2 * m / (brho[:, None] + brho[None, :]) * c6 * np.clip(h * h - d2, 0, None) ** 3)[..., None] * dv
          * off[..., None]).sum(1)
    # coefficients pass the 6-digit settings round trip: 1e-5 relative
    assert np.abs(F[:, :3] - acc).max() <= 1e-5 * np.abs(acc).max()
    assert np.abs(X[:, :3] - xs).max() <= 1e-5 * np.abs(xs).max()


def test_pair_terms_are_antisymmetric():
    """With equal masses the pressure, viscosity and XSPH pair terms cancel in
    the sum over all particles (momentum conservation)."""
    sc = scenes.dam_custom(12, 12, 6, seed=5)
    sp, gp, gs = O.derive_params(sc.settings)
    rng = np.random.default_rng(6)
    pos = sc.pos.astype(np.float64)
    pos[:, :3] += rng.normal(scale=0.004, size=(len(pos), 3))
    ve = np.zeros_like(pos)
    ve[:, :3] = rng.normal(scale=0.3, size=(len(pos), 3))
    k, i, start, end, nc, ps, vs, es = one_step_front(pos, sp, gp, gs, ve, ve)
    rho = O.density(ps, start, end, sp, gs)
    F, X = O.force(ps, rho, es, start, end, sp, gs)
    assert np.abs(F[:, :3].sum(0)).max() <= 1e-10 * np.abs(F[:, :3]).sum()
    assert np.abs(X[:, :3].sum(0)).max() <= 1e-10 * np.abs(X[:, :3]).sum()


def test_two_particles_pressure_pushes_apart():
    sp, gp, gs = params()
    pos = np.array([[0.5, 0.5, 0.5, 1], [0.51, 0.5, 0.5, 1]], np.float64)
    k, i, start, end, nc, ps, vs, es = one_step_front(pos, sp, gp, gs)
    rho = np.array([1200.0, 1300.0])          # compressed: positive pressures
    F, X = O.force(ps, rho, es, start, end, sp, gs)
    a = np.flatnonzero(ps[:, 0] < 0.505)[0]
    b = 1 - a
    assert F[a, 0] < 0 < F[b, 0]
    # closed form for particle a (Mueller pressure term, a = f / rho_a)
    r = 0.01
    want = 0.008 * (3 * 200 + 3 * 300) / (2 * rho[b]) * 45 / (math.pi * H ** 6) * (H - r) ** 2 / rho[a]
    assert abs(-F[a, 0] - want) <= 1e-5 * want


# ---------------------------------------------------------------- walls and integration
def test_wall_repulsion_and_friction_quirk():
    sp, gp, gs = params()
    bd = 0.5 * float(sp.rest_distance)
    pos = np.array([[0.5, 0.5, 0.004, 1], [0.5, 0.5, 0.004, 1], [0.5, 0.5, 0.5, 1]], np.float64)
    vel = np.array([[0, 0, 0, 0], [0.3, 0.2, -0.1, 0], [1, 1, 1, 0]], np.float64)
    f = O.collision_wall(pos, vel, np.zeros_like(pos), sp, gs)
    diff = bd - 0.004
    assert abs(f[0, 2] - 20000 * diff) <= 1e-9 and f[0, 0] == 0 and f[0, 1] == 0
    # moving particle: damping -256 * (n . v) and friction -v_t with v_n = v (n . v)
    vz = -0.1
    vt = np.array([0.3, 0.2, -0.1]) - np.array([0.3, 0.2, -0.1]) * vz
    assert np.mean(vt) > 0
    want = np.array([0, 0, 20000 * diff - 256 * vz]) - vt
    np.testing.assert_allclose(f[1, :3], want, rtol=1e-12)
    assert np.all(f[2] == 0)                               # far from every wall


def test_gravity_point_sources():
    """gravity.cl:9-36 closed forms: outside the cutoff (0.5 scale) a source
    pulls with alpha M / d^2 (d in scaled units) along the unit direction;
    inside it pushes away with alpha M / (0.5 scale)^2; sources superpose
    onto the incoming acceleration.  scale = 2 catches a missing factor."""
    src = np.array([[1.0, 0.0, 0.0, 1.0], [0.0, 0.0, 5.0, 1.0]], np.float32)
    mass = np.array([3.0, 2.0], np.float32)
    alpha = np.array([0.5, 0.25], np.float32)
    pos = np.array([[0.0, 0.0, 0.0, 1.0], [0.9, 0.0, 0.0, 1.0]], np.float32)
    f0 = np.array([[0.1, 0.2, 0.3, 0.0], [0.0, 0.0, 0.0, 0.0]], np.float32)
    s = 2.0
    for dt in (np.float32, np.float64):
        f = O.gravity(pos.astype(dt), f0.astype(dt), src, mass, alpha, s)
        # particle 0: source 0 at scaled distance 2 (outside), source 1 at 10 (outside)
        e0 = f0[0, :3] + 1.5 / 4.0 * np.array([1, 0, 0]) + 0.5 / 100.0 * np.array([0, 0, 1])
        # particle 1: source 0 at scaled distance 0.2 < cutoff 1 -> pushed away with 1.5 / 1
        d1 = 2.0 * (src[1, :3] - pos[1, :3])
        e1 = -1.5 * np.array([1, 0, 0]) + 0.5 / np.dot(d1, d1) * d1 / np.linalg.norm(d1)
        np.testing.assert_allclose(f[0, :3], e0, rtol=1e-6)
        np.testing.assert_allclose(f[1, :3], e1, rtol=1e-6)


def test_free_fall_and_clamp():
    st = scenes.dam_settings(1.0, 16)
    sp, gp, gs = O.derive_params(st)
    dt = 0.003
    pos = np.array([[0.5, 0.5, 0.8, 1.0]], np.float64)
    s = O.State(pos, np.zeros_like(pos), sp, gp, gs, dtype=np.float64)
    g = float(sp.gravity.z)          # -9.8 as FP32
    dt = float(np.float32(dt))
    for n in range(1, 11):
        s.step(dt, 0)
        assert abs(s.vel_u[0, 2] - g * n * dt) <= 1e-12
        assert abs(s.pos_u[0, 2] - (0.8 + g * dt * dt * n * (n + 1) / 2)) <= 1e-12
        np.testing.assert_allclose(s.veleval_u[0, 2], g * dt * (n - 0.5), rtol=1e-12)
    # velocity clamp: |f| > 600 is scaled to 600
    f = np.array([[0, 0, 1e4, 0]], np.float64)
    pu, vu, eu = O.leapfrog(pos, np.zeros_like(pos), f, np.zeros_like(pos), np.zeros(1, np.uint32), sp, dt,
                            np.zeros_like(pos), np.zeros_like(pos), np.zeros_like(pos))
    assert abs(vu[0, 2] - 600 * dt) <= 1e-9


def test_euler_free_fall_ignores_settings_gravity_and_limit():
    """euler.cl:33-78: gravity is a hard-coded -9.8 on z whatever the settings
    say, the clamp a hard-coded 600 (velocity_limit is not read), no XSPH, and
    veleval is left as it was.  Closed forms of semi-implicit Euler free fall:
    v_n = -9.8 n dt, p_n = p_0 - 9.8 dt^2 n (n + 1) / 2."""
    st = scenes.dam_settings(1.0, 16)
    st["gravity"] = "2 0 -3 0"
    st["velocity_limit"] = "0.001"
    st["integrator"] = "euler"
    sp, gp, gs = O.derive_params(st)
    dt = float(np.float32(0.003))
    pos = np.array([[0.5, 0.5, 0.8, 1.0]], np.float64)
    ve0 = np.array([[0.1, 0.2, 0.3, 0.0]], np.float64)
    s = O.State(pos, np.zeros_like(pos), sp, gp, gs, dtype=np.float64, veleval=ve0)
    g = float(np.float32(-9.8))
    for n in range(1, 11):
        s.step(dt, O.FLAG_EULER)
        assert s.vel_u[0, 0] == 0.0 and s.pos_u[0, 0] == 0.5
        assert abs(s.vel_u[0, 2] - g * n * dt) <= 1e-12
        assert abs(s.pos_u[0, 2] - (0.8 + g * dt * dt * n * (n + 1) / 2)) <= 1e-12
        np.testing.assert_array_equal(s.veleval_u, ve0)
    # the same particle under leapfrog follows the settings' gravity instead
    s2 = O.State(pos, np.zeros_like(pos), sp, gp, gs, dtype=np.float64)
    s2.step(dt, 0)
    assert s2.vel_u[0, 0] != 0.0
    # clamp at 600 on |f.xyz| (after the -9.8), scaling f.w too; none at 100
    z = np.zeros_like(pos)
    f = np.array([[0.0, 0.0, 1e4, 5.0]], np.float64)
    pu, vu = O.euler(pos, z, f, np.zeros(1, np.uint32), sp, dt, z, z)
    k = 600.0 / (1e4 + g)
    assert abs(vu[0, 2] - 600.0 * dt) <= 1e-9 and abs(vu[0, 3] - 5.0 * k * dt) <= 1e-12
    f = np.array([[0.0, 100.0, 0.0, 0.0]], np.float64)
    pu, vu = O.euler(pos, z, f, np.zeros(1, np.uint32), sp, dt, z, z)
    assert abs(vu[0, 1] - 100.0 * dt) <= 1e-12 and abs(vu[0, 2] - g * dt) <= 1e-12
    assert pu[0, 3] == 1.0


def test_deletion_path():
    sc = scenes.dam_custom(8, 8, 8, seed=2)
    sp, gp, gs = O.derive_params(sc.settings)
    pos = sc.pos.copy()
    pos[::10, 0] = 50.0
    s = O.State(pos, sc.vel, sp, gp, gs)
    d = s.step(sc.dt, O.FLAG_COLLISION)
    assert d == len(pos[::10]) and s.num == len(pos) - d
    # survivors keep their physics: same as a system built from the in-grid particles only
    keep = np.ones(len(pos), bool)
    keep[::10] = False
    s2 = O.State(pos[keep], sc.vel[keep], sp, gp, gs)
    s2.step(sc.dt, O.FLAG_COLLISION)
    a = s.pos_u[: s.num]
    b = s2.pos_u
    np.testing.assert_allclose(np.sort(a[:, 0]), np.sort(b[:, 0]), rtol=1e-6)


def test_fp32_tracks_fp64():
    sc = scenes.dam16k()
    sp, gp, gs = O.derive_params(sc.settings)
    a = O.State(sc.pos, sc.vel, sp, gp, gs, dtype=np.float32)
    b = O.State(sc.pos, sc.vel, sp, gp, gs, dtype=np.float64)
    for _ in range(5):
        a.step(sc.dt, O.FLAG_COLLISION)
        b.step(sc.dt, O.FLAG_COLLISION)
    np.testing.assert_allclose(a.pos_u[:, :3], b.pos_u[:, :3], rtol=1e-5)
    assert np.abs(a.density - b.density).max() <= 1e-4 * np.abs(b.density).max()


def _tl_vs_fd(pos, vel, st, steps, origin_note):
    """Forward-mode tangents of the FP64 step (TLState) against FP64 central
    differences along the same directions: one random initial-velocity
    direction, the viscosity and the gas constant (their FP32 parameter
    values perturbed by +-1e-4 relative, divided by the FP32 difference)."""
    sp, gp, gs = O.derive_params(st)
    n = len(pos)
    rng = np.random.default_rng(3)
    u = np.zeros((3, n, 4))
    u[0, :, :3] = rng.normal(size=(n, 3))
    dt = float(np.float32(st["time_step"]))
    tl = O.TLState(pos.astype(np.float64), vel.astype(np.float64), sp, gp, gs, dvel0=u, dmu=[0, 1, 0],
                   dK=[0, 0, 1])
    for _ in range(steps):
        tl.step(dt, O.FLAG_COLLISION)

    def run(v, mu, K):
        sp2, _, _ = O.derive_params(st)
        sp2.viscosity, sp2.K = mu, K
        s = O.State(pos.astype(np.float64), v, sp2, gp, gs, dtype=np.float64)
        for _ in range(steps):
            s.step(dt, O.FLAG_COLLISION)
        return s.pos_u.copy()

    v = vel.astype(np.float64)
    # the tangent state's values are the FP64 oracle's, bit for bit
    np.testing.assert_array_equal(tl.pos_u[:, :, 0], run(v, sp.viscosity, sp.K))
    d = 1e-6
    fd = (run(v + d * u[0], sp.viscosity, sp.K) - run(v - d * u[0], sp.viscosity, sp.K)) / (2 * d)
    assert np.abs(fd - tl.pos_u[:, :, 1]).max() <= 1e-6 * np.abs(fd).max(), origin_note
    for k, name in ((2, "viscosity"), (3, "K")):
        base = sp.viscosity if name == "viscosity" else sp.K
        hi, lo = float(np.float32(base * (1 + 1e-4))), float(np.float32(base * (1 - 1e-4)))
        a = run(v, hi, sp.K) if name == "viscosity" else run(v, sp.viscosity, hi)
        b = run(v, lo, sp.K) if name == "viscosity" else run(v, sp.viscosity, lo)
        fd = (a - b) / (hi - lo)
        assert np.abs(fd - tl.pos_u[:, :, k]).max() <= 1e-5 * np.abs(fd).max(), (origin_note, name)


def test_tangent_linear_step_matches_central_differences():
    """Pins the forward-mode oracle (the reference for the adjoint's
    dot-product tests) where the step is smooth along the directions: a block
    falling freely (10 steps) and a block resting on the floor and two walls
    (wall penalty, damping and friction active; 10 steps)."""
    pos = scenes.lattice(8, 8, 4, 21, origin=(0.34, 0.34, 0.45))
    vel = np.zeros_like(pos)
    vel[:, :3] = scenes.normal(22, 3 * len(pos), 0.05).reshape(-1, 3)
    _tl_vs_fd(pos, vel, scenes.dam_settings(1.0, len(pos), dt=0.002), 10, "free block")
    pos = scenes.lattice(8, 8, 5, 23, origin=(0.0, 0.0, 0.0))
    vel = np.zeros_like(pos)
    vel[:, :3] = scenes.normal(24, 3 * len(pos), 0.05).reshape(-1, 3)
    _tl_vs_fd(pos, vel, scenes.dam_settings(0.5, len(pos), dt=0.002), 10, "walls")


def test_add_rect_lattice_closed_form():
    """addRect (IV.cpp:165-202): count per axis = floor((max - min) / s) + 1
    where the box is not within rounding of a whole number of spacings,
    coordinates min + k s within float accumulation error (k ulps), z outer /
    x inner order, w = 1, and the count capped at num part way through a row."""
    lo, hi, sp = (0.1, 0.2, 0.3), (0.33, 0.41, 0.51), 0.02
    per = [int(np.floor((h - l) / sp)) + 1 for l, h in zip(lo, hi)]
    assert all(abs((h - l) / sp - round((h - l) / sp)) > 1e-3 for l, h in zip(lo, hi))
    full = per[0] * per[1] * per[2]
    p = O.add_rect(10 ** 6, lo, hi, sp)
    assert len(p) == full
    k = np.arange(full)
    ix, iy, iz = k % per[0], (k // per[0]) % per[1], k // (per[0] * per[1])
    for c, idx in ((0, ix), (1, iy), (2, iz)):
        exact = lo[c] + idx * sp
        assert np.all(np.abs(p[:, c] - exact) <= (idx + 2) * 2e-8 + 1e-7), c
    assert np.all(p[:, 3] == 1.0)
    # float accumulation, not multiplication: x_k = fl(x_{k-1} + s)
    x = np.float32(lo[0])
    for j in range(per[0]):
        assert p[j, 0] == x
        x = np.float32(x + np.float32(sp))
    q = O.add_rect(per[0] + 3, lo, hi, sp)     # stops three points into the second row
    np.testing.assert_array_equal(q, p[: per[0] + 3])


def _rigid_setup(st, bpos):
    sp, gp, gs = O.derive_params(st)
    keys, idx = O.hash(bpos, gp)
    keys, idx = O.sort_stable(keys, idx)
    start, end, _ = O.cellindices(keys, gp.nb_cells)
    return sp, gp, gs, bpos[idx], idx, start, end


@pytest.mark.parametrize("dynamic", [False, True])
def test_rigid_force_two_particles_closed_form(dynamic):
    """One fluid particle and one boundary particle at r < h on +x:
    F = -s0 m (2h - r) n + s1 sqrt(s0 m mn) (vrel . n) n, mn = m m_j / (m + m_j)
    (m / 2 for the static set), force += F / m -+ m / rho^2 mu c V with
    V = w (h - r): w = vrel (static, minus sign) or its tangential part
    (dynamic, plus sign); vrel = v_j - e_i (static: -e_i)."""
    st = scenes.dam_settings(1.0, 16)
    s0, s1 = 3.0, 0.5
    pos = np.array([[0.5, 0.5, 0.5, 1.0]])
    e = np.array([[0.3, -0.2, 0.1, 0.0]])
    r = 0.013
    bpos = np.array([[0.5 + r, 0.5, 0.5, 1.0]], np.float32)
    bvel = np.array([[0.05, 0.4, 0.0, 0.0]], np.float32)
    bmass = np.array([0.02], np.float32)
    sp, gp, gs, bps, idx, start, end = _rigid_setup(st, bpos)
    rho = np.array([950.0])
    f0 = np.zeros((1, 4))
    f = O.rigid_force(pos, e, rho, f0, bps, start, end, sp, gs, s0, s1,
                      *((bvel, bmass) if dynamic else ()))
    m, mu, c, h = sp.mass, sp.viscosity, sp.wvisc_dd_coef, sp.smoothing_distance
    rr = float(np.float32(0.5 + r)) - 0.5
    n = np.array([1.0, 0.0, 0.0])
    mj = float(bmass[0]) if dynamic else m
    mn = m * mj / (m + mj)
    vrel = (bvel[0, :3].astype(np.float64) if dynamic else 0.0) - e[0, :3]
    F = -s0 * m * (2 * h - rr) * n + s1 * np.sqrt(s0 * m * mn) * (vrel @ n) * n
    w = vrel - (vrel @ n) * n if dynamic else vrel
    V = w * (h - rr)
    expect = F / m + (1 if dynamic else -1) * m / rho[0] ** 2 * mu * c * V
    np.testing.assert_allclose(f[0, :3], expect, rtol=1e-9, atol=1e-12)


def test_rigid_force_brute_force():
    """Against an O(N M) loop over every (fluid, boundary) pair accepted by
    the reference's rule (boundary cell within +-1 of the fluid particle's
    cell on each axis, |r| <= 2h), in FP64, both forms."""
    st = scenes.dam_settings(0.6, 4096)
    rng = np.random.default_rng(12)
    pos = np.c_[rng.uniform(0.05, 0.55, (600, 3)), np.ones(600)]
    e = np.c_[rng.normal(0, 0.1, (600, 3)), np.zeros(600)]
    rho = rng.uniform(800, 1200, 600)
    bpos = np.c_[rng.uniform(0.0, 0.6, (900, 3)), np.ones(900)].astype(np.float32)
    bvel = np.c_[rng.normal(0, 0.1, (900, 3)), np.zeros(900)].astype(np.float32)
    bmass = rng.uniform(0.005, 0.02, 900).astype(np.float32)
    sp, gp, gs, bps, idx, start, end = _rigid_setup(st, bpos)
    h, eps, m = sp.smoothing_distance, sp.EPSILON, sp.mass
    s0, s1 = 2.0, 0.75
    cell = lambda p: np.trunc((p[:, :3] - [gs.grid_min.x, gs.grid_min.y, gs.grid_min.z])  # noqa: E731
                              * [gs.grid_delta.x, gs.grid_delta.y, gs.grid_delta.z]).astype(int)
    ci, cj = cell(pos), cell(bpos.astype(np.float64))
    for dynamic in (False, True):
        got = O.rigid_force(pos, e, rho, np.zeros((600, 4)), bps, start, end, sp, gs, s0, s1,
                            *((bvel[idx], bmass[idx]) if dynamic else ()))
        ref = np.zeros((600, 3))
        for i in range(600):
            ok = np.all(np.abs(cj - ci[i]) <= 1, axis=1)
            r = bpos[ok, :3].astype(np.float64) - pos[i, :3]
            rl = np.linalg.norm(r, axis=1)
            acc = rl <= 2 * h
            r, rl = r[acc], np.maximum(rl[acc], eps)
            n = r / rl[:, None]
            mj = bmass[ok][acc].astype(np.float64) if dynamic else np.full(len(rl), m)
            mn = m * mj / (m + mj)
            vrel = (bvel[ok][acc, :3].astype(np.float64) if dynamic else 0.0) - e[i, :3]
            un = np.sum(vrel * n, axis=1)
            F = (-s0 * m * (2 * h - rl))[:, None] * n + (s1 * np.sqrt(s0 * m * mn) * un)[:, None] * n
            w = vrel - un[:, None] * n if dynamic else vrel + 0 * n
            V = np.sum(w * (h - rl)[:, None], axis=0)
            ref[i] = F.sum(0) / m + (1 if dynamic else -1) * m / rho[i] ** 2 * sp.viscosity * sp.wvisc_dd_coef * V
        assert np.count_nonzero(np.abs(ref).sum(1)) > 300
        np.testing.assert_allclose(got[:, :3], ref, rtol=1e-9, atol=1e-9 * np.abs(ref).max())
