"""Particle hoses (rtps_hose_*, host code; no GPU needed).

The library's hose is checked bit for bit against the restatement in
oracle/hose_oracle.py (Hose.cpp:30-128, IV.cpp:313-365, same SplitMix64
stream).  The restatement itself is pinned by what the algorithm fixes
without retyping it: SplitMix64's published outputs, the disc axes being
orthonormal and perpendicular to the velocity, every particle inside the
disc radius and within 10 |v| dt of the disc plane, the spray cadence
em = 1 + spacing / dt / |v| / 10, a disc of about pi r^2 / spacing^2
particles, and the particle budget (total_n, refill).
"""
import math

import numpy as np
import pytest

import hose_oracle as O

HOSES = [  # dt, total_n, center, velocity, radius, spacing, seed
    (0.003, 400, (1.0, 1.2, 0.8, 1.0), (0.0, 0.0, -2.0, 0.0), 0.12, 0.025, 1),
    (0.001, 1000, (0.5, 0.5, 0.5, 1.0), (1.0, -0.5, 0.3, 0.0), 0.2, 0.02, 2),
    (0.002, 50, (0.0, 0.3, 0.1, 1.0), (3.0, 0.0, 0.0, 0.0), 0.1, 0.015, 3),   # v.y = v.z = 0: guarded axes
    (0.0005, 300, (2.0, 2.0, 2.0, 1.0), (0.2, 0.7, 0.0, 0.0), 0.06, 0.01, 4),
]


@pytest.fixture(scope="module")
def rtps():
    from enjaparticles_amd import rtps
    return rtps


def test_splitmix64_published_outputs():
    r = O.SplitMix64(0)
    assert [r.next() for _ in range(3)] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    r = O.SplitMix64(1234567)
    assert [r.next() for _ in range(5)] == [6457827717110365317, 3203168211198807973, 9817491932198370423,
                                           4593380528125082431, 16408922859458223821]
    r = O.SplitMix64(9)
    u = [r.uniform01() for _ in range(20000)]
    assert 0.0 <= min(u) and max(u) < 1.0 and abs(np.mean(u) - 0.5) < 0.01


@pytest.mark.parametrize("spec", HOSES)
def test_library_hose_equals_restatement(rtps, spec):
    dt, n, c, v, r, s, seed = spec
    lib_h = rtps.Hose(dt, n, c, v, r, s, seed)
    ref = O.Hose(dt, n, c, v, r, s, seed)
    st = lib_h.state()
    np.testing.assert_array_equal(st["u"], np.array(ref.u, np.float32))
    np.testing.assert_array_equal(st["w"], np.array(ref.w, np.float32))
    np.testing.assert_array_equal(st["velocity"], np.array(ref.velocity, np.float32))
    sprays = 0
    for call in range(60):
        if call == 25:   # move and refill mid-way
            lib_h.update((c[0] + 0.1, c[1], c[2], 1.0), (v[0], v[1] + 0.5, v[2], 0.0), r * 1.5, s)
            ref.update((c[0] + 0.1, c[1], c[2], 1.0), (v[0], v[1] + 0.5, v[2], 0.0), r * 1.5, s)
        if call == 40:
            lib_h.refill(n // 2)
            ref.refill(n // 2)
        a = lib_h.spray()
        b = np.array(ref.spray(), np.float32).reshape(-1, 4)
        np.testing.assert_array_equal(a, b, err_msg=f"call {call}")
        st = lib_h.state()
        assert (st["n_left"], st["em_count"], st["em"]) == (ref.n_count, ref.em_count, ref.em)
        sprays += len(a) > 0
    assert sprays >= 2
    lib_h.close()


@pytest.mark.parametrize("spec", HOSES)
def test_disc_geometry(spec):
    dt, n, c, v, r, s, seed = spec
    h = O.Hose(dt, 10 ** 6, c, v, r, s, seed)
    vel = np.array(h.velocity[:3], np.float64)
    u, w = np.array(h.u[:3], np.float64), np.array(h.w[:3], np.float64)
    vhat = vel / np.linalg.norm(vel)
    # axes: unit, perpendicular to each other and to the velocity
    for a in (u, w):
        assert abs(np.linalg.norm(a) - 1) < 1e-6
        assert abs(a @ vhat) < 1e-5
    assert abs(u @ w) < 1e-5
    parts = []
    while not parts:
        parts = h.spray()
    p = np.array(parts, np.float64)[:, :3] - np.array(c[:3], np.float64)
    # inside the disc radius, and at most 10 |v| dt from its plane
    assert (np.linalg.norm(p, axis=1) < r * (1 + 1e-6)).all()
    assert (np.abs(p @ vhat) <= 10 * np.linalg.norm(vel) * dt * (1 + 1e-5)).all()
    # a perturbed spacing grid clipped to the disc: about pi r^2 / s^2 points,
    # fewer where the axial spread pushes points out of the ball
    full = math.pi * r * r / (s * s)
    assert 0.25 * full <= len(p) <= 1.3 * full + 8, (len(p), full)
    # in-plane coordinates of distinct grid points are about a spacing apart
    q = np.stack([p @ u, p @ w], 1)
    d = np.linalg.norm(q[:, None] - q[None], axis=2) + np.eye(len(q)) * 1e9
    assert d.min() > 0.5 * s


@pytest.mark.parametrize("spec", HOSES)
def test_cadence_and_budget(spec):
    dt, n, c, v, r, s, seed = spec
    h = O.Hose(dt, n, c, v, r, s, seed)
    em = O.em_of(s, dt, v)
    calls, emitted, fired = 0, 0, []
    while h.n_count > 0 and calls < 10000:
        calls += 1
        out = h.spray()
        assert h.em == em
        if out:
            fired.append(calls)
            emitted += len(out)
    # sprays come every em calls, and exactly total_n particles leave the hose
    assert fired == [em * (k + 1) for k in range(len(fired))]
    assert emitted == n and h.n_count == 0
    assert h.spray() == [] and h.n_count == 0
    h.refill(7)
    got = []
    for _ in range(em):
        got += h.spray()
    assert len(got) == 7 and h.n_count == 0   # the last spray stops at the budget


def test_arguments(rtps):
    from enjaparticles_amd import _native as N
    with pytest.raises(N.RTPSError):
        rtps.Hose(0.001, 10, (0, 0, 0), (0, 0, 0), 0.1, 0.02)   # no direction
    with pytest.raises(N.RTPSError):
        rtps.Hose(0.0, 10, (0, 0, 0), (1, 0, 0), 0.1, 0.02)
    h = rtps.Hose(0.001, 10, (0, 0, 0), (1, 0, 0), 0.1, 0.02)
    out = np.zeros((4, 4), np.float32)
    n = N.C.c_int()
    assert N.lib().rtps_hose_spray(h._h, out.ctypes.data, 4, N.C.byref(n)) == N.RTPS_ERR_CAPACITY
    with pytest.raises(N.RTPSError):
        h.refill(-1)
    h.close()
