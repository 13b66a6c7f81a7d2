"""Marching-cubes case table and iso-surface restatement (no GPU needed).

The library's case table (rtps_mc_case_table) and edge ends
(rtps_mc_geometry) are the reference's: their SHA-256 fingerprints equal
those of triTable (traversehp.cl:74), nrOfTriangles (classifycubes.cl:37)
and offsets3 (traversehp.cl:35) recorded in tests/golden/mc_tables.json
(tests/golden/make_mc_digest.py, run where the reference tree exists).  The
corner bit order equals the restatement in oracle/surface_oracle.py
(classifycubes.cl:44-55 with indexmap.h's cubeOffsets).  The table is also
checked for what a case table must satisfy: each case's triangles use exactly
its crossing edges, at most five, all wound alike; a sphere extracted with it
is closed with Euler characteristic 2 and a torus with 0 (no ambiguous faces
at these resolutions), vertices of a linear field lie on its plane and every
vertex of a random field interpolates the iso level on a crossing edge.
"""
import ctypes as C
import hashlib
import json
import os
from collections import Counter

import numpy as np
import pytest

import surface_oracle as S

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "mc_tables.json")


@pytest.fixture(scope="module")
def table():
    from enjaparticles_amd import _native as N
    tri = np.zeros((256, 16), np.int8)
    ntri = np.zeros(256, np.uint8)
    N.check(N.lib().rtps_mc_case_table(tri.ctypes.data, ntri.ctypes.data))
    return tri, ntri


def test_library_table_is_the_references(table):
    from enjaparticles_amd import _native as N
    g = json.load(open(GOLDEN))
    tri, ntri = table
    assert hashlib.sha256(tri.tobytes()).hexdigest() == g["triTable_int8_sha256"]
    assert hashlib.sha256(ntri.tobytes()).hexdigest() == g["nrOfTriangles_uint8_sha256"]
    assert int(ntri.astype(int).sum()) == g["triangles_total"]
    corners = np.zeros(24, np.int8)
    edges = np.zeros(72, np.int8)
    N.check(N.lib().rtps_mc_geometry(corners.ctypes.data, edges.ctypes.data))
    assert hashlib.sha256(edges.tobytes()).hexdigest() == g["offsets3_int8_sha256"]
    np.testing.assert_array_equal(corners.reshape(8, 3), np.array(S.CORNERS))
    np.testing.assert_array_equal(edges.reshape(12, 2, 3), np.array(S.EDGE_ENDS))


def test_cases_use_exactly_their_crossing_edges(table):
    tri, ntri = table
    assert ntri[0] == 0 and ntri[255] == 0 and ntri.max() == 5
    sign = set()
    for m in range(256):
        above = [(m >> k) & 1 for k in range(8)]
        crossing = {e for e, (a, b) in enumerate(S.EDGE_ENDS) if above[S.corner_of(a)] != above[S.corner_of(b)]}
        used = tri[m, : 3 * ntri[m]]
        assert set(used.tolist()) == crossing, m
        assert (tri[m, 3 * ntri[m]:] == -1).all()
        # winding: every triangle (vertices at edge midpoints) faces the same
        # way relative to its edges' above-iso ends
        for t in range(ntri[m]):
            es = used[3 * t: 3 * t + 3]
            p = [0.5 * (np.array(S.EDGE_ENDS[e][0]) + np.array(S.EDGE_ENDS[e][1])) for e in es]
            n = np.cross(p[1] - p[0], p[2] - p[0])
            d = sum((np.array(S.EDGE_ENDS[e][1]) - np.array(S.EDGE_ENDS[e][0]))
                    * (1 if above[S.corner_of(S.EDGE_ENDS[e][0])] else -1) for e in es)
            sign.add(bool(n @ d > 0))
    assert len(sign) == 1


def _closed(keys, res):
    """Every vertex edge inside the lattice is shared by exactly two
    triangles' sides, traversed in opposite directions."""
    sides = Counter()
    for tk in keys:
        for q in range(3):
            sides[(tk[q], tk[(q + 1) % 3])] += 1
    for (a, b), c in sides.items():
        assert c == 1, (a, b, c)
        interior = all(0 < v < res - 1 for v in a[0] + a[1] + b[0] + b[1])
        if interior:
            assert sides.get((b, a), 0) == 1, (a, b)
    return sides


def _euler(keys):
    verts = {k for tk in keys for k in tk}
    edges = {frozenset((tk[q], tk[(q + 1) % 3])) for tk in keys for q in range(3)}
    return len(verts) - len(edges) + len(keys)


def _lattice(res):
    ax = np.linspace(-1.0, 1.0, res)
    z, y, x = np.meshgrid(ax, ax, ax, indexing="ij")
    return x, y, z


def test_sphere_is_closed_with_outward_normals(table):
    res = 20
    x, y, z = _lattice(res)
    f = (0.7 - np.sqrt(x * x + y * y + z * z)).astype(np.float32)
    v, n, k = S.marching_cubes(f, 0.0, table, (-1, -1, -1), (1, 1, 1))
    _closed(k, res)
    assert _euler(k) == 2
    r = np.linalg.norm(v, axis=2)
    assert np.abs(r - 0.7).max() < 0.02
    assert ((n * v).sum(2) > 0).all()   # normals away from the inside (higher field)
    # the table winds every triangle alike (here: against the normals)
    g = np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0])
    w = (g * v.mean(1)).sum(1)
    assert (w < 0).all() or (w > 0).all()


def test_torus_is_closed(table):
    res = 28
    x, y, z = _lattice(res)
    q = np.sqrt(x * x + y * y) - 0.55
    f = (0.25 - np.sqrt(q * q + z * z)).astype(np.float32)
    _, _, k = S.marching_cubes(f, 0.0, table, (-1, -1, -1), (1, 1, 1))
    _closed(k, res)
    assert _euler(k) == 0


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_random_field_vertices_interpolate_the_iso_level(table, seed):
    """Every vertex lies on a lattice edge whose ends straddle the iso level,
    at the linearly interpolated crossing; each lattice edge gets one vertex
    position however many cubes share it."""
    res = 12
    f = np.random.default_rng(seed).normal(size=(res, res, res)).astype(np.float32)
    v, _, k = S.marching_cubes(f, 0.0, table, (0, 0, 0), (res - 1, res - 1, res - 1))
    assert len(k) > 100
    seen = {}
    for tv, tk in zip(v, k):
        for q in range(3):
            a, b = tk[q]
            fa, fb = f[a[2], a[1], a[0]], f[b[2], b[1], b[0]]
            assert (fa > 0) != (fb > 0)
            t = fa / (fa - fb)
            np.testing.assert_allclose(tv[q], np.array(a) + t * (np.array(b) - np.array(a)), atol=2e-5)
            assert np.allclose(seen.setdefault((a, b), tv[q]), tv[q], atol=2e-5)


def test_linear_field_gives_its_plane(table):
    res = 10
    x, y, z = _lattice(res)
    a = np.array([0.3, -0.5, 0.8])
    f = (a[0] * x + a[1] * y + a[2] * z).astype(np.float32)
    v, n, k = S.marching_cubes(f, 0.1, table, (-1, -1, -1), (1, 1, 1))
    assert len(v) > 0
    assert np.abs(v @ a - 0.1).max() < 1e-5
    # normals: the negative gradient wherever both central differences are
    # whole (edges off the lattice boundary; the clamped ones are one-sided)
    inner = np.array([all(0 < c < res - 1 for c in pa + pb) for tk in k for (pa, pb) in tk])
    assert inner.sum() > 100
    np.testing.assert_allclose(n.reshape(-1, 3)[inner], np.tile(-a / np.linalg.norm(a), (inner.sum(), 1)), atol=1e-5)


def test_color_field_restatement_closed_form():
    """One particle: the field is mass * Wspiky(r) / rho at every node within h."""
    h, mass, rho, eps = 0.04, 0.008, 1000.0, 1e-6
    p = np.array([[0.5, 0.5, 0.5, 1.0]])
    res = 9
    f = S.color_field(p, [rho], mass, h, eps, (0.46, 0.46, 0.46), (0.54, 0.54, 0.54), res)
    ax = np.linspace(0.46, 0.54, res)
    z, y, x = np.meshgrid(ax, ax, ax, indexing="ij")
    r = np.sqrt((x - 0.5) ** 2 + (y - 0.5) ** 2 + (z - 0.5) ** 2)
    want = np.where(r <= h, mass * 15 / np.pi / h ** 6 * (h - np.maximum(r, eps)) ** 3 / rho, 0.0)
    np.testing.assert_allclose(f, want, rtol=1e-12, atol=0)
    # at the particle: r = 0 clamps to eps
    assert f[4, 4, 4] == pytest.approx(mass * 15 / np.pi / h ** 6 * (h - eps) ** 3 / rho, rel=1e-12)
