"""CPU restatement of the colour field and the iso-surface, for tests only.

Test infrastructure: only tests/ may import this module, as the checker of
rtps_sph_color_field / rtps_sph_surface / rtps_mc_case_table.

  color_field     colorfield.cl:55-125 (ForNeighbor, colorfield_update) with
                  Wspiky (cl_sph_kernels.h:71-78), brute force over all
                  particles instead of the 27-cell loop (the same sum: the
                  cells cover every particle within h of a node)
  mc_case_table   the library's generated case table, restated: per face,
                  each run of above-iso corners is cut off by one segment,
                  directed with the run on its left seen from outside; the
                  segments chain into loops from their lowest edge, each
                  triangulated without a diagonal inside a cube face
  marching_cubes  classifycubes.cl:44-57 (cases), traversehp.cl:470-515
                  (edge interpolation, normals from negative central
                  differences) in cube order, vertices mapped to the box
"""
from __future__ import annotations

import math

import numpy as np

f32 = np.float32

EDGES = [(0, 1), (2, 3), (4, 5), (6, 7), (0, 2), (1, 3), (4, 6), (5, 7), (0, 4), (1, 5), (2, 6), (3, 7)]
FACES = [(0, 2, 6, 4), (1, 3, 7, 5), (0, 1, 5, 4), (2, 3, 7, 6), (0, 1, 3, 2), (4, 5, 7, 6)]


def corner(k):
    return np.array([k & 1, (k >> 1) & 1, (k >> 2) & 1], np.float64)


def _edge(a, b):
    return EDGES.index((min(a, b), max(a, b)))


def _faces_ccw():
    """Face corner cycles, counter-clockwise seen from outside the cube."""
    out = []
    for f in FACES:
        p = [corner(k) for k in f]
        n = np.cross(p[1] - p[0], p[2] - p[1])
        if n @ (np.mean(p, 0) - 0.5) < 0:
            f = (f[0], f[3], f[2], f[1])
        out.append(f)
    return out


def _share_face(e1, e2):
    (a1, b1), (a2, b2) = EDGES[e1], EDGES[e2]
    for axis in range(3):
        bit = 1 << axis
        if (a1 & bit) != (b1 & bit) or (a2 & bit) != (b2 & bit):
            continue
        if (a1 & bit) == (a2 & bit):
            return True
    return False


def _triangulate(poly):
    """Triangles of a loop polygon with no diagonal between two vertices on
    one cube face; side (p0, p_last) takes the first apex that works.
    Triangles (a, b, c) in loop order are returned as (a, c, b); None if
    impossible."""
    k = len(poly)
    if k < 3:
        return []
    if k == 3:
        return [(poly[0], poly[2], poly[1])]
    for j in range(1, k - 1):
        if j > 1 and _share_face(poly[0], poly[j]):
            continue
        if j + 2 < k and _share_face(poly[j], poly[k - 1]):
            continue
        left, right = _triangulate(poly[: j + 1]), _triangulate(poly[j:])
        if left is None or right is None:
            continue
        return left + right + [(poly[0], poly[k - 1], poly[j])]
    return None


def mc_case_table():
    """(tri[256, 16] int8 edge ids -1 padded, ntri[256] uint8)."""
    faces = _faces_ccw()
    tri = np.full((256, 16), -1, np.int8)
    ntri = np.zeros(256, np.uint8)
    for m in range(256):
        above = [(m >> k) & 1 for k in range(8)]
        nxt = {}
        for f in faces:
            if all(above[c] for c in f):
                continue
            for i in range(4):
                if not above[f[i]] or above[f[i - 1]]:
                    continue
                j = i
                while above[f[(j + 1) % 4]]:
                    j = (j + 1) % 4
                e_in, e_out = _edge(f[i - 1], f[i]), _edge(f[j], f[(j + 1) % 4])
                assert e_out not in nxt
                nxt[e_out] = e_in     # run of above-iso corners on the left, seen from outside
        seen, nt = set(), 0
        for e0 in sorted(nxt):
            if e0 in seen:
                continue
            loop, cur = [], e0
            while True:
                loop.append(cur)
                seen.add(cur)
                cur = nxt[cur]
                if cur == e0:
                    break
            tris = _triangulate(loop)
            assert tris is not None, m
            for t3 in tris:
                tri[m, 3 * nt: 3 * nt + 3] = t3
                nt += 1
        ntri[m] = nt
    return tri, ntri


def color_field(pos_scaled, dens, mass, h, eps, bmin, bmax, res, chunk=4096):
    """Field at the res^3 nodes, float64, indexed [z, y, x]."""
    p = np.asarray(pos_scaled, np.float64)[:, :3]
    rho = np.asarray(dens, np.float64)
    tmp = 1.0 / (res - 1)
    ax = [np.arange(res) * tmp * (bmax[k] - bmin[k]) + bmin[k] for k in range(3)]
    zz, yy, xx = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
    nodes = np.stack([xx.ravel(), yy.ravel(), zz.ravel()], 1)
    alpha = 15.0 / math.pi / h ** 6
    out = np.zeros(len(nodes))
    for c0 in range(0, len(nodes), chunk):
        x = nodes[c0:c0 + chunk]
        r2 = ((x[:, None, :] - p[None, :, :]) ** 2).sum(-1)
        rl = np.maximum(np.sqrt(r2), eps)
        w = np.where(r2 <= h * h, alpha * (h - rl) ** 3 / rho[None, :], 0.0)
        out[c0:c0 + chunk] = w.sum(1)
    return (out * mass).reshape(res, res, res)


def marching_cubes(field, iso, table, wmin, wmax):
    """Triangles of the iso-surface of a [z, y, x] float32 lattice, cube
    order x fastest: (vertices, normals, edge keys), each (T, 3, ...).  Edge
    keys name the lattice edge a vertex lies on (for topology checks)."""
    tri, ntri = table
    f = np.asarray(field, np.float32)
    res = f.shape[0]
    m = res - 1
    cases = np.zeros((m, m, m), np.int64)
    for k in range(8):
        dx, dy, dz = k & 1, (k >> 1) & 1, (k >> 2) & 1
        cases |= (f[dz:dz + m, dy:dy + m, dx:dx + m] > f32(iso)).astype(np.int64) << k
    wmin = np.asarray(wmin, np.float32)[:3]
    wmax = np.asarray(wmax, np.float32)[:3]
    inv = f32(1.0) / f32(res - 1)

    def val(x, y, z):
        return f[z, y, x]

    def ngrad(x, y, z):
        c = lambda v: min(max(v, 0), res - 1)   # noqa: E731
        return np.array([val(c(x - 1), y, z) - val(c(x + 1), y, z), val(x, c(y - 1), z) - val(x, c(y + 1), z),
                         val(x, y, c(z - 1)) - val(x, y, c(z + 1))], np.float32)

    V, Nn, K = [], [], []
    zs, ys, xs = np.nonzero(ntri[cases] > 0)
    order = np.lexsort((xs, ys, zs))
    for q in order:
        x, y, z = int(xs[q]), int(ys[q]), int(zs[q])
        cs = cases[z, y, x]
        for t in range(ntri[cs]):
            tv, tn, tk = [], [], []
            for v in range(3):
                a, b = EDGES[tri[cs, 3 * t + v]]
                pa = (x + (a & 1), y + ((a >> 1) & 1), z + ((a >> 2) & 1))
                pb = (x + (b & 1), y + ((b >> 1) & 1), z + ((b >> 2) & 1))
                v0, v1 = val(*pa), val(*pb)
                d = (f32(iso) - v0) / (v1 - v0)
                g = np.array([f32(pa[i]) + d * f32(pb[i] - pa[i]) for i in range(3)], np.float32)
                n0, n1 = ngrad(*pa), ngrad(*pb)
                n = n0 + d * (n1 - n0)
                ln = float(np.sqrt((n.astype(np.float64) ** 2).sum()))
                tv.append(wmin + g * inv * (wmax - wmin))
                tn.append(n / ln if ln > 0 else n * 0)
                tk.append((pa, pb) if pa < pb else (pb, pa))
            V.append(tv)
            Nn.append(tn)
            K.append(tk)
    return (np.array(V, np.float32).reshape(-1, 3, 3), np.array(Nn, np.float32).reshape(-1, 3, 3), K)
