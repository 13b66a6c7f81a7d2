"""CPU restatement of the colour field and the iso-surface, for tests only.

Test infrastructure: only tests/ may import this module, as the checker of
rtps_sph_color_field / rtps_sph_surface.

  color_field     colorfield.cl:55-125 (ForNeighbor, colorfield_update) with
                  Wspiky (cl_sph_kernels.h:71-78), brute force over all
                  particles instead of the 27-cell loop (the same sum: the
                  cells cover every particle within h of a node)
  CORNERS         the lattice offset of case bit k: classifycubes.cl:44-55
                  reads the nodes cubeOffsets[0, 1, 3, 2, 4, 5, 7, 6]
                  (indexmap.h:29-38) into bits 0..7
  EDGE_ENDS       the two ends of edge e, traversehp.cl:35-72 (offsets3)
  marching_cubes  classifycubes.cl:44-57 (cases), traversehp.cl:470-515
                  (edge interpolation from the first end to the second,
                  normals from negative central differences), with the case
                  table passed in (the library's, whose digest equals the
                  reference's triTable), cubes in lattice order; vertices in
                  the box's world coordinates or the reference's frame
                  (lattice / (res - 1) * 10)
"""
from __future__ import annotations

import math

import numpy as np

f32 = np.float32

CUBE_OFFSETS = [(0, 0, 0), (1, 0, 0), (0, 0, 1), (1, 0, 1), (0, 1, 0), (1, 1, 0), (0, 1, 1), (1, 1, 1)]
CORNERS = [CUBE_OFFSETS[i] for i in (0, 1, 3, 2, 4, 5, 7, 6)]
EDGE_ENDS = [((0, 0, 0), (1, 0, 0)), ((1, 0, 0), (1, 0, 1)), ((1, 0, 1), (0, 0, 1)), ((0, 0, 1), (0, 0, 0)),
             ((0, 1, 0), (1, 1, 0)), ((1, 1, 0), (1, 1, 1)), ((1, 1, 1), (0, 1, 1)), ((0, 1, 1), (0, 1, 0)),
             ((0, 0, 0), (0, 1, 0)), ((1, 0, 0), (1, 1, 0)), ((1, 0, 1), (1, 1, 1)), ((0, 0, 1), (0, 1, 1))]


def corner_of(offset):
    """Case bit of the cube corner at a lattice offset."""
    return CORNERS.index(tuple(offset))


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


def marching_cubes(field, iso, table, wmin, wmax, frame=0):
    """Triangles of the iso-surface of a [z, y, x] float32 lattice, cubes in
    lattice order (x fastest): (vertices, normals, edge keys), each (T, 3, ...).
    Edge keys name the lattice edge a vertex lies on (for topology checks).
    frame 0: world coordinates of [wmin, wmax]; 1: lattice / (res - 1) * 10."""
    tri, ntri = table
    f = np.asarray(field, np.float32)
    res = f.shape[0]
    m = res - 1
    cases = np.zeros((m, m, m), np.int64)
    for k, (dx, dy, dz) in enumerate(CORNERS):
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
                a, b = EDGE_ENDS[tri[cs, 3 * t + v]]
                pa = (x + a[0], y + a[1], z + a[2])
                pb = (x + b[0], y + b[1], z + b[2])
                v0, v1 = val(*pa), val(*pb)
                d = (f32(iso) - v0) / (v1 - v0)
                g = np.array([f32(pa[i]) + d * f32(pb[i] - pa[i]) for i in range(3)], np.float32)
                n0, n1 = ngrad(*pa), ngrad(*pb)
                n = n0 + d * (n1 - n0)
                ln = float(np.sqrt((n.astype(np.float64) ** 2).sum()))
                if frame == 1:
                    tv.append(g / f32(res - 1) * f32(10.0))
                else:
                    tv.append(wmin + g * inv * (wmax - wmin))
                tn.append(n / ln if ln > 0 else n * 0)
                tk.append((pa, pb) if pa < pb else (pb, pa))
            V.append(tv)
            Nn.append(tn)
            K.append(tk)
    return (np.array(V, np.float32).reshape(-1, 3, 3), np.array(Nn, np.float32).reshape(-1, 3, 3), K)
