"""Colour field and iso-surface on the GPU (rtps_sph_color_field,
rtps_sph_surface, rtps_sph_post_process) against surface_oracle.

Bars: the field within 1e-4 of its maximum of the FP64 brute-force sum
(same sorted positions and densities; FP32 sums over ~100 neighbours); the
surface of the GPU field equal to the restatement's with the reference's case
table (whose digest test_surface_cpu pins), triangle by triangle in cube
order (identical cases; vertices within 1e-5 of the box size, unit normals
within 1e-4), in world coordinates and in the reference's [0, 10]^3 frame.
"""
import numpy as np
import pytest

import oracle as O
import surface_oracle as SO
from enjaparticles_amd import scenes
from enjaparticles_amd.rtps import RTPS

pytestmark = pytest.mark.gpu


def _state(steps=30, seed=21):
    sc = scenes.dam_custom(16, 16, 8, seed=seed)
    s = RTPS.generateSystemInstance(sc.settings)
    s.pushParticles(sc.pos, sc.vel)
    s.step(steps)
    return sc, s


def _table():
    from enjaparticles_amd import _native as N
    tri = np.zeros((256, 16), np.int8)
    ntri = np.zeros(256, np.uint8)
    N.check(N.lib().rtps_mc_case_table(tri.ctypes.data, ntri.ctypes.data))
    return tri, ntri


def _box(g):
    return (g.bnd_min.x, g.bnd_min.y, g.bnd_min.z), (g.bnd_max.x, g.bnd_max.y, g.bnd_max.z)


def _derived_state(steps=10, seed=23):
    """A lattice with the reference's derived parameters (simulation scale
    0.31, not 1): the colour field's scaled-position path."""
    st = scenes.dam_settings(1.0, 16384)
    for k in ("mass", "rest_distance", "smoothing_distance", "simulation_scale"):
        del st[k]
    sp, _, _ = O.derive_params(st)
    pos = scenes.lattice(16, 16, 8, seed=seed, spacing=sp.rest_distance / sp.simulation_scale)
    sc = scenes.Scene("derived", st, pos, np.zeros_like(pos), steps, 0.003)
    s = RTPS.generateSystemInstance(sc.settings)
    s.pushParticles(sc.pos, sc.vel)
    s.step(steps)
    return sc, s


@pytest.mark.parametrize("derived", [False, True])
def test_color_field_matches_brute_force(cuda, derived):
    sc, s = _derived_state() if derived else _state()
    res = 24
    f = s.colorField(res)
    sp, gp, gs = s.params()
    n = s.getNum()
    pos = s.read("pos_s")[:n].astype(np.float64) * np.float64(np.float32(sp.simulation_scale))
    ref = SO.color_field(pos, s.read("density")[:n], sp.mass, sp.smoothing_distance, sp.EPSILON, *_box(gs), res)
    err = np.abs(f - ref).max() / np.abs(ref).max()
    print(f"colour field: max {ref.max():.4f}, nonzero nodes {(ref > 0).sum()}, max rel err {err:.2e}")
    assert (ref > 0).sum() > 100
    assert err <= 1e-4
    s.close()


@pytest.mark.parametrize("iso,frame", [(0.5, 0), (0.0001, 0), (0.5, 1)])
def test_surface_matches_restatement(cuda, iso, frame):
    sc, s = _state()
    res = 28
    s.setOption(13, frame)      # RTPS_OPT_SURFACE_FRAME
    v, n = s.surface(res, iso)
    f = s.colorField(res)
    _, gp, _ = s.params()
    wmin, wmax = _box(gp)
    rv, rn, keys = SO.marching_cubes(f, iso, _table(), wmin, wmax, frame)
    print(f"iso {iso} frame {frame}: {len(v)} triangles")
    assert len(v) == len(rv) > 50
    ext = 10.0 if frame else max(np.subtract(wmax, wmin))
    np.testing.assert_allclose(v, rv, rtol=0, atol=1e-5 * ext)
    np.testing.assert_allclose(n, rn, rtol=0, atol=1e-4)
    np.testing.assert_allclose(np.linalg.norm(n, axis=2), 1.0, atol=1e-4)
    if frame:
        assert v.min() >= -1e-4 and v.max() <= 10.0 + 1e-4
    s.close()


def test_surface_above_the_field_is_empty(cuda):
    """An iso level above the field's maximum: no cube is cut, the
    restatement and the GPU both give no triangles."""
    sc, s = _state(steps=5, seed=9)
    res = 20
    f = s.colorField(res)
    iso = float(f.max()) * 1.5 + 1.0
    v, n = s.surface(res, iso)
    _, gp, _ = s.params()
    rv, _, _ = SO.marching_cubes(f, iso, _table(), *_box(gp))
    assert v.shape == n.shape == (0, 3, 3)
    assert len(rv) == 0
    s.close()


def test_post_process_keeps_the_settings_surface(cuda):
    sc, s = _state(steps=10, seed=5)
    s.postProcess()
    assert s.getMCMesh() is None            # use_color_field defaults to 0
    s.setSetting("use_color_field", "1")
    s.setSetting("color_field_res", "20")
    s.postProcess()
    v, n = s.getMCMesh()
    v2, n2 = s.surface(20, 0.0001)          # the reference's isolevel
    np.testing.assert_array_equal(v, v2)
    np.testing.assert_array_equal(n, n2)
    s.close()
