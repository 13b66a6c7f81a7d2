"""CPU oracle for the SPH step -- TEST INFRASTRUCTURE ONLY.

ctypes/numpy front end of ``oracle/sph_oracle.cpp`` (a CPU restatement of
EnjaParticles' OpenCL SPH step; the file:line citations live there).  Only
``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline``
leg may import this module, and only as the checker / the CPU baseline --
never as the product path.

The library is built by ``oracle/Makefile`` into ``oracle/_build``; if it is
missing this module builds it once.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libsph_oracle.so")


class F4(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float), ("w", C.c_float)]


class SPHParams(C.Structure):
    """Byte layout of rtpslib/system/SPHSettings.h:38-136 (144 bytes)."""
    _fields_ = [(n, C.c_float) for n in (
        "rest_density", "mass", "rest_distance", "smoothing_distance", "simulation_scale",
        "boundary_stiffness", "boundary_dampening", "boundary_distance", "K",
        "viscosity", "velocity_limit", "xsph_factor",
        "friction_coef", "restitution_coef", "shear", "attraction", "spring",
        "EPSILON",
        "wpoly6_coef", "wpoly6_d_coef", "wpoly6_dd_coef", "wspiky_coef", "wspiky_d_coef",
        "wspiky_dd_coef", "wvisc_coef", "wvisc_d_coef", "wvisc_dd_coef")] + [
        ("num", C.c_int), ("max_num", C.c_int), ("_pad", C.c_int * 3), ("gravity", F4)]


class GridParams(C.Structure):
    """Byte layout of rtpslib/domain/Domain.h:144-172 (128 bytes)."""
    _fields_ = [(n, F4) for n in (
        "grid_size", "grid_min", "grid_max", "bnd_min", "bnd_max", "grid_res", "grid_delta")] + [
        ("nb_cells", C.c_int), ("_pad", C.c_int * 3)]


assert C.sizeof(SPHParams) == 144 and SPHParams.gravity.offset == 128
assert C.sizeof(GridParams) == 128


def _build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def _load():
    if not os.path.exists(LIB_PATH):
        _build()
    return C.CDLL(LIB_PATH)


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load()
        _declare(_lib)
    return _lib


P = C.c_void_p


def _declare(L):
    i, u, f, d = C.c_int, C.c_uint, C.c_float, C.c_double
    sig = {
        "oracle_hash_f32": (None, [i, P, P, P, P]),
        "oracle_hash_f64": (None, [i, P, P, P, P]),
        "oracle_sort_stable": (None, [i, P, P]),
        "oracle_sort_bitonic": (None, [i, P, P]),
        "oracle_cellindices": (i, [i, P, i, P, P]),
        "oracle_permute_f32": (None, [i, P, P, P, P, P, P, P]),
        "oracle_permute_f64": (None, [i, P, P, P, P, P, P, P]),
        "oracle_density_f32": (None, [i, P, P, P, P, P, P]),
        "oracle_density_f64": (None, [i, P, P, P, P, P, P]),
        "oracle_force_f32": (None, [i, P, P, P, P, P, P, P, P, P]),
        "oracle_force_f64": (None, [i, P, P, P, P, P, P, P, P, P]),
        "oracle_gravity_f32": (None, [i, i, P, P, P, P, P, f]),
        "oracle_gravity_f64": (None, [i, i, P, P, P, P, P, d]),
        "oracle_collision_wall_f32": (None, [i, P, P, P, P, P]),
        "oracle_collision_wall_f64": (None, [i, P, P, P, P, P]),
        "oracle_leapfrog_f32": (None, [i, P, P, P, P, P, P, f, P, P, P]),
        "oracle_leapfrog_f64": (None, [i, P, P, P, P, P, P, d, P, P, P]),
        "oracle_euler_f32": (None, [i, P, P, P, P, P, f, P, P]),
        "oracle_euler_f64": (None, [i, P, P, P, P, P, d, P, P]),
        "oracle_count_pairs": (None, [i, P, P, P, P, P, P, P]),
        "oracle_step_f32": (i, [P, P, P, P, P, P, P, f, i, i, P, P, P] + [P] * 10),
        "oracle_step_f64": (i, [P, P, P, P, P, P, P, d, i, i, P, P, P] + [P] * 10),
        "oracle_step_tl": (i, [P, P, P, P, P, P, P, d, i, i, P, P, P, P, P] + [P] * 10),
        "oracle_tl_directions": (i, []),
        "oracle_derive_params": (i, [P, P, i, P, P, P]),
        "oracle_add_rect": (i, [i, P, P, f, f, P]),
        "oracle_rigid_force_f32": (None, [i] + [P] * 11 + [f, f, i]),
        "oracle_rigid_force_f64": (None, [i] + [P] * 11 + [f, f, i]),
        "oracle_set_rigid": (None, [i, i, P, P, P, f, f]),
        "oracle_num_threads": (i, []),
        "oracle_set_threads": (None, [i]),
        "oracle_last_stage_times": (None, [P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args


def _p(a):
    if a is None:
        return None
    if isinstance(a, C.Structure):
        return C.addressof(a)
    assert a.flags["C_CONTIGUOUS"], "oracle arrays must be C-contiguous"
    return a.ctypes.data


def _sfx(dtype):
    return "f32" if np.dtype(dtype) == np.float32 else "f64"


def num_threads() -> int:
    return lib().oracle_num_threads()


STAGE_NAMES = ("hash", "sort", "cellindices", "permute", "density", "force", "gravity+walls", "leapfrog")


def last_stage_times() -> dict:
    """Wall-clock seconds of each stage of the last State.step (steady_clock)."""
    out = np.zeros(8, np.float64)
    lib().oracle_last_stage_times(out.ctypes.data)
    return dict(zip(STAGE_NAMES, out.tolist()))


def set_threads(n: int) -> None:
    """OpenMP threads of the following oracle calls (n <= 0: all host cores)."""
    lib().oracle_set_threads(int(n))


# ---------------------------------------------------------------------------
def derive_params(settings: dict):
    """Replay SPH::calculate / updateParams / setupDomain on a settings map."""
    keys = [k.encode() for k in settings]
    vals = [str(v).encode() for v in settings.values()]
    ka = (C.c_char_p * len(keys))(*keys)
    va = (C.c_char_p * len(vals))(*vals)
    sp, gp, gs = SPHParams(), GridParams(), GridParams()
    rc = lib().oracle_derive_params(C.cast(ka, P), C.cast(va, P), len(keys), _p(sp), _p(gp), _p(gs))
    assert rc == 0
    return sp, gp, gs


def add_rect(num, bmin, bmax, spacing, scale=1.0):
    """addRect (IV.cpp:165-202): (count, 4) float32 lattice, z outer, x inner."""
    out = np.zeros((max(num, 0), 4), np.float32)
    lo = np.ascontiguousarray(bmin[:3], np.float32)
    hi = np.ascontiguousarray(bmax[:3], np.float32)
    k = lib().oracle_add_rect(int(num), _p(lo), _p(hi), float(spacing), float(scale), _p(out))
    return out[:k].copy()


def hash(pos_u, gp):
    n = pos_u.shape[0]
    keys = np.empty(n, np.uint32)
    idx = np.empty(n, np.uint32)
    getattr(lib(), "oracle_hash_" + _sfx(pos_u.dtype))(n, _p(pos_u), _p(gp), _p(keys), _p(idx))
    return keys, idx


def sort_stable(keys, idx):
    k, v = keys.copy(), idx.copy()
    lib().oracle_sort_stable(len(k), _p(k), _p(v))
    return k, v


def sort_bitonic(keys, idx):
    k, v = keys.copy(), idx.copy()
    lib().oracle_sort_bitonic(len(k), _p(k), _p(v))
    return k, v


def cellindices(keys, nb_cells, end=None):
    start = np.empty(nb_cells + 1, np.uint32)
    if end is None:
        end = np.full(nb_cells + 1, 666, np.uint32)  # SPH.cpp:512-517 initial fill
    nc = lib().oracle_cellindices(len(keys), _p(keys), nb_cells, _p(start), _p(end))
    return start, end, nc


def permute(idx, pos_u, vel_u, veleval_u):
    n = len(idx)
    ps, vs, es = (np.empty((n, 4), pos_u.dtype) for _ in range(3))
    getattr(lib(), "oracle_permute_" + _sfx(pos_u.dtype))(
        n, _p(idx), _p(pos_u), _p(vel_u), _p(veleval_u), _p(ps), _p(vs), _p(es))
    return ps, vs, es


def density(pos_s, start, end, sphp, gps):
    n = pos_s.shape[0]
    out = np.empty(n, pos_s.dtype)
    getattr(lib(), "oracle_density_" + _sfx(pos_s.dtype))(
        n, _p(pos_s), _p(start), _p(end), _p(sphp), _p(gps), _p(out))
    return out


def force(pos_s, dens, veleval_s, start, end, sphp, gps):
    n = pos_s.shape[0]
    f = np.empty((n, 4), pos_s.dtype)
    x = np.empty((n, 4), pos_s.dtype)
    getattr(lib(), "oracle_force_" + _sfx(pos_s.dtype))(
        n, _p(pos_s), _p(dens), _p(veleval_s), _p(start), _p(end), _p(sphp), _p(gps), _p(f), _p(x))
    return f, x


def gravity(pos_s, force_s, sources, masses, alphas, scale):
    f = force_s.copy()
    src = np.ascontiguousarray(sources, np.float32).reshape(-1, 4)
    m = np.ascontiguousarray(masses, np.float32)
    a = np.ascontiguousarray(alphas, np.float32)
    getattr(lib(), "oracle_gravity_" + _sfx(pos_s.dtype))(
        pos_s.shape[0], len(src), _p(src), _p(m), _p(a), _p(pos_s), _p(f), scale)
    return f


def collision_wall(pos_s, vel_s, force_s, sphp, gps):
    f = force_s.copy()
    getattr(lib(), "oracle_collision_wall_" + _sfx(pos_s.dtype))(
        pos_s.shape[0], _p(pos_s), _p(vel_s), _p(f), _p(sphp), _p(gps))
    return f


def leapfrog(pos_s, vel_s, force_s, xsph_s, idx, sphp, dt, pos_u, vel_u, veleval_u):
    """Returns updated copies of (pos_u, vel_u, veleval_u)."""
    pu, vu, eu = pos_u.copy(), vel_u.copy(), veleval_u.copy()
    getattr(lib(), "oracle_leapfrog_" + _sfx(pos_s.dtype))(
        pos_s.shape[0], _p(pos_s), _p(vel_s), _p(force_s), _p(xsph_s), _p(idx), _p(sphp), dt,
        _p(pu), _p(vu), _p(eu))
    return pu, vu, eu


def euler(pos_s, vel_s, force_s, idx, sphp, dt, pos_u, vel_u):
    """euler.cl:33-78; returns updated copies of (pos_u, vel_u)."""
    pu, vu = pos_u.copy(), vel_u.copy()
    getattr(lib(), "oracle_euler_" + _sfx(pos_s.dtype))(
        pos_s.shape[0], _p(pos_s), _p(vel_s), _p(force_s), _p(idx), _p(sphp), dt, _p(pu), _p(vu))
    return pu, vu


def rigid_force(pos_s, veleval_s, dens, force_s, bpos_s, bstart, bend, sphp, gps, spring, dampening,
                bvel_s=None, bmass_s=None):
    """Rigid-body forces of SPH::interact added to a copy of force_s:
    staticrigidbody_force.cl (bvel_s None) or rigidbody_force.cl (dynamic)."""
    f = np.ascontiguousarray(force_s).copy()
    bp = np.ascontiguousarray(bpos_s, np.float32)
    dyn = bvel_s is not None
    bv = np.ascontiguousarray(bvel_s if dyn else np.zeros_like(bp), np.float32)
    bm = np.ascontiguousarray(bmass_s if dyn else np.zeros(len(bp)), np.float32)
    getattr(lib(), "oracle_rigid_force_" + _sfx(pos_s.dtype))(
        pos_s.shape[0], _p(pos_s), _p(veleval_s), _p(dens), _p(f), _p(bp), _p(bv), _p(bm), _p(bstart), _p(bend),
        _p(sphp), _p(gps), spring, dampening, int(dyn))
    return f


def set_rigid(kind: str, pos=None, spring=0.0, dampening=0.0, vel=None, mass=None):
    """Rigid-body particles coupled into the following State.step calls
    (kind "dynamic" or "static"; pos None removes the set)."""
    k = 1 if kind == "static" else 0
    if pos is None:
        lib().oracle_set_rigid(k, 0, None, None, None, 0.0, 0.0)
        return
    p_ = np.ascontiguousarray(pos, np.float32).reshape(-1, 4)
    v_ = None if vel is None else np.ascontiguousarray(vel, np.float32).reshape(-1, 4)
    m_ = None if mass is None else np.ascontiguousarray(mass, np.float32)
    lib().oracle_set_rigid(k, len(p_), _p(p_), _p(v_), _p(m_), float(spring), float(dampening))


def count_pairs(pos_s, start, end, sphp, gps):
    K, M = C.c_longlong(), C.c_longlong()
    lib().oracle_count_pairs(pos_s.shape[0], _p(pos_s), _p(start), _p(end), _p(sphp), _p(gps),
                             C.byref(K), C.byref(M))
    return K.value, M.value


FLAG_COLLISION = 1
FLAG_BITONIC = 2
FLAG_EULER = 4      # integrator "euler" (euler.cl) instead of leapfrog


class State:
    """Unsorted particle state plus the last step's sorted intermediates."""

    def __init__(self, pos, vel, sphp, gp, gps, dtype=np.float32, veleval=None,
                 sources=None, masses=None, alphas=None):
        n = pos.shape[0]
        self.dtype = np.dtype(dtype)
        self.num = n
        self.pos_u = np.ascontiguousarray(pos, dtype).copy()
        self.vel_u = np.ascontiguousarray(vel, dtype).copy()
        self.veleval_u = (np.zeros((n, 4), dtype) if veleval is None
                          else np.ascontiguousarray(veleval, dtype).copy())
        self.sphp, self.gp, self.gps = sphp, gp, gps
        C_ = gp.nb_cells
        self.keys = np.zeros(n, np.uint32)
        self.idx = np.zeros(n, np.uint32)
        self.start = np.full(C_ + 1, 666, np.uint32)
        self.end = np.full(C_ + 1, 666, np.uint32)
        self.pos_s, self.vel_s, self.veleval_s, self.force, self.xsph = (
            np.zeros((n, 4), dtype) for _ in range(5))
        self.density = np.zeros(n, dtype)
        if sources is None or alphas is None or not np.any(np.asarray(alphas)):
            self.nsrc = 0
            self.src = np.zeros((1, 4), np.float32)
            self.smass = np.zeros(1, np.float32)
            self.alpha = np.zeros(1, np.float32)
        else:
            self.src = np.ascontiguousarray(sources, np.float32).reshape(-1, 4)
            self.smass = np.ascontiguousarray(masses, np.float32)
            self.alpha = np.ascontiguousarray(alphas, np.float32)
            self.nsrc = len(self.src)
        self.deleted = 0

    def step(self, dt, flags=FLAG_COLLISION):
        num = C.c_int(self.num)
        fn = getattr(lib(), "oracle_step_" + _sfx(self.dtype))
        d = fn(C.byref(num), _p(self.pos_u), _p(self.vel_u), _p(self.veleval_u), _p(self.sphp),
               _p(self.gp), _p(self.gps), dt, flags, self.nsrc, _p(self.src), _p(self.smass),
               _p(self.alpha), _p(self.keys), _p(self.idx), _p(self.start), _p(self.end),
               _p(self.pos_s), _p(self.vel_s), _p(self.veleval_s), _p(self.density),
               _p(self.force), _p(self.xsph))
        self.num = num.value
        self.deleted += d
        return d


class TLState:
    """FP64 state carrying kTL tangent directions (forward-mode derivatives of
    the oracle step, branches fixed as the values decide them).

    pos/vel/veleval hold (n, 4, 1 + kTL) doubles: [..., 0] the value (the FP64
    oracle's), [..., 1 + k] the derivative along direction k.  Seed the
    directions with ``dvel0`` ((kTL, n, 4) initial-velocity tangents) and
    ``dmu`` / ``dK`` (kTL viscosity / gas-constant tangents, applied at every
    step).  After the steps, ``pos_u[:, :, 1 + k]`` is J_k, the derivative of
    the final positions along direction k.
    """

    def __init__(self, pos, vel, sphp, gp, gps, dvel0=None, dmu=None, dK=None):
        L = lib()
        self.k = L.oracle_tl_directions()
        n = pos.shape[0]
        self.num = n
        w = 1 + self.k
        self.pos_u = np.zeros((n, 4, w))
        self.vel_u = np.zeros((n, 4, w))
        self.veleval_u = np.zeros((n, 4, w))
        self.pos_u[:, :, 0] = pos
        self.vel_u[:, :, 0] = vel
        if dvel0 is not None:
            for k in range(self.k):
                self.vel_u[:, :, 1 + k] = dvel0[k]
        self.dmu = np.zeros(self.k) if dmu is None else np.asarray(dmu, np.float64)
        self.dK = np.zeros(self.k) if dK is None else np.asarray(dK, np.float64)
        self.sphp, self.gp, self.gps = sphp, gp, gps
        self.keys = np.zeros(n, np.uint32)
        self.idx = np.zeros(n, np.uint32)
        self.start = np.full(gp.nb_cells + 1, 666, np.uint32)
        self.end = np.full(gp.nb_cells + 1, 666, np.uint32)
        self.pos_s, self.vel_s, self.veleval_s, self.force, self.xsph = (np.zeros((n, 4, w)) for _ in range(5))
        self.density = np.zeros((n, w))
        self.src = np.zeros((1, 4), np.float32)
        self.smass = np.zeros(1, np.float32)
        self.alpha = np.zeros(1, np.float32)

    def step(self, dt, flags=FLAG_COLLISION):
        num = C.c_int(self.num)
        d = lib().oracle_step_tl(C.byref(num), _p(self.pos_u), _p(self.vel_u), _p(self.veleval_u), _p(self.sphp),
                                 _p(self.gp), _p(self.gps), dt, flags, 0, _p(self.src), _p(self.smass),
                                 _p(self.alpha), _p(self.dmu), _p(self.dK), _p(self.keys), _p(self.idx),
                                 _p(self.start), _p(self.end), _p(self.pos_s), _p(self.vel_s), _p(self.veleval_s),
                                 _p(self.density), _p(self.force), _p(self.xsph))
        self.num = num.value
        return d
