"""ctypes binding of librtps_hip.so (the C ABI declared in include/rtps_c.h).

The library is built in-tree by ``enjaparticles_amd/csrc/Makefile`` into
``enjaparticles_amd/_build/librtps_hip.so``.  Loading fails loudly when it is
missing: there is no CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "librtps_hip.so")

RTPS_OK = 0
RTPS_ERR_ARG = -1
RTPS_ERR_HIP = -2
RTPS_ERR_CAPACITY = -3
RTPS_ERR_UNKNOWN_SYSTEM = -4
RTPS_ERR_STATE = -5

FIELD = {
    "pos": 0, "vel": 1, "veleval": 2, "pos_s": 3, "vel_s": 4, "veleval_s": 5, "density": 6,
    "force": 7, "xsph": 8, "keys": 9, "idx": 10, "cell_start": 11, "cell_end": 12,
    "color": 13, "color_s": 14, "mass": 15, "mass_s": 16,
}
(OPT_COLLISION, OPT_GRAPH, OPT_FORCE_OUT, OPT_TIMERS, OPT_FORCE_VARIANT, OPT_WINDOW, OPT_TILE_REMAP, OPT_AUTOTUNE,
 OPT_SORT_TILE, OPT_COLOR, OPT_DENSITY_STAGE, OPT_BLOCK_DENSITY, OPT_BLOCK_FORCE, OPT_SURFACE_FRAME,
 OPT_TILE_QUEUES, OPT_BUCKET_SORT, OPT_SHARED_ROWS) = range(17)
# The library's process-wide default of OPT_FORCE_VARIANT (the force loop of
# handles that are not autotuned).
DEFAULT_FORCE_LOOP = 5
NUM_STAGES = 6
STAGE_NAMES = ("hash", "sort", "cellindices", "permute", "density", "force")
MAX_GRAVITY_SOURCES = 16


class F4(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float), ("w", C.c_float)]


class SPHParams(C.Structure):
    """rtps_sph_params (SPHSettings.h:38-136 layout, 144 bytes)."""
    _fields_ = [(n, C.c_float) for n in (
        "rest_density", "mass", "rest_distance", "smoothing_distance", "simulation_scale",
        "boundary_stiffness", "boundary_dampening", "boundary_distance", "K",
        "viscosity", "velocity_limit", "xsph_factor",
        "friction_coef", "restitution_coef", "shear", "attraction", "spring",
        "EPSILON",
        "wpoly6_coef", "wpoly6_d_coef", "wpoly6_dd_coef", "wspiky_coef", "wspiky_d_coef",
        "wspiky_dd_coef", "wvisc_coef", "wvisc_d_coef", "wvisc_dd_coef")] + [
        ("num", C.c_int), ("max_num", C.c_int), ("pad_", C.c_int * 3), ("gravity", F4)]


class GridParams(C.Structure):
    """rtps_grid_params (Domain.h:144-172 layout, 128 bytes)."""
    _fields_ = [(n, F4) for n in (
        "grid_size", "grid_min", "grid_max", "bnd_min", "bnd_max", "grid_res", "grid_delta")] + [
        ("nb_cells", C.c_int), ("pad_", C.c_int * 3)]


assert C.sizeof(SPHParams) == 144 and SPHParams.gravity.offset == 128
assert C.sizeof(GridParams) == 128

P = C.c_void_p
_i, _f, _sz = C.c_int, C.c_float, C.c_size_t

SIGNATURES = {
    "rtps_last_error": (C.c_char_p, []),
    "rtps_version": (_i, []),
    "rtps_derive_params": (_i, [P, P, _i, P, P, P]),
    "rtps_hash": (_i, [_i, _i, P, P, P, P, P]),
    "rtps_sort_key_bits": (_i, [_i]),
    "rtps_sort_temp_bytes": (_sz, [_i, _i]),
    "rtps_sort_pairs": (_i, [_i, _i, P, P, P, _sz, P]),
    "rtps_sort_pairs_noalloc": (_i, [_i, _i, P, P, P]),
    "rtps_sort_segments": (_i, [_i, _i, _i, P, P, P, P, P, _sz, P]),
    "rtps_sph_set_rigid_particles": (_i, [P, P, P, P, _i, _f, _f]),
    "rtps_sph_set_static_boundary": (_i, [P, P, _i, _f, _f]),
    "rtps_sph_interact": (_i, [P]),
    "rtps_mc_geometry": (_i, [P, P]),
    "rtps_batch_autotune": (_i, [P, _i, P, P, P, P]),
    "rtps_cellindices": (_i, [_i, P, _i, P, P, P, P]),
    "rtps_permute": (_i, [_i, P, P, P, P, P, P, P, P]),
    "rtps_density": (_i, [_i, P, P, P, P, P, P, P]),
    "rtps_force": (_i, [_i, P, P, P, P, P, P, P, P, P, P]),
    "rtps_gravity": (_i, [_i, _i, P, P, P, P, P, _f, P]),
    "rtps_collision_wall": (_i, [_i, P, P, P, P, P, P]),
    "rtps_leapfrog": (_i, [_i, P, P, P, P, P, P, _f, P, P, P, P]),
    "rtps_sph_create": (_i, [P, P, _i, P]),
    "rtps_sph_destroy": (_i, [P]),
    "rtps_sph_push_particles": (_i, [P, P, P, _i]),
    "rtps_sph_push_particles_color": (_i, [P, P, P, _i, F4]),
    "rtps_sph_update": (_i, [P]),
    "rtps_sph_integrate": (_i, [P]),
    "rtps_sph_step": (_i, [P, _i]),
    "rtps_sph_synchronize": (_i, [P]),
    "rtps_sph_get_num": (_i, [P, P]),
    "rtps_sph_read": (_i, [P, _i, P, _i]),
    "rtps_sph_write": (_i, [P, _i, P, _i]),
    "rtps_sph_device_ptr": (_i, [P, _i, P]),
    "rtps_sph_read_async": (_i, [P, _i, P, _i]),
    "rtps_sph_load_triangles": (_i, [P, P, _i]),
    "rtps_collision_triangles": (_i, [_i, P, P, P, P, _i, P, P]),
    "rtps_sph_color_field": (_i, [P, _i, P]),
    "rtps_sph_surface": (_i, [P, _i, _f, _i, P, P, P]),
    "rtps_sph_post_process": (_i, [P]),
    "rtps_sph_mesh": (_i, [P, _i, P, P, P]),
    "rtps_mc_case_table": (_i, [P, P]),
    "rtps_host_alloc": (_i, [_sz, P]),
    "rtps_host_free": (_i, [P]),
    "rtps_sph_set_setting": (_i, [P, C.c_char_p, C.c_char_p]),
    "rtps_sph_get_setting": (_i, [P, C.c_char_p, P, _i]),
    "rtps_sph_get_params": (_i, [P, P, P, P]),
    "rtps_sph_set_option": (_i, [P, _i, _i]),
    "rtps_sph_autotune": (_i, [P, _i, P, P, P]),
    "rtps_sph_stage_times": (_i, [P, P, P, _i]),
    "rtps_sph_reset_timers": (_i, [P]),
    "rtps_sph_stream": (_i, [P, P]),
    "rtps_sph_count_pairs": (_i, [P, P, P, P]),
    "rtps_sph_add_hose": (_i, [P, _i, F4, F4, _f, F4, P]),
    "rtps_sph_update_hose": (_i, [P, _i, F4, F4, _f, F4]),
    "rtps_sph_refill_hose": (_i, [P, _i, _i]),
    "rtps_sph_test_delete": (_i, [P]),
    "rtps_hose_create": (_i, [_f, _i, F4, F4, _f, _f, C.c_ulonglong, P]),
    "rtps_hose_destroy": (_i, [P]),
    "rtps_hose_update": (_i, [P, F4, F4, _f, _f]),
    "rtps_hose_refill": (_i, [P, _i]),
    "rtps_hose_spray": (_i, [P, P, _i, P]),
    "rtps_hose_state": (_i, [P, P, P, P, P, P, P]),
    "rtps_batch_create": (_i, [_i, P, P, P, P]),
    "rtps_batch_destroy": (_i, [P]),
    "rtps_batch_num_scenes": (_i, [P, P, P]),
    "rtps_batch_push_particles": (_i, [P, _i, P, P, _i]),
    "rtps_batch_step": (_i, [P, _i]),
    "rtps_batch_synchronize": (_i, [P]),
    "rtps_batch_get_num": (_i, [P, _i, P]),
    "rtps_batch_read": (_i, [P, _i, _i, P, _i]),
    "rtps_batch_set_option": (_i, [P, _i, _i]),
    "rtps_batch_stage_times": (_i, [P, P, P, _i]),
    "rtps_batch_reset_timers": (_i, [P]),
    "rtps_adj_create": (_i, [P, P, _i, _i, _i, _i, P]),
    "rtps_adj_destroy": (_i, [P]),
    "rtps_adj_forward": (_i, [P, P, P, _i, _f, _f, _i, P, P, P]),
    "rtps_adj_backward": (_i, [P, P, P, P, P, P, P, P]),
    "rtps_badj_create": (_i, [_i, P, P, P, _i, _i, P]),
    "rtps_badj_destroy": (_i, [P]),
    "rtps_badj_layout": (_i, [P, P, P, P]),
    "rtps_badj_forward": (_i, [P, P, P, _i, P, P, _i, P, P, P]),
    "rtps_badj_backward": (_i, [P, P, P, P, P, P, P, P]),
}


class RTPSError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"rtps error {code}: {msg}")
        self.code = code


_lib = None


def lib():
    """The loaded library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
                "or `make -C enjaparticles_amd/csrc` (there is no CPU fallback)")
        # torch ships its own HIP runtime.  If this library were loaded first it
        # would bind the system runtime and a later `import torch` would bring a
        # second runtime into the process, which then cannot open the device.
        # Loading torch first makes both share torch's runtime.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        # RTLD_NOW: every symbol is bound at load, so a library with an
        # unresolved reference fails here instead of at its first use.
        L = C.CDLL(LIB_PATH, mode=os.RTLD_NOW)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int) -> int:
    if rc < 0:
        raise RTPSError(rc, lib().rtps_last_error().decode(errors="replace"))
    return rc


def settings_arrays(settings: dict):
    keys = [str(k).encode() for k in settings]
    vals = [str(v).encode() for v in settings.values()]
    ka = (C.c_char_p * max(1, len(keys)))(*keys)
    va = (C.c_char_p * max(1, len(vals)))(*vals)
    return ka, va, len(keys)


def derive_params(settings: dict):
    ka, va, n = settings_arrays(settings)
    sp, gp, gs = SPHParams(), GridParams(), GridParams()
    check(lib().rtps_derive_params(C.cast(ka, P), C.cast(va, P), n, C.byref(sp), C.byref(gp), C.byref(gs)))
    return sp, gp, gs
