"""Stage operators on device tensors (torch is used only for device memory and
streams).  Each wraps one C-ABI entry point of include/rtps_c.h and launches on
torch's current stream; arguments are validated on the host before launch.
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _native as N


def _stream() -> int:
    return torch.cuda.current_stream().cuda_stream


def _ptr(t: torch.Tensor | None) -> int | None:
    if t is None:
        return None
    assert t.is_cuda and t.is_contiguous(), "device tensors must be contiguous CUDA tensors"
    return t.data_ptr()


def _f4(t: torch.Tensor, n: int, name: str) -> None:
    if t.dtype != torch.float32 or t.dim() != 2 or t.shape[1] != 4 or t.shape[0] < n:
        raise ValueError(f"{name} must be float32 (>= {n}, 4), got {tuple(t.shape)} {t.dtype}")


def _u32(t: torch.Tensor, n: int, name: str) -> None:
    if t.dtype not in (torch.int32, torch.uint32) or t.dim() != 1 or t.shape[0] < n:
        raise ValueError(f"{name} must be a 32-bit integer vector of length >= {n}")


def hash(pos_u: torch.Tensor, gp: N.GridParams, num: int | None = None, n_pad: int | None = None):
    num = pos_u.shape[0] if num is None else num
    n_pad = num if n_pad is None else n_pad
    _f4(pos_u, num, "pos_u")
    keys = torch.empty(n_pad, dtype=torch.int32, device=pos_u.device)
    idx = torch.empty(n_pad, dtype=torch.int32, device=pos_u.device)
    N.check(N.lib().rtps_hash(num, n_pad, _ptr(pos_u), _ptr(keys), _ptr(idx), C.byref(gp), _stream()))
    return keys, idx


def sort_pairs(keys: torch.Tensor, vals: torch.Tensor, key_bits: int) -> None:
    """In-place stable ascending sort by the low key_bits bits."""
    n = keys.shape[0]
    _u32(keys, n, "keys")
    _u32(vals, n, "vals")
    nbytes = N.lib().rtps_sort_temp_bytes(n, key_bits)
    temp = torch.empty(max(1, nbytes), dtype=torch.uint8, device=keys.device)
    N.check(N.lib().rtps_sort_pairs(n, key_bits, _ptr(keys), _ptr(vals), _ptr(temp), nbytes, _stream()))


def sort_segments(dst_keys: torch.Tensor, dst_vals: torch.Tensor, src_keys: torch.Tensor, src_vals: torch.Tensor,
                  batch: int, array_length: int, direction: int) -> None:
    """Bitonic<T>::Sort(batch, arrayLength, dir, dst..., src...)
    (BitonicSort.h:26-28): batch arrays of array_length pairs, src -> dst,
    ascending for direction != 0, on all 32 key bits."""
    n = batch * array_length
    for t, name in ((dst_keys, "dst_keys"), (dst_vals, "dst_vals"), (src_keys, "src_keys"), (src_vals, "src_vals")):
        _u32(t, n, name)
    nbytes = N.lib().rtps_sort_temp_bytes(array_length, 32)
    temp = torch.empty(max(1, nbytes), dtype=torch.uint8, device=dst_keys.device)
    N.check(N.lib().rtps_sort_segments(batch, array_length, direction, _ptr(dst_keys), _ptr(dst_vals),
                                       _ptr(src_keys), _ptr(src_vals), _ptr(temp), nbytes, _stream()))


def key_bits(nb_cells: int) -> int:
    return N.check(N.lib().rtps_sort_key_bits(nb_cells))


def cellindices(keys: torch.Tensor, nb_cells: int, num: int | None = None, end: torch.Tensor | None = None):
    num = keys.shape[0] if num is None else num
    _u32(keys, num, "keys")
    start = torch.empty(nb_cells + 1, dtype=torch.int32, device=keys.device)
    if end is None:
        end = torch.full((nb_cells + 1,), 666, dtype=torch.int32, device=keys.device)
    nc = torch.empty(1, dtype=torch.int32, device=keys.device)
    N.check(N.lib().rtps_cellindices(num, _ptr(keys), nb_cells, _ptr(start), _ptr(end), _ptr(nc), _stream()))
    return start, end, nc


def permute(idx, pos_u, vel_u, veleval_u, num: int | None = None):
    num = idx.shape[0] if num is None else num
    for t, nm in ((pos_u, "pos_u"), (vel_u, "vel_u"), (veleval_u, "veleval_u")):
        _f4(t, 1, nm)
    _u32(idx, num, "idx")
    n_u = min(pos_u.shape[0], vel_u.shape[0], veleval_u.shape[0])
    if num and int(idx[:num].max().item()) >= n_u:
        raise ValueError("idx points outside the unsorted arrays")
    ps, vs, es = (torch.empty((num, 4), dtype=torch.float32, device=idx.device) for _ in range(3))
    N.check(N.lib().rtps_permute(num, _ptr(idx), _ptr(pos_u), _ptr(vel_u), _ptr(veleval_u), _ptr(ps), _ptr(vs),
                                 _ptr(es), _stream()))
    return ps, vs, es


def _check_cells(start, end, nb_cells, num):
    if start.shape[0] < nb_cells + 1 or end.shape[0] < nb_cells + 1:
        raise ValueError("start/end must hold nb_cells + 1 entries")


def density(pos_s, start, end, sphp: N.SPHParams, gps: N.GridParams):
    num = pos_s.shape[0]
    _f4(pos_s, num, "pos_s")
    _check_cells(start, end, gps.nb_cells, num)
    out = torch.empty(num, dtype=torch.float32, device=pos_s.device)
    N.check(N.lib().rtps_density(num, _ptr(pos_s), _ptr(start), _ptr(end), C.byref(sphp), C.byref(gps), _ptr(out),
                                 _stream()))
    return out


def force(pos_s, dens, veleval_s, start, end, sphp: N.SPHParams, gps: N.GridParams):
    num = pos_s.shape[0]
    _f4(pos_s, num, "pos_s")
    _f4(veleval_s, num, "veleval_s")
    if dens.shape[0] < num:
        raise ValueError("density too short")
    _check_cells(start, end, gps.nb_cells, num)
    f = torch.empty((num, 4), dtype=torch.float32, device=pos_s.device)
    x = torch.empty((num, 4), dtype=torch.float32, device=pos_s.device)
    N.check(N.lib().rtps_force(num, _ptr(pos_s), _ptr(dens), _ptr(veleval_s), _ptr(start), _ptr(end), C.byref(sphp),
                               C.byref(gps), _ptr(f), _ptr(x), _stream()))
    return f, x


def gravity(pos_s, force_s, sources, masses, alphas, scale: float):
    import numpy as np
    num = pos_s.shape[0]
    src = np.ascontiguousarray(sources, np.float32).reshape(-1, 4)
    m = np.ascontiguousarray(masses, np.float32)
    a = np.ascontiguousarray(alphas, np.float32)
    f = force_s.clone()
    N.check(N.lib().rtps_gravity(num, len(src), src.ctypes.data, m.ctypes.data, a.ctypes.data, _ptr(pos_s), _ptr(f),
                                 float(scale), _stream()))
    return f


def collision_wall(pos_s, vel_s, force_s, sphp, gps):
    num = pos_s.shape[0]
    f = force_s.clone()
    N.check(N.lib().rtps_collision_wall(num, _ptr(pos_s), _ptr(vel_s), _ptr(f), C.byref(sphp), C.byref(gps),
                                        _stream()))
    return f


def collision_triangles(pos_s, vel_s, force_s, triangles, sphp):
    """CollisionTriangle::execute: force_s plus the triangle collision forces
    (triangles: (n, 3, 3) world vertex coordinates, host)."""
    from .rtps import triangles_array
    num = pos_s.shape[0]
    f = force_s.clone()
    t = triangles_array(triangles)
    N.check(N.lib().rtps_collision_triangles(num, _ptr(pos_s), _ptr(vel_s), _ptr(f),
                                             t.ctypes.data if len(t) else None, len(t), C.byref(sphp), _stream()))
    return f


def leapfrog(pos_s, vel_s, force_s, xsph_s, idx, sphp, dt, pos_u, vel_u, veleval_u):
    """Returns updated clones of (pos_u, vel_u, veleval_u)."""
    num = pos_s.shape[0]
    if num and int(idx[:num].max().item()) >= pos_u.shape[0]:
        raise ValueError("idx points outside the unsorted arrays")
    pu, vu, eu = pos_u.clone(), vel_u.clone(), veleval_u.clone()
    N.check(N.lib().rtps_leapfrog(num, _ptr(pos_s), _ptr(vel_s), _ptr(force_s), _ptr(xsph_s), _ptr(idx),
                                  C.byref(sphp), float(dt), _ptr(pu), _ptr(vu), _ptr(eu), _stream()))
    return pu, vu, eu
