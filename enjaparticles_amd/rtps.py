"""Python mirror of rtpslib's System/SPH API over the C ABI.

Names follow the reference: ``RTPSSettings`` (rtpslib/RTPSSettings.h:43-107),
``RTPS.generateSystemInstance`` (rtpslib/RTPS.h:45, RTPS.cpp:37-67), and the
``SPH`` system (rtpslib/system/SPH.h:56-155, System.h:61-271): ``update``,
``integrate``, ``pushParticles``, ``addBox``, ``getNum`` and the sorted /
unsorted buffer getters.  Every call goes to librtps_hip.so; buffers come back
as numpy copies (``device_ptr`` gives zero-copy device access).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N


class RTPSSettings:
    """String map with typed access (values are kept as strings)."""

    def __init__(self, init: dict | None = None):
        self._s: dict[str, str] = {}
        self._changed = False
        for k, v in (init or {}).items():
            self.SetSetting(k, v)

    def SetSetting(self, key: str, value) -> None:
        if isinstance(value, (list, tuple, np.ndarray)):
            value = " ".join(str(x) for x in np.asarray(value).ravel())
        self._s[key] = str(value)
        self._changed = True

    def GetSettingAs(self, kind, key: str, default: str = "0"):
        raw = self._s.get(key, default)
        if kind is str:
            return raw
        vals = raw.split()
        if kind in (list, np.ndarray):
            return [float(x) for x in vals]
        return kind(vals[0]) if vals else kind(0)

    def Exists(self, key: str) -> bool:
        return key in self._s

    def hasChanged(self) -> bool:
        return self._changed

    def updated(self) -> None:
        self._changed = False

    def items(self):
        return self._s.items()

    def as_dict(self) -> dict:
        return dict(self._s)


def rigid_spring_dampening(penetration_factor: float, velocity_limit: float, smoothing_distance: float,
                           restitution: float):
    """The rigid-body system's "spring" and "dampening" settings as
    ParticleRigidBody::updateParams derives them (ParticleRigidBody.cpp:603-608):
    spring = penetration_factor * velocity_limit / (4 h^2),
    dampening = -2 ln(restitution) / sqrt(ln(restitution)^2 + pi^2)."""
    import math
    spring = penetration_factor * velocity_limit / (4.0 * smoothing_distance * smoothing_distance)
    ln_res = math.log(restitution)
    return spring, -2.0 * ln_res / math.sqrt(ln_res * ln_res + math.pi * math.pi)


class SPH:
    """Handle of one SPH system on the GPU (rtps_sph in include/rtps_c.h)."""

    def __init__(self, settings: RTPSSettings | dict):
        if isinstance(settings, dict):
            settings = RTPSSettings(settings)
        self.settings = settings
        ka, va, n = N.settings_arrays(settings.as_dict())
        h = C.c_void_p()
        N.check(N.lib().rtps_sph_create(C.cast(ka, C.c_void_p), C.cast(va, C.c_void_p), n, C.byref(h)))
        self._h = h
        self.max_num = int(settings.GetSettingAs(int, "max_num_particles"))
        sp, gp, gs = self.params()
        self.nb_cells = gp.nb_cells

    # -- lifetime ---------------------------------------------------------
    def close(self) -> None:
        if getattr(self, "_h", None):
            N.lib().rtps_sph_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- System API -------------------------------------------------------
    def pushParticles(self, pos, vel=None, color=None) -> None:
        """SPH::pushParticles (SPH.cpp:560-602); color is one float4 for all
        pushed particles (default (1, 0, 0, 1), System.h:160)."""
        pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 4)
        vptr = None
        if vel is not None:
            vel = np.ascontiguousarray(vel, np.float32).reshape(-1, 4)
            assert vel.shape == pos.shape
            vptr = vel.ctypes.data
        col = _f4(DEFAULT_COLOR if color is None else color)
        N.check(N.lib().rtps_sph_push_particles_color(self._h, pos.ctypes.data, vptr, pos.shape[0], col))

    def addBox(self, nn: int, bmin, bmax, spacing: float | None = None, color=None) -> int:
        """System::addBox (System.cpp:331-338) with addRect (IV.cpp:165-202):
        float-accumulated lattice, z outer / x inner, at most nn particles."""
        if spacing is None:
            spacing = float(self.getSetting("spacing"))
        sp = np.float32(spacing)
        out = []
        z = np.float32(bmin[2])
        while z <= np.float32(bmax[2]) and len(out) < nn:
            y = np.float32(bmin[1])
            while y <= np.float32(bmax[1]) and len(out) < nn:
                x = np.float32(bmin[0])
                while x <= np.float32(bmax[0]) and len(out) < nn:
                    out.append((x, y, z, 1.0))
                    x = np.float32(x + sp)
                y = np.float32(y + sp)
            z = np.float32(z + sp)
        pos = np.array(out, np.float32).reshape(-1, 4)
        if len(pos):
            self.pushParticles(pos, color=color)
        return len(pos)

    def addHose(self, total_n: int, center, velocity, radius: float, color=None, mass=None) -> int:
        """SPH::addHose (SPH.cpp:523-532): a hose sprayed at the start of every
        update; radius is in units of the particle spacing.  Returns its index.
        Sprayed particles get the hose colour; mass is accepted for the
        reference's signature (every particle has the system's mass)."""
        idx = C.c_int()
        col = _f4(DEFAULT_COLOR if color is None else color)
        N.check(N.lib().rtps_sph_add_hose(self._h, int(total_n), _f4(center), _f4(velocity), float(radius), col,
                                          C.byref(idx)))
        return idx.value

    def updateHose(self, index: int, center, velocity, radius: float, color=None) -> None:
        """SPH::updateHose (SPH.cpp:533-541)."""
        col = _f4(DEFAULT_COLOR if color is None else color)
        N.check(N.lib().rtps_sph_update_hose(self._h, int(index), _f4(center), _f4(velocity), float(radius), col))

    def refillHose(self, index: int, refill: int) -> None:
        """SPH::refillHose (SPH.cpp:542-545)."""
        N.check(N.lib().rtps_sph_refill_hose(self._h, int(index), int(refill)))

    def loadTriangles(self, triangles) -> None:
        """SPH::loadTriangles: triangles (n, 3, 3) or (n, 3, 4) of vertex
        coordinates in world units; particles collide with them after the
        walls in every update (collision_tri.cl).  An empty array removes them."""
        t = triangles_array(triangles)
        N.check(N.lib().rtps_sph_load_triangles(self._h, t.ctypes.data if len(t) else None, len(t)))

    def testDelete(self) -> None:
        """System::testDelete (System.cpp:366-376): moves the first 40 particles
        to (100, 100, 100), outside the grid; the next update deletes them."""
        N.check(N.lib().rtps_sph_test_delete(self._h))

    def update(self) -> None:
        N.check(N.lib().rtps_sph_update(self._h))

    def interact(self) -> None:
        """SPH::interact (SPH.cpp:610-663): the loaded rigid-body particle sets
        add their forces to those of the preceding update()."""
        N.check(N.lib().rtps_sph_interact(self._h))

    def setRigidParticles(self, pos, vel, mass, spring: float, dampening: float) -> None:
        """A moving rigid body's particles (rigidbody_force.cl): positions,
        velocities (world units) and masses; spring / dampening are the rigid
        system's settings.  An empty array removes the set."""
        pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 4)
        vel = np.ascontiguousarray(vel, np.float32).reshape(-1, 4)
        mass = np.ascontiguousarray(mass, np.float32).reshape(-1)
        assert vel.shape == pos.shape and mass.shape[0] == pos.shape[0]
        n = pos.shape[0]
        N.check(N.lib().rtps_sph_set_rigid_particles(self._h, pos.ctypes.data if n else None,
                                                      vel.ctypes.data if n else None,
                                                      mass.ctypes.data if n else None, n, float(spring),
                                                      float(dampening)))

    def setStaticBoundary(self, pos, spring: float, dampening: float) -> None:
        """Static boundary particles of a rigid-body system
        (staticrigidbody_force.cl).  An empty array removes the set."""
        pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 4)
        n = pos.shape[0]
        N.check(N.lib().rtps_sph_set_static_boundary(self._h, pos.ctypes.data if n else None, n, float(spring),
                                                      float(dampening)))

    def integrate(self) -> None:
        N.check(N.lib().rtps_sph_integrate(self._h))

    def postProcess(self) -> None:
        """SPH::postProcess (SPH.cpp:338-386): with setting use_color_field = 1,
        the colour-field iso-surface (color_field_res, color_field_iso),
        returned by getMCMesh()."""
        N.check(N.lib().rtps_sph_post_process(self._h))

    def getMCMesh(self):
        """The mesh of the last postProcess(): (vertices, normals), each
        (ntris, 3, 3) float32, or None when there is none."""
        n = C.c_int()
        N.check(N.lib().rtps_sph_mesh(self._h, 0, None, None, C.byref(n)))
        if n.value == 0:
            return None
        v = np.empty((n.value, 3, 3), np.float32)
        nn = np.empty((n.value, 3, 3), np.float32)
        N.check(N.lib().rtps_sph_mesh(self._h, n.value, v.ctypes.data, nn.ctypes.data, C.byref(n)))
        return v, nn

    def colorField(self, res: int) -> np.ndarray:
        """ColorField::execute on the last step's sorted state: (res, res, res)
        float32 indexed [z, y, x] over the boundary box."""
        out = np.empty((res, res, res), np.float32)
        N.check(N.lib().rtps_sph_color_field(self._h, int(res), out.ctypes.data))
        return out

    def surface(self, res: int, iso: float = 0.0001):
        """Iso-surface of colorField(res) at iso: (vertices, normals), each
        (ntris, 3, 3) float32, vertices in world coordinates."""
        n = C.c_int()
        N.check(N.lib().rtps_sph_surface(self._h, int(res), float(iso), 0, None, None, C.byref(n)))
        v = np.empty((n.value, 3, 3), np.float32)
        nn = np.empty((n.value, 3, 3), np.float32)
        if n.value:
            N.check(N.lib().rtps_sph_mesh(self._h, n.value, v.ctypes.data, nn.ctypes.data, C.byref(n)))
        return v, nn

    def step(self, nsteps: int = 1) -> None:
        """nsteps x (update + integrate), fused kernels, one hipGraph replay per step."""
        N.check(N.lib().rtps_sph_step(self._h, int(nsteps)))

    def synchronize(self) -> None:
        N.check(N.lib().rtps_sph_synchronize(self._h))

    def getNum(self) -> int:
        v = C.c_int()
        N.check(N.lib().rtps_sph_get_num(self._h, C.byref(v)))
        return v.value

    def setOption(self, option: int, value: int) -> None:
        N.check(N.lib().rtps_sph_set_option(self._h, option, int(value)))

    def setSetting(self, key: str, value) -> None:
        self.settings.SetSetting(key, value)
        N.check(N.lib().rtps_sph_set_setting(self._h, key.encode(), str(value).encode()))

    def getSetting(self, key: str) -> str:
        buf = C.create_string_buffer(256)
        N.check(N.lib().rtps_sph_get_setting(self._h, key.encode(), buf, 256))
        return buf.value.decode()

    def params(self):
        sp, gp, gs = N.SPHParams(), N.GridParams(), N.GridParams()
        N.check(N.lib().rtps_sph_get_params(self._h, C.byref(sp), C.byref(gp), C.byref(gs)))
        return sp, gp, gs

    # -- buffers ----------------------------------------------------------
    def read(self, field: str, count: int | None = None) -> np.ndarray:
        f = N.FIELD[field]
        if count is None:
            count = self.nb_cells + 1 if field in ("cell_start", "cell_end") else self.getNum()
        if field in ("keys", "idx", "cell_start", "cell_end"):
            out = np.empty(count, np.uint32)
        elif field in ("density", "mass", "mass_s"):
            out = np.empty(count, np.float32)
        else:
            out = np.empty((count, 4), np.float32)
        if count:
            N.check(N.lib().rtps_sph_read(self._h, f, out.ctypes.data, count))
        return out

    def readAsync(self, field: str, out: "HostBuffer", count: int | None = None) -> None:
        """Enqueue a copy of a float4 field (or idx / cell arrays) into a
        page-locked HostBuffer; it is there after synchronize()."""
        count = out.count if count is None else count
        assert count <= out.count
        N.check(N.lib().rtps_sph_read_async(self._h, N.FIELD[field], out.ptr, int(count)))

    def write(self, field: str, arr) -> None:
        a = np.ascontiguousarray(arr, np.float32).reshape(-1, 4)
        N.check(N.lib().rtps_sph_write(self._h, N.FIELD[field], a.ctypes.data, a.shape[0]))

    def device_ptr(self, field: str) -> int:
        p = C.c_void_p()
        N.check(N.lib().rtps_sph_device_ptr(self._h, N.FIELD[field], C.byref(p)))
        return p.value

    def autotune(self, measure: bool = True) -> dict:
        """One step that (re)tunes the neighbour launches; returns the choice."""
        bd, bf, var = C.c_int(), C.c_int(), C.c_int()
        N.check(N.lib().rtps_sph_autotune(self._h, int(measure), C.byref(bd), C.byref(bf), C.byref(var)))
        return {"block_density": bd.value, "block_force": bf.value, "force_variant": var.value}

    def countPairs(self):
        """(K, M, K_tested): 27-cell candidates, accepted pairs (self included) and
        the candidates the windowed rows test, on the last step's state."""
        K, M, Kt = C.c_longlong(), C.c_longlong(), C.c_longlong()
        N.check(N.lib().rtps_sph_count_pairs(self._h, C.byref(K), C.byref(M), C.byref(Kt)))
        return K.value, M.value, Kt.value

    def stream(self) -> int:
        p = C.c_void_p()
        N.check(N.lib().rtps_sph_stream(self._h, C.byref(p)))
        return p.value or 0

    # reference getter names (System.h:168-177, SPH.h:71)
    def getPositionBufferUnsorted(self): return self.read("pos")
    def getVelocityBufferUnsorted(self): return self.read("vel")
    def getPositionBuffer(self): return self.read("pos_s")
    def getVelocityBuffer(self): return self.read("vel_s")
    def getDensityBuffer(self): return self.read("density")
    def getColorBuffer(self): return self.read("color_s")
    def getColorBufferUnsorted(self): return self.read("color")
    def getMassBuffer(self): return self.read("mass_s")
    def getMassBufferUnsorted(self): return self.read("mass")
    def getCellStartBuffer(self): return self.read("cell_start")
    def getCellEndBuffer(self): return self.read("cell_end")

    # -- timers (SPH::printTimers, SPH.cpp:664-672) -------------------------
    def stageTimes(self) -> dict:
        ms = (C.c_float * N.NUM_STAGES)()
        cnt = (C.c_int * N.NUM_STAGES)()
        N.check(N.lib().rtps_sph_stage_times(self._h, ms, cnt, N.NUM_STAGES))
        return {name: (ms[i], cnt[i]) for i, name in enumerate(N.STAGE_NAMES)}

    def resetTimers(self) -> None:
        N.check(N.lib().rtps_sph_reset_timers(self._h))

    def printTimers(self) -> str:
        lines = ["Timers [All times in ms (1/1000 s)]:"]
        for name, (avg, cnt) in self.stageTimes().items():
            lines.append(f"{name:<40s}|  avg: {avg:10.4f}  |  count= {cnt:6d}")
        text = "\n".join(lines)
        print(text)
        return text


DEFAULT_COLOR = (1.0, 0.0, 0.0, 1.0)   # System.h:160


def triangles_array(triangles) -> np.ndarray:
    """(n, 4, 4) float32 rtps_triangle records (three vertices, normal slot 0)
    from (n, 3, 3) or (n, 3, 4) vertex coordinates."""
    v = np.asarray(triangles, np.float32)
    if v.size == 0:
        return np.zeros((0, 4, 4), np.float32)
    assert v.ndim == 3 and v.shape[1] == 3 and v.shape[2] in (3, 4)
    out = np.zeros((len(v), 4, 4), np.float32)
    out[:, :3, :3] = v[:, :, :3]
    return out


class HostBuffer:
    """Page-locked host memory (rtps_host_alloc) seen as a numpy array of
    `count` float4 records (or uint32 with dtype=np.uint32)."""

    def __init__(self, count: int, dtype=np.float32):
        self.count = int(count)
        shape = (self.count, 4) if np.dtype(dtype) == np.float32 else (self.count,)
        nbytes = int(np.prod(shape)) * np.dtype(dtype).itemsize
        p = C.c_void_p()
        N.check(N.lib().rtps_host_alloc(nbytes, C.byref(p)))
        self.ptr = p.value
        raw = (C.c_char * max(nbytes, 1)).from_address(self.ptr)
        self.array = np.frombuffer(raw, dtype=dtype, count=int(np.prod(shape))).reshape(shape)

    def close(self) -> None:
        if getattr(self, "ptr", None):
            self.array = None
            N.lib().rtps_host_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _f4(v) -> N.F4:
    v = [float(c) for c in v] + [0.0] * (4 - len(v))
    return N.F4(*v[:4])


class Hose:
    """Hose (rtpslib/system/common/Hose.h, Hose.cpp:30-128): sprays a perturbed
    disc of particles (addDiscRandom, IV.cpp:313-365) perpendicular to its
    velocity every ``em`` calls of ``spray``, until ``total_n`` have left.
    Host code; random numbers come from a SplitMix64 stream started at
    ``seed`` (the reference uses libc rand())."""

    def __init__(self, dt, total_n, center, velocity, radius, spacing, seed: int = 0x5EED):
        h = C.c_void_p()
        N.check(N.lib().rtps_hose_create(float(dt), int(total_n), _f4(center), _f4(velocity), float(radius),
                                         float(spacing), int(seed), C.byref(h)))
        self._h = h

    def close(self) -> None:
        if getattr(self, "_h", None):
            N.lib().rtps_hose_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def update(self, center, velocity, radius, spacing) -> None:
        N.check(N.lib().rtps_hose_update(self._h, _f4(center), _f4(velocity), float(radius), float(spacing)))

    def refill(self, num: int) -> None:
        N.check(N.lib().rtps_hose_refill(self._h, int(num)))

    def spray(self) -> np.ndarray:
        """The particles sprayed by this call, (n, 4) float32 (w = 1)."""
        left = self.state()["n_left"]
        out = np.zeros((max(left, 1), 4), np.float32)
        n = C.c_int()
        N.check(N.lib().rtps_hose_spray(self._h, out.ctypes.data, left, C.byref(n)))
        return out[:n.value].copy()

    def state(self) -> dict:
        n, emc, em = C.c_int(), C.c_int(), C.c_int()
        u, w, v = N.F4(), N.F4(), N.F4()
        N.check(N.lib().rtps_hose_state(self._h, C.byref(n), C.byref(emc), C.byref(em), C.byref(u), C.byref(w),
                                         C.byref(v)))
        f = lambda q: np.array([q.x, q.y, q.z, q.w], np.float32)   # noqa: E731
        return {"n_left": n.value, "em_count": emc.value, "em": em.value, "u": f(u), "w": f(w),
                "velocity": f(v)}

    def getVelocity(self) -> np.ndarray:
        return self.state()["velocity"]


class RTPS:
    @staticmethod
    def generateSystemInstance(settings: RTPSSettings | dict):
        """Returns an SPH system, or None for an unknown "system" (RTPS.cpp:62-66)."""
        s = settings if isinstance(settings, RTPSSettings) else RTPSSettings(settings)
        if s.GetSettingAs(str, "system") != "sph":
            return None
        return SPH(s)


class BatchedSPH:
    """Many independent SPH scenes stepped together (rtps_batch in include/rtps_c.h).

    ``scenes`` is a list of settings dicts (one per scene); particles are pushed
    per scene; ``step(n)`` advances every scene n steps with one launch per
    stage.  Each scene computes what an ``SPH`` system with its settings does.
    """

    def __init__(self, scene_settings: list):
        flat_k, flat_v, counts = [], [], []
        for st in scene_settings:
            st = st.as_dict() if isinstance(st, RTPSSettings) else dict(st)
            counts.append(len(st))
            flat_k += [str(k).encode() for k in st]
            flat_v += [str(v).encode() for v in st.values()]
        ka = (C.c_char_p * max(1, len(flat_k)))(*flat_k)
        va = (C.c_char_p * max(1, len(flat_v)))(*flat_v)
        ca = (C.c_int * len(counts))(*counts)
        h = C.c_void_p()
        N.check(N.lib().rtps_batch_create(len(counts), ca, C.cast(ka, C.c_void_p), C.cast(va, C.c_void_p),
                                          C.byref(h)))
        self._h = h
        self.caps = [int(dict(st.as_dict() if isinstance(st, RTPSSettings) else st)["max_num_particles"])
                     for st in scene_settings]
        self.num_scenes = len(counts)

    def close(self):
        if getattr(self, "_h", None):
            N.lib().rtps_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def pushParticles(self, scene: int, pos, vel=None):
        pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 4)
        vptr = None
        if vel is not None:
            vel = np.ascontiguousarray(vel, np.float32).reshape(-1, 4)
            vptr = vel.ctypes.data
        N.check(N.lib().rtps_batch_push_particles(self._h, scene, pos.ctypes.data, vptr, pos.shape[0]))

    def step(self, nsteps: int = 1):
        N.check(N.lib().rtps_batch_step(self._h, int(nsteps)))

    def synchronize(self):
        N.check(N.lib().rtps_batch_synchronize(self._h))

    def getNum(self, scene: int) -> int:
        v = C.c_int()
        N.check(N.lib().rtps_batch_get_num(self._h, scene, C.byref(v)))
        return v.value

    def totalSlots(self) -> int:
        t = C.c_int()
        N.check(N.lib().rtps_batch_num_scenes(self._h, None, C.byref(t)))
        return t.value

    def setOption(self, option: int, value: int):
        N.check(N.lib().rtps_batch_set_option(self._h, option, int(value)))

    def read(self, scene: int, field: str, count: int | None = None) -> np.ndarray:
        if count is None:
            count = self.getNum(scene)
        if field in ("keys", "idx"):
            out = np.empty(count, np.uint32)
        elif field == "density":
            out = np.empty(count, np.float32)
        else:
            out = np.empty((count, 4), np.float32)
        if count:
            N.check(N.lib().rtps_batch_read(self._h, scene, N.FIELD[field], out.ctypes.data, count))
        return out

    def autotune(self, measure: bool = True) -> dict:
        """rtps_batch_autotune: one step that tunes the pair kernels' block size
        and register budget; returns the configurations in use."""
        v = [C.c_int() for _ in range(4)]
        N.check(N.lib().rtps_batch_autotune(self._h, int(measure), *(C.byref(x) for x in v)))
        return {"density": {"block": v[0].value, "waves": v[1].value},
                "force": {"block": v[2].value, "waves": v[3].value}}

    def stageTimes(self) -> dict:
        ms = (C.c_float * N.NUM_STAGES)()
        cnt = (C.c_int * N.NUM_STAGES)()
        N.check(N.lib().rtps_batch_stage_times(self._h, ms, cnt, N.NUM_STAGES))
        return {name: (ms[i], cnt[i]) for i, name in enumerate(N.STAGE_NAMES)}

    def resetTimers(self):
        N.check(N.lib().rtps_batch_reset_timers(self._h))
