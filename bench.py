#!/usr/bin/env python3
"""Benchmark of the SPH step on one MI355X.

Workload: the plain dam break with 1,048,576 FP32 particles (128x128x64
jittered lattice, h = 0.04, rest density 1000, gas constant 3, viscosity
0.01, dt = 0.003, walls on), synthetic seeded inputs resident in HBM.  A step
is one full pass of the pipeline (hash, sort, cell start/end, permute,
density, force + walls + leapfrog), replayed from one hipGraph.

Prints one JSON line: particle-updates/s (N * steps / s), the per-stage
times, the roofline of the dominant kernel (force, VALU-bound) and a CPU
baseline (the FP32 CPU restatement of the reference step, OpenMP, timed on a
bounded sample of the same workload).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-updates/sec at 1M SPH (plain vs batched vs fwd+adjoint); HBM % peak"
PEAK_HBM_GBS = 8000.0          # MI355X HBM3E
PEAK_FP32_TFLOPS = 157.3       # MI355X published FP32 vector peak (packed: two lanes per VALU op)
PEAK_VALU_FP32_TFLOPS = 78.6   # non-packed FP32 VALU rate (the kernels are compiled without packed FP32)
# The force kernel's HBM bytes per launch on the timed state (steps 6-25 of
# the dam break) come from the committed counter summary of that state: a line
#   traffic <kernel> <grid> <block> <bytes per launch>
# written from two rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE) scaled by a
# device copy of known size in the same runs.
COUNTER_SUMMARY = os.path.join(ROOT, "profiles", "dam1m_timed_counters.txt")


def measured_traffic(kernel: str = "k_force", grid: int = 1048576, block: int = 64):
    """HBM bytes per launch of `kernel` from the counter summary, or None."""
    try:
        with open(COUNTER_SUMMARY) as f:
            for line in f:
                p = line.split()
                if len(p) == 5 and p[0] == "traffic" and p[1] == kernel and int(p[2]) == grid and int(p[3]) == block:
                    return float(p[4])
    except (OSError, ValueError):
        pass
    return None


class _Marker:
    """roctx range around the timed region (recorded by rocprofv3
    --marker-trace, so a kernel trace can be cut to the timed steps); a no-op
    where no roctx library can be loaded.  rocprofv3 records the ranges of
    the rocprofiler-sdk roctx library; the older libroctx64 is a fallback for
    tracers that hook that one."""

    def __init__(self):
        import ctypes
        self.lib = None
        for name in ("librocprofiler-sdk-roctx.so.1", "/opt/rocm/lib/librocprofiler-sdk-roctx.so.1",
                     "libroctx64.so.4", "/opt/rocm/lib/libroctx64.so.4"):
            try:
                self.lib = ctypes.CDLL(name)
                self.lib.roctxRangePushA.argtypes = [ctypes.c_char_p]
                break
            except OSError:
                continue

    def push(self, name: str):
        if self.lib:
            self.lib.roctxRangePushA(name.encode())

    def pop(self):
        if self.lib:
            self.lib.roctxRangePop()


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(scene, steps: int = 2, single_steps: int = 2):
    """Oracle FP32 step (bitonic sort, OpenMP) on a bounded sample: `steps`
    steps of the same 1M-particle workload from its initial state on all
    host threads, then `single_steps` more on one thread."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    sp, gp, gs = O.derive_params(scene.settings)
    st = O.State(scene.pos, scene.vel, sp, gp, gs)
    flags = O.FLAG_COLLISION | O.FLAG_BITONIC
    st.step(scene.dt, flags)   # untimed first touch
    threads = O.num_threads()
    per_step = []   # per-stage wall-clock seconds of every timed step (steady_clock inside the oracle)
    t0 = time.perf_counter()
    for _ in range(steps):
        st.step(scene.dt, flags)
        per_step.append(O.last_stage_times())
    dt = time.perf_counter() - t0
    names = list(per_step[0]) if per_step else []
    stage_med = {k: round(1e3 * float(np.median([p[k] for p in per_step])), 3) for k in names}
    single = None
    if single_steps > 0:
        O.set_threads(1)
        try:
            single_stages = []
            t1 = time.perf_counter()
            for _ in range(single_steps):
                st.step(scene.dt, flags)
                single_stages.append(O.last_stage_times())
            d1 = time.perf_counter() - t1
        finally:
            O.set_threads(threads)
        single = {"value": scene.n * single_steps / d1, "ms_per_step": 1e3 * d1 / single_steps,
                  "steps": single_steps,
                  "stages_ms_median": {k: round(1e3 * float(np.median([p[k] for p in single_stages])), 3)
                                       for k in names}}
    return {
        "value": scene.n * steps / dt,
        "unit": "particle-updates/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{steps} steps (after 1 untimed) of the {scene.n}-particle dam break, FP32 CPU restatement "
                  f"of EnjaParticles' OpenCL kernels (hash, bitonic sort, cell indices, permute, density, "
                  f"force, walls, leapfrog), {threads} OpenMP threads",
        "ms_per_step": 1e3 * dt / steps,
        "stages_ms_median": stage_med,
        "cpu_model": _cpu_model(),
        "host_cpus": os.cpu_count(),
        "single_thread": single,
    }


def batched_leg(steps: int, num_scenes: int = 256, seed: int = 1003):
    """Batched scenes (256 scenes, counts log-uniform in [4096, 65536]) stepped
    with one launch per stage, against the same scenes stepped one after
    another as plain systems (one graph replay per scene-step)."""
    import torch
    from enjaparticles_amd import scenes
    from enjaparticles_amd.rtps import SPH, BatchedSPH
    scs = scenes.batched(num_scenes, 4096, 65536, seed=seed)
    total = sum(sc.n for sc in scs)
    b = BatchedSPH([sc.settings for sc in scs])
    for i, sc in enumerate(scs):
        b.pushParticles(i, sc.pos, sc.vel)
    b.step(2)
    b.synchronize()
    t0 = time.perf_counter()
    b.step(steps)
    b.synchronize()
    tb = time.perf_counter() - t0
    b.setOption(3, 1)
    b.resetTimers()
    b.step(3)
    stages = {k: round(v[0], 4) for k, v in b.stageTimes().items()}
    b.close()
    plains = []
    for sc in scs:
        p = SPH(sc.settings)
        p.pushParticles(sc.pos, sc.vel)
        p.step(2)
        plains.append(p)
    torch.cuda.synchronize()
    for p in plains:
        p.synchronize()
    t0 = time.perf_counter()
    for p in plains:
        p.step(steps)
        p.synchronize()
    ts = time.perf_counter() - t0
    for p in plains:
        p.close()
    vb, vs = total * steps / tb, total * steps / ts
    return {"scenes": num_scenes, "particles": total, "steps": steps, "value": vb,
            "unit": "particle-updates/s", "sequential_plain_value": vs, "ratio": vb / vs,
            "ms_per_step": 1e3 * tb / steps, "stages_ms": stages}


def adjoint_leg(steps: int = 200, checkpoint_every: int = 20):
    """Forward + adjoint of the 65,536-particle rollout (200 steps, dt 0.002,
    checkpoints every 20 steps): the forward keeps only checkpoints, the
    backward replays each segment and runs the adjoint steps.  Also times the
    same scene as a plain system (graph replay) for the cost ratio."""
    import numpy as np
    import torch
    from enjaparticles_amd import scenes
    from enjaparticles_amd.adjoint import AdjointSPH
    from enjaparticles_amd.rtps import SPH
    sc = scenes.adjoint64k()
    dev = torch.device("cuda:0")
    sim = AdjointSPH(sc.settings, sc.n, checkpoint_every=checkpoint_every, max_steps=steps)
    p0 = torch.from_numpy(sc.pos).to(dev)
    v0 = torch.from_numpy(sc.vel).to(dev)
    g = torch.from_numpy(np.random.default_rng(5).normal(size=sc.pos.shape).astype(np.float32)).to(dev)
    g[:, 3] = 0.0
    mu, gas = 0.01, 3.0
    sim.forward(p0, v0, steps, mu, gas)          # untimed first pass
    sim.backward(g)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sim.forward(p0, v0, steps, mu, gas)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    sim.backward(g)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    sim.close()
    p = SPH(sc.settings)
    p.pushParticles(sc.pos, sc.vel)
    p.step(2)
    p.synchronize()
    t3 = time.perf_counter()
    p.step(steps)
    p.synchronize()
    t4 = time.perf_counter()
    p.close()
    tf, tb, tp = t1 - t0, t2 - t1, t4 - t3
    return {"particles": sc.n, "steps": steps, "checkpoint_every": checkpoint_every,
            "value": sc.n * steps / (tf + tb), "unit": "particle-updates/s (forward + adjoint)",
            "forward_ms_per_step": 1e3 * tf / steps, "adjoint_ms_per_step": 1e3 * tb / steps,
            "plain_ms_per_step": 1e3 * tp / steps, "adjoint_over_forward": tb / tf,
            "adjoint_over_plain": tb / tp}


def batched_adjoint_leg(steps: int = 100, num_scenes: int = 64, checkpoint_every: int = 20, seed: int = 1005):
    """Batched + differentiable: 64 scenes with counts log-uniform in
    [16384, 131072] (per-scene viscosity and gas constant), forward and
    per-scene adjoint of `steps` steps in one batched launch per stage, next
    to the same scenes' plain batched forward."""
    import numpy as np
    import torch
    from enjaparticles_amd import scenes
    from enjaparticles_amd.adjoint import BatchedAdjointSPH
    from enjaparticles_amd.rtps import BatchedSPH
    scs = scenes.batched(num_scenes, 16384, 131072, seed=seed)
    total = sum(sc.n for sc in scs)
    dev = torch.device("cuda:0")
    pos0 = torch.from_numpy(np.concatenate([sc.pos for sc in scs])).to(dev)
    vel = np.zeros((total, 4), np.float32)
    vel[:, :3] = scenes.normal(seed + 1, 3 * total, 0.05).reshape(total, 3)
    vel0 = torch.from_numpy(vel).to(dev)
    mus = [sc.meta["viscosity"] for sc in scs]
    gases = [sc.meta["gas_constant"] for sc in scs]
    b = BatchedAdjointSPH([sc.settings for sc in scs], checkpoint_every=checkpoint_every, max_steps=steps)
    g = torch.from_numpy(np.random.default_rng(6).normal(size=(total, 4)).astype(np.float32)).to(dev)
    g[:, 3] = 0.0
    b.forward(pos0, vel0, checkpoint_every, mus, gases)      # untimed first touch (one segment)
    b.backward(g)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b.forward(pos0, vel0, steps, mus, gases)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    _, _, dmu, dK = b.backward(g)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    b.close()
    p = BatchedSPH([sc.settings for sc in scs])
    for i, sc in enumerate(scs):
        p.pushParticles(i, sc.pos, vel[sum(x.n for x in scs[:i]):][:sc.n])
    p.step(2)
    p.synchronize()
    t3 = time.perf_counter()
    p.step(steps)
    p.synchronize()
    t4 = time.perf_counter()
    p.close()
    tf, tb, tp = t1 - t0, t2 - t1, t4 - t3
    return {"scenes": num_scenes, "particles": total, "steps": steps, "checkpoint_every": checkpoint_every,
            "value": total * steps / (tf + tb), "unit": "particle-updates/s (forward + adjoint)",
            "forward_ms_per_step": 1e3 * tf / steps, "adjoint_ms_per_step": 1e3 * tb / steps,
            "plain_batched_ms_per_step": 1e3 * tp / steps, "adjoint_over_forward": tb / tf,
            "adjoint_over_plain": tb / tp, "finite_gradients": bool(np.all(np.isfinite(dmu + dK)))}


def triangles_leg(s, ntri: int, steps: int = 50):
    """Steps of the current state with ntri collision triangles loaded (a
    wavy sheet over the whole box at a fifth of its height, inside the fluid,
    facing down) against the same number of steps without: the split
    collision path and its triangle kernel."""
    import numpy as np
    _, gp, _ = s.params()
    lo = np.array([gp.bnd_min.x, gp.bnd_min.y, gp.bnd_min.z])
    hi = np.array([gp.bnd_max.x, gp.bnd_max.y, gp.bnd_max.z])
    side = max(1, int(round((ntri / 2) ** 0.5)))
    xs = np.linspace(lo[0], hi[0], side + 1)
    ys = np.linspace(lo[1], hi[1], side + 1)
    zz = lo[2] + 0.2 * (hi[2] - lo[2]) + 0.01 * (hi[2] - lo[2]) * np.sin(np.add.outer(xs, ys) * 7.0)
    tris = []
    for i in range(side):
        for j in range(side):
            a = (xs[i], ys[j], zz[i, j])
            b = (xs[i + 1], ys[j], zz[i + 1, j])
            c = (xs[i + 1], ys[j + 1], zz[i + 1, j + 1])
            d = (xs[i], ys[j + 1], zz[i, j + 1])
            tris += [(a, c, b), (a, d, c)]          # wound to face down
    tris = np.array(tris, np.float32)

    def timed():
        s.step(2)
        s.synchronize()
        t0 = time.perf_counter()
        s.step(steps)
        s.synchronize()
        return (time.perf_counter() - t0) * 1e3 / steps

    base = timed()
    s.loadTriangles(tris)
    with_tri = timed()
    s.loadTriangles(np.zeros((0, 3, 3), np.float32))
    return {"triangles": len(tris), "steps": steps, "ms_per_step_without": base, "ms_per_step_with": with_tri,
            "extra_ms": with_tri - base}


def long_run_leg(s, steps: int, first: int, timer_steps: int = 100):
    """The timed system stepped `steps` more steps (graph replays, stage
    timers off): the dam's representative state, after the fluid has left the
    lattice (more neighbours per particle than in the first steps).  The
    per-stage times come from `timer_steps` further steps with the stage
    timers on."""
    import torch
    import enjaparticles_amd._native as N
    s.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    s.step(steps)
    s.synchronize()
    dt = time.perf_counter() - t0
    K, M, Kt = s.countPairs()
    n = s.getNum()
    s.setOption(N.OPT_TIMERS, 1)
    s.resetTimers()
    s.step(timer_steps)
    s.synchronize()
    stages = {k: round(v[0], 5) for k, v in s.stageTimes().items() if v[1] > 0}
    s.setOption(N.OPT_TIMERS, 0)
    last = first + steps - 1
    return {"steps": steps, "first_step": first, "last_step": last,
            "ms_per_step": 1e3 * dt / steps, "value": n * steps / dt, "unit": "particle-updates/s",
            "stages_ms": stages, "stages_steps": [last + 1, last + timer_steps],
            "pairs_per_particle_at_end": {"candidates_27cell": K / n, "accepted_incl_self": M / n,
                                          "candidates_tested": Kt / n}}


def surface_leg(s, res: int, iso: float = 0.5, reps: int = 10):
    """Colour field and iso-surface of the current state (SPH::postProcess):
    wall time per call, each ending in a stream synchronisation."""
    import ctypes as C
    from enjaparticles_amd import _native as N
    lib = N.lib()
    n = C.c_int()
    N.check(lib.rtps_sph_color_field(s._h, res, None))               # buffers, first launch
    N.check(lib.rtps_sph_surface(s._h, res, iso, 0, None, None, C.byref(n)))
    t0 = time.perf_counter()
    for _ in range(reps):
        N.check(lib.rtps_sph_color_field(s._h, res, None))
    field_ms = (time.perf_counter() - t0) * 1e3 / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        N.check(lib.rtps_sph_surface(s._h, res, iso, 0, None, None, C.byref(n)))
    surf_ms = (time.perf_counter() - t0) * 1e3 / reps
    return {"res": res, "nodes": res ** 3, "iso": iso, "triangles": n.value, "field_ms": field_ms,
            "field_plus_surface_ms": surf_ms, "nodes_per_s": res ** 3 / (field_ms * 1e-3)}


def slab_leg(steps: int = 100, warmup: int = 20, slabs: int = 8):
    """One GPU's share of the 1M dam break cut into `slabs` z-slabs: the
    bottom slab's 128x128x8 lattice layers (131,072 particles) and the 2-cell
    halo above them (4 layers, 65,536 particles, the next slab's particles a
    sharded step reads), the first layers of the same jittered lattice in the
    same box, stepped as one plain system.  The halo is stepped too (a sharded
    step would only read it), so the time is the stage cost of one slab with
    its halo; particle-updates/s counts the slab's own particles."""
    import torch
    from enjaparticles_amd import scenes
    from enjaparticles_amd.rtps import RTPS
    import enjaparticles_amd._native as N
    sc = scenes.dam1m()
    own_layers, halo_layers = 64 // slabs, 4
    own = 128 * 128 * own_layers
    pos = scenes.lattice(128, 128, 64, 1002, count=128 * 128 * (own_layers + halo_layers))
    s = RTPS.generateSystemInstance(sc.settings)
    s.pushParticles(pos, np.zeros_like(pos))
    s.setOption(N.OPT_GRAPH, 1)
    s.setOption(N.OPT_TIMERS, 1)
    s.autotune(measure=True)
    s.step(max(0, warmup - 1))
    s.synchronize()
    s.resetTimers()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    s.step(steps)
    s.synchronize()
    dt = time.perf_counter() - t0
    stages = {k: round(v[0], 5) for k, v in s.stageTimes().items() if v[1] > 0}
    n = s.getNum()
    s.close()
    return {"own_particles": own, "halo_particles": n - own, "steps": steps, "ms_per_step": 1e3 * dt / steps,
            "value": own * steps / dt, "unit": "particle-updates/s (own particles)", "stages_ms": stages}


def compare4m_leg(steps: int = 100, launch: str = ""):
    """Plain and batched forward at 4,194,304 particles: one 256x128x128 dam
    break, and 64 scenes of 64x64x16, stepped `steps` steps each (hipGraph),
    with per-stage times.  HBM and L2 counters come from profiling this mode.
    launch: "BD,BF,LOOP" fixes the plain system's neighbour launches."""
    from enjaparticles_amd import scenes
    from enjaparticles_amd.rtps import RTPS, BatchedSPH
    import enjaparticles_amd._native as N
    res = {}
    sc = scenes.dam4m()
    s = RTPS.generateSystemInstance(sc.settings)
    s.pushParticles(sc.pos, sc.vel)
    if launch:
        bd, bf, loop = (int(v) for v in launch.split(","))
        s.setOption(N.OPT_FORCE_VARIANT, loop)
        s.setOption(N.OPT_BLOCK_DENSITY, bd)
        s.setOption(N.OPT_BLOCK_FORCE, bf)
        s.step(1)
    else:
        s.autotune(measure=True)
    s.step(9)
    s.synchronize()
    t0 = time.perf_counter()
    s.step(steps)
    s.synchronize()
    tp = time.perf_counter() - t0
    s.setOption(N.OPT_TIMERS, 1)
    s.resetTimers()
    s.step(5)
    K, M, _ = s.countPairs()
    res["plain"] = {"particles": sc.n, "value": sc.n * steps / tp, "ms_per_step": 1e3 * tp / steps,
                    "stages_ms": {k: round(v[0], 4) for k, v in s.stageTimes().items()},
                    "pairs_per_particle_at_end": {"candidates_27cell": K / sc.n, "accepted_incl_self": M / sc.n}}
    s.close()
    scs = scenes.batched_equal(64, 64, 64, 16, seed=1007)
    total = sum(x.n for x in scs)
    b = BatchedSPH([x.settings for x in scs])
    for i, x in enumerate(scs):
        b.pushParticles(i, x.pos, x.vel)
    b.step(10)
    b.synchronize()
    t0 = time.perf_counter()
    b.step(steps)
    b.synchronize()
    tb = time.perf_counter() - t0
    b.setOption(3, 1)
    b.resetTimers()
    b.step(5)
    res["batched"] = {"scenes": 64, "particles": total, "value": total * steps / tb, "ms_per_step": 1e3 * tb / steps,
                      "stages_ms": {k: round(v[0], 4) for k, v in b.stageTimes().items()}}
    b.close()
    # The two workloads evolve differently (the slabs are 16 particles deep, the
    # dam 128): neighbour counts of one scene, stepped alone as a plain system
    # the same number of steps, put the two on a per-pair footing.
    p0 = RTPS.generateSystemInstance(scs[0].settings)
    p0.pushParticles(scs[0].pos, scs[0].vel)
    p0.step(10 + steps + 5)
    K0, M0, _ = p0.countPairs()
    n0 = p0.getNum()
    p0.close()
    res["batched"]["pairs_per_particle_at_end_scene0"] = {"candidates_27cell": K0 / n0, "accepted_incl_self": M0 / n0}
    mp = res["plain"]["pairs_per_particle_at_end"]["accepted_incl_self"]
    res["pair_updates_ratio_batched_over_plain"] = (res["batched"]["value"] * M0 / n0) / (res["plain"]["value"] * mp)
    res["steps"] = steps
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--chunk", type=int, default=100, help="timed steps between pair counts")
    ap.add_argument("--no-stage-timers", action="store_true", help="time the graph without any events")
    ap.add_argument("--stage-steps", type=int, default=20,
                    help="steps after the timed ones with events around every stage (per-stage breakdown)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=100)
    ap.add_argument("--cpu-single-steps", type=int, default=2)
    ap.add_argument("--profile", action="store_true", help="timed loop only (for rocprofv3 runs)")
    ap.add_argument("--launch", default="", metavar="BD,BF,LOOP",
                    help="fixed neighbour launches instead of autotuning: density and force threads per block and "
                         "the force loop (profiling runs: counters perturb the autotuner's timings)")
    ap.add_argument("--shared-rows", type=int, default=-1, choices=(-1, 0, 1),
                    help="RTPS_OPT_SHARED_ROWS of the plain system (-1 = the library default)")
    ap.add_argument("--tile-queues", type=int, default=0, choices=(0, 1),
                    help="1 = density and force as resident grids over per-XCD tile queues (RTPS_OPT_TILE_QUEUES)")
    ap.add_argument("--force-variant", type=int, default=5,
                    help="process-default force loop before autotuning: 5 = LDS-staged rows with accepted-candidate "
                         "bits, 4 = LDS-staged rows, 0 = direct")
    ap.add_argument("--long-steps", type=int, default=1000,
                    help="steps of the long-run leg that continues the timed system (0 = skip)")
    ap.add_argument("--triangles", type=int, default=2048,
                    help="collision triangles of the triangle leg on the final state (0 = skip)")
    ap.add_argument("--surface-res", type=int, default=128,
                    help="colour-field lattice of the surface leg on the final state (0 = skip)")
    ap.add_argument("--batched-steps", type=int, default=30, help="steps of the batched leg (0 = skip)")
    ap.add_argument("--adjoint-steps", type=int, default=200, help="steps of the fwd+adjoint leg (0 = skip)")
    ap.add_argument("--badj-steps", type=int, default=100, help="steps of the batched fwd+adjoint leg (0 = skip)")
    ap.add_argument("--slab-steps", type=int, default=100,
                    help="steps of the one-z-slab leg (one GPU's share of an 8-way split; 0 = skip)")
    ap.add_argument("--compare-4m", type=int, default=0, metavar="STEPS",
                    help="only run the plain vs batched comparison at 4,194,304 particles for STEPS steps")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: runs on one GPU only (asked for --gpus {args.gpus}); "
              "the multi-GPU path (z-slab sharding with halo exchange) is a design only and is not built",
              file=sys.stderr)
        sys.exit(2)

    import torch
    if not torch.cuda.is_available():
        print("bench.py: no GPU visible", file=sys.stderr)
        sys.exit(1)
    if args.compare_4m > 0:
        print(json.dumps({"compare_4m": compare4m_leg(args.compare_4m, args.launch)}))
        return
    from enjaparticles_amd import scenes
    from enjaparticles_amd.rtps import RTPS
    import enjaparticles_amd._native as N

    sc = scenes.dam1m()
    s = RTPS.generateSystemInstance(sc.settings)
    s.pushParticles(sc.pos, sc.vel)
    s.setOption(N.OPT_GRAPH, 1)
    s.setOption(N.OPT_FORCE_VARIANT, args.force_variant)
    s.setOption(N.OPT_TILE_QUEUES, args.tile_queues)
    if args.shared_rows >= 0:
        s.setOption(N.OPT_SHARED_ROWS, args.shared_rows)
    # the timed steps record HIP events around the force stage on every fourth step (RTPS_OPT_TIMERS 3:
    # two events inside the graph of steps 0, 4, 8, ... of the timed region, the others replay the
    # graph without events, whose records cost a few microseconds each); the per-stage breakdown comes
    # from the steps that follow
    s.setOption(N.OPT_TIMERS, 0 if args.no_stage_timers else 3)
    if args.launch:
        bd, bf, loop = (int(v) for v in args.launch.split(","))
        s.setOption(N.OPT_FORCE_VARIANT, loop)
        s.setOption(N.OPT_BLOCK_DENSITY, bd)
        s.setOption(N.OPT_BLOCK_FORCE, bf)
        tune = {"block_density": bd, "block_force": bf, "force_variant": loop}
        s.step(1)
    else:
        tune = s.autotune(measure=True)      # the first warmup step tunes the neighbour launches
    s.step(max(0, args.warmup - 1))
    s.synchronize()
    print(f"# warmup {args.warmup} steps done", file=sys.stderr, flush=True)

    # Timed region: `steps` graph replays in chunks; every fourth step records
    # HIP events around its force kernel on the handle's stream inside the
    # graph.  The pair counts of the FLOP model are taken between chunks,
    # outside the timing.
    s.resetTimers()
    counts = [s.countPairs()]
    elapsed, done = 0.0, 0
    marker = _Marker()
    while done < args.steps:
        k = min(args.chunk, args.steps - done)
        s.synchronize()
        torch.cuda.synchronize()
        marker.push("timed")
        t0 = time.perf_counter()
        s.step(k)
        s.synchronize()
        torch.cuda.synchronize()
        elapsed += time.perf_counter() - t0
        marker.pop()
        done += k
        counts.append(s.countPairs())
        print(f"# {done} steps, {elapsed:.3f} s", file=sys.stderr, flush=True)
    n = s.getNum()
    if args.profile:
        # a plain 256 MiB device copy of known size, to calibrate the HBM
        # byte counters (FETCH_SIZE / WRITE_SIZE) of the same profiling run
        src = torch.empty(64 << 20, dtype=torch.float32, device="cuda").fill_(1.0)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print("# reference copy: 268435456 bytes read + written", file=sys.stderr, flush=True)
        return
    force_timed = s.stageTimes()["force"] if not args.no_stage_timers else None
    stages = {}
    if not args.no_stage_timers and args.stage_steps > 0:
        # per-stage breakdown: the next stage_steps steps with events around every stage
        s.setOption(N.OPT_TIMERS, 1)
        s.resetTimers()
        s.step(args.stage_steps)
        s.synchronize()
        stages = s.stageTimes()
    s.setOption(N.OPT_TIMERS, 0)
    # pair counts per particle, averaged over the chunk boundaries (trapezoid)
    sizes = [min(args.chunk, args.steps - c * args.chunk) for c in range(len(counts) - 1)]
    avg = [sum(0.5 * (counts[c][q] + counts[c + 1][q]) * sizes[c] for c in range(len(sizes))) / args.steps / n
           for q in range(3)]
    Kp, Mp, Ktp = avg
    # force FLOP model: 8 per candidate pair + 35 per accepted pair (self excluded)
    force_flop = n * (8.0 * Kp + 35.0 * (Mp - 1.0))
    dens_flop = n * (8.0 * Kp + 4.0 * Mp)
    ms_per_step = 1e3 * elapsed / args.steps
    out = {
        "metric": METRIC,
        "value": n * args.steps / elapsed,
        "unit": "particle-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded jittered lattice, SplitMix64)",
        "config": {"workload": "plain SPH dam break, 1,048,576 particles (128x128x64), h=0.04, dt=0.003, "
                               "walls on, hipGraph per step", "particles": n, "steps_per_s": 1e3 / ms_per_step},
        "pairs_per_particle": {"candidates_27cell": Kp, "accepted_incl_self": Mp, "candidates_tested": Ktp},
        "autotune": tune,
    }
    if force_timed is not None and force_timed[1] > 0:
        # the dominant kernel: HIP events around the force stage of every fourth timed step
        force_ms = force_timed[0]
        achieved = force_flop / (force_ms * 1e-3) / 1e12
        out["roofline"] = {
            "bound": "valu",
            "kernel": "k_force (fused force + walls + leapfrog)",
            "achieved": achieved,
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32_TFLOPS,
            "peak_nonpacked": PEAK_VALU_FP32_TFLOPS,
            "frac_nonpacked": achieved / PEAK_VALU_FP32_TFLOPS,
            "traffic": measured_traffic("k_force", n, int(tune.get("block_force", 64))),
            "traffic_unit": "HBM bytes per launch on the timed state (rocprofv3 FETCH_SIZE/WRITE_SIZE, calibrated)",
            "algorithmic_bytes": 100.0 * n,
            "flop_per_launch": force_flop,
            "avg_launch_ms": force_ms,
            "launches": force_timed[1],
        }
    if stages:
        # per-stage breakdown (the steps right after the timed ones); HBM side:
        # algorithmic (minimum) bytes of each stage over its measured time
        import numpy as np
        cells = s.nb_cells + 1
        occupied = int(np.count_nonzero(s.read("cell_start") != np.uint32(0xFFFFFFFF)))
        key_bits = (s.nb_cells + 1).bit_length() + 4      # cell key bits (incl. pad) + 4 sub-cell bits
        passes = (key_bits + 7) // 8
        low = max(key_bits - 16, min(key_bits, 8))         # the bucket sort's bits inside a bucket
        bucket = low <= 8 and (1 << (key_bits - low)) * 1024 >= n
        # the step runs the hash inside the sort's first kernel and the permute
        # inside the cellindices kernel: those stage pairs are reported together
        alg = {
            # bucket sort: pos read, keys written, start[] cleared; keys read, pairs
            # placed by bucket; pairs read and written sorted.  LSD radix sort: pos
            # read, key + idx written, start[] cleared; per pass pairs read + written
            "hash+sort": (("hash", "sort"), (48.0 * n if bucket else 24.0 * n + passes * 16.0 * n) + 4.0 * cells),
            # keys read, start/end of occupied cells written; idx + 3 float4 gathered, 3 float4 written
            "cellindices+permute": (("cellindices", "permute"), 4.0 * n + 8.0 * occupied + 100.0 * n),
            # own position read, density written (neighbours: cache)
            "density": (("density",), 20.0 * n),
            # pos, veleval|density, vel, idx read; 3 float4 written
            "force": (("force",), 100.0 * n),
        }
        hbm = {}
        for k, (names, b) in alg.items():
            ms = sum(stages[q][0] for q in names)
            gbs = b / (ms * 1e-3) / 1e9
            hbm[k] = {"bytes": b, "ms": round(ms, 5), "GBps": round(gbs, 1), "frac_of_peak": round(gbs / PEAK_HBM_GBS, 4)}
        alg = {k: b for k, (_, b) in alg.items()}
        step_bytes = sum(alg.values())
        hbm["step"] = {"bytes": step_bytes, "ms": round(ms_per_step, 5),
                       "GBps": round(step_bytes / (ms_per_step * 1e-3) / 1e9, 1),
                       "frac_of_peak": round(step_bytes / (ms_per_step * 1e-3) / 1e9 / PEAK_HBM_GBS, 4)}
        out["hbm"] = {"peak_GBps": PEAK_HBM_GBS, "occupied_cells": occupied, "stages": hbm}
        first = args.warmup + args.steps + 1
        out["stages_ms"] = {k: round(v[0], 5) for k, v in stages.items()}
        out["stages_steps"] = [first, first + args.stage_steps - 1]
        dens_ms = stages["density"][0]
        out["density_roofline"] = {
            "achieved_tflops": dens_flop / (dens_ms * 1e-3) / 1e12,
            "frac": dens_flop / (dens_ms * 1e-3) / 1e12 / PEAK_FP32_TFLOPS,
            "frac_nonpacked": dens_flop / (dens_ms * 1e-3) / 1e12 / PEAK_VALU_FP32_TFLOPS,
            "avg_launch_ms": dens_ms,
        }
    if args.long_steps > 0 and not args.profile:
        print("# long-run leg ...", file=sys.stderr, flush=True)
        out["long_run"] = long_run_leg(s, args.long_steps,
                                       args.warmup + args.steps + (args.stage_steps if stages else 0) + 1)
        print(f"# long run: {out['long_run']['value']:.3e} p-u/s over steps {out['long_run']['first_step']}-"
              f"{out['long_run']['last_step']}", file=sys.stderr, flush=True)
    if args.surface_res > 0 and not args.profile:
        print("# surface leg ...", file=sys.stderr, flush=True)
        out["surface"] = surface_leg(s, args.surface_res)
        print(f"# surface: field {out['surface']['field_ms']:.3f} ms, +surface "
              f"{out['surface']['field_plus_surface_ms']:.3f} ms, {out['surface']['triangles']} triangles",
              file=sys.stderr, flush=True)
    if args.triangles > 0 and not args.profile:
        print("# triangle leg ...", file=sys.stderr, flush=True)
        out["triangle_collision"] = triangles_leg(s, args.triangles)
        print(f"# triangles: {out['triangle_collision']['ms_per_step_with']:.3f} ms/step with "
              f"{out['triangle_collision']['triangles']} vs {out['triangle_collision']['ms_per_step_without']:.3f}",
              file=sys.stderr, flush=True)
    s.close()
    if args.batched_steps > 0:
        print("# batched leg ...", file=sys.stderr, flush=True)
        out["batched"] = batched_leg(args.batched_steps)
        print(f"# batched: {out['batched']['value']:.3e} p-u/s, x{out['batched']['ratio']:.2f} vs sequential plain",
              file=sys.stderr, flush=True)
    if args.adjoint_steps > 0:
        print("# fwd+adjoint leg ...", file=sys.stderr, flush=True)
        out["fwd_adjoint"] = adjoint_leg(args.adjoint_steps)
        print(f"# fwd+adjoint: {out['fwd_adjoint']['value']:.3e} p-u/s, adjoint/forward "
              f"{out['fwd_adjoint']['adjoint_over_forward']:.2f}", file=sys.stderr, flush=True)
    if args.slab_steps > 0:
        print("# one-slab leg ...", file=sys.stderr, flush=True)
        out["one_slab"] = slab_leg(args.slab_steps)
        print(f"# one slab: {out['one_slab']['ms_per_step']:.3f} ms/step, {out['one_slab']['value']:.3e} p-u/s "
              f"(own particles)", file=sys.stderr, flush=True)
    if args.badj_steps > 0:
        print("# batched fwd+adjoint leg ...", file=sys.stderr, flush=True)
        out["batched_fwd_adjoint"] = batched_adjoint_leg(args.badj_steps)
        print(f"# batched fwd+adjoint: {out['batched_fwd_adjoint']['value']:.3e} p-u/s, adjoint/forward "
              f"{out['batched_fwd_adjoint']['adjoint_over_forward']:.2f}", file=sys.stderr, flush=True)
    if not args.no_cpu_baseline:
        print("# cpu baseline ...", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(sc, args.cpu_steps, args.cpu_single_steps)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
