"""Benchmark: cell-steps/s of the 2D step on one MI355X.

Workload (the single-grid configuration the headline metric is quoted on):
a 2048 x 2048 MAC grid, all interior cells FLUID inside one-cell SOLID walls,
semi-Lagrangian mode.  One step = semi-Lagrangian advection of u, v and a
passive density, then the MIC(0)-preconditioned CG projection to a relative
infinity-norm residual of 1e-6 and the velocity update with extrapolation.  Initial velocity: i.i.d. N(0,1) from
numpy.random.default_rng(1) (u then v), Gaussian-smoothed with sigma = 8 cells
(periodic padding), projected divergence-free on the GPU; density: 16 Gaussian
blobs drawn from the same generator; dt = 1.0 * dx / max|u| (CFL 1) fixed
after the projection.  Gravity is off (a closed box of smoke-like fluid).

MIC(0) runs with tau = 0.97 (the value the 64x64 dam-break configuration
names): with the reference's default 0.999 the FP32 preconditioner on this
all-Neumann (singular) system stalls above the 1e-6 relative residual and
every step runs into the iteration cap.

`roofline` reports the dominant kernel class: algorithmic bytes per launch
(BYTES_PER_CELL x cells) over its average launch time, measured with HIP
events on the library's stream inside the timed region.  With the tuning's
fused_pcg (the default) a PCG iteration is two launches: the backward-solve
class also runs the next iteration's SpMV (reported as
`pcg_backward_solve+spmv`, with both kernels' arrays) and the forward-solve
class the p/r update (`pcg_forward_solve+axpy`).  `--no-events` drops every
HIP event (for profiler runs; no roofline).  Launches that start
after convergence exit at once; the average is taken over the launches that
did work (PCG iterations + 1 per step for the two solves).  Each timed launch
adds two events to the stream, and events around every kernel cost ~8% of
the step, so the timed region times the dominant class only; the dominant
class and `kernel_ms_per_step` come from 2 steps with every class timed,
run after the warmup and before the timed region.

Value = cells x steps / wall time of the 2048^2 steps taken by a simulation
whose fs2d_step replays captured HIP graphs (fs2d_set_graph), with no events;
a second simulation, in lockstep with it up to the timed regions (same state,
bit-identical results, the same iteration counts), takes the same steps
uncaptured with events on the dominant class (`no_graph_*`, and the source of
`roofline` and `kernel_ms_per_step`).  `--no-graph` leaves the captured twin
out; the value is then the uncaptured one.

Both simulations, and both ensemble batches, run the launch shapes the
autotuner picks (fs2d_autotune / fs2d_batch_autotune, which change no
result), read from `--tune-cache` when it holds the problem's key; the choice
is reported as `tuning`.

`ensemble` (the batched half of the metric): 256 independent PIC/FLIP dam
breaks, side uniform in {64, 96, ..., 512} (seed 2), fluid block width and
height uniform in [0.2, 0.7], 4 particles per fluid cell, FLIP blend 0.95,
dt 0.002, PCG to 1e-6 with converged members masked, stepped as one batch
whose step replays captured HIP graphs (fs2d_batch_set_graph; the value),
by uncaptured batch steps of a second batch in lockstep (`no_graph_*`) and,
for comparison, by looping fs2d_step over the same 256 simulations.  Value =
total cells x steps / wall time, with no events; `kernel_ms_per_step` comes
from 2 further uncaptured steps with every class timed.

`pic`: the 512 x 512 PIC/FLIP dam break (fluid block 40% x 60% of the
domain, 4 particles per fluid cell seeded uniformly at random with seed 0,
FLIP blend 0.95, RK2 particle advection, dt 0.002, PCG to 1e-6): captured
steps timed without events, and each particle stage's particles/s from 2
further steps with every class timed.

`adjoint`: the differentiable rollout, 256 x 256 all-fluid smoke, 64 steps,
dt 0.01, velocity seeded like the 2048^2 workload (seed 3), loss = L2 of the
final density against 8 blobs (seed 4): forward rollout with a checkpoint
every 8 steps, then the reverse pass (recomputing each segment); ratio =
reverse time / forward time, with checkpoints and with every state kept.

`grid3d`: the 3D MAC grid (include/MACGrid3D.h) at 256^3 with N(0,1) FP32
u, v, w: velInterp (Catmull-Rom in the velocity tap order) at 2^24 uniform
random points, once in random order and once sorted by cell, and velDiv at
every cell (16 algorithmic bytes per cell); HIP events on the grid's stream.

`cpu_baseline` times the oracle (FP64 CPU restatement of the reference step,
OpenMP, one thread per physical core of this process's affinity mask, at most
16, OMP_PROC_BIND=close) on the same starting state with PCG capped at three
iteration counts, and extrapolates to the GPU's mean iteration count: step =
fixed + iters x per-iteration, fitted on the outer two caps and checked on the
middle one.  A full oracle step at 2048^2 takes minutes.  The host's CPU model,
logical and physical core counts are reported beside it.

`config.pcg_iters_max` and `config.pcg_capped_steps` show whether any timed
step stopped at the iteration cap instead of converging.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "fluid-sim_amd"))
sys.path.insert(0, ROOT)

METRIC = "cell-steps/s (advect+PCG project) at 2048² & batched; % of HBM peak GB/s"
HBM_PEAK_GBS = 8000.0
PROFILE_STEPS = 2  # steps with every kernel class timed, outside the timed regions

# Algorithmic HBM bytes per cell per launch of each kernel class: every array
# the kernel must read or write, counted once (neighbour re-reads are cache
# hits by design).  FP32 fields, uint8 flags, FP64 level set.
BYTES_PER_CELL = {
    "pcg_forward_solve": 20,    # r, P, fx, fy in; q out
    "pcg_backward_solve": 20,   # q, P, gx, gy in; z out
    # the backward solve with the next SpMV in one launch: q, P, gx, gy in, z
    # out; s, Adiag, coupling byte in, s', w out (z and q counted once)
    "pcg_backward_solve+spmv": 37,
    # the p/r update with the forward solve in one launch: p, r, s, w in, p, r
    # out; P, fx, fy in, q out (r counted once each way)
    "pcg_forward_solve+axpy": 40,
    "pcg_spmv_dot": 21,         # z, s, Adiag, coupling byte in; s', w out
    "pcg_axpy_norm": 24,        # p, r, s, w in; p, r out
    "mic0_factor": 25,          # Adiag, coupling byte in; P, fx, fy, gx, gy out
    "assemble": 34,             # cell, phi (f64), u, v in; Adiag, coupling, r, p, s out
    "sl_advect": 24,            # u, v, density in and out
}
# launches that did work per step: `iters` loop launches (+1 initial solve)
EFFECTIVE = {"pcg_forward_solve": 1, "pcg_backward_solve": 1, "pcg_spmv_dot": 0, "pcg_axpy_norm": 0}


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "pcg_2048_hbm_traffic.md")


def measured_traffic(kernel, n):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary (calibrated against a copy kernel), for the 2048^2
    workload it was measured on; None otherwise."""
    if n != 2048 or not os.path.exists(TRAFFIC_FILE):
        return None
    for line in open(TRAFFIC_FILE):
        f = line.split()
        if len(f) == 3 and f[0] == "TRAFFIC_BYTES" and f[1] == kernel:
            return float(f[2])
    return None


def make_state(n, seed=1):
    from scipy.ndimage import gaussian_filter
    rng = np.random.default_rng(seed)
    u = gaussian_filter(rng.standard_normal((n, n + 1)), 8.0, mode="wrap")
    v = gaussian_filter(rng.standard_normal((n + 1, n)), 8.0, mode="wrap")
    J, I = np.mgrid[0:n, 0:n].astype(np.float64) + 0.5
    dens = np.zeros((n, n))
    for _ in range(16):
        cx, cy = rng.uniform(0.15, 0.85, size=2) * n
        rad = rng.uniform(n / 64.0, n / 16.0)
        dens += np.exp(-((I - cx) ** 2 + (J - cy) ** 2) / (2 * rad * rad))
    return u.astype(np.float32), v.astype(np.float32), dens.astype(np.float32)


def walls_all_fluid(n):
    c = np.full((n, n), 1, dtype=np.uint8)
    c[0, :] = c[-1, :] = 2
    c[:, 0] = c[:, -1] = 2
    return c


def build_gpu_sim(n, tol, tau, max_iters):
    import fs2d
    cells = walls_all_fluid(n)
    sim = fs2d.Sim(cells, k=0, dt=1e-3, dx=1.0 / n, mode=fs2d.SEMILAGRANGIAN, gravity=(0.0, 0.0),
                   tau=tau, tol=tol, max_iters=max_iters, static_cells=True, density=True)
    u, v, d = make_state(n)
    sim.set(fs2d.F_U, u)
    sim.set(fs2d.F_V, v)
    sim.set(fs2d.F_DENSITY, d)
    sim.run_stage(fs2d.STAGE_PROJECT)
    sim.run_stage(fs2d.STAGE_UPDATE_VELOCITY)
    u, v = sim.get(fs2d.F_U), sim.get(fs2d.F_V)
    dt = 1.0 * (1.0 / n) / float(max(np.abs(u).max(), np.abs(v).max()))
    sim.set_dt(dt)
    return sim, cells, dt


def host_cpus():
    """What the host offers this process: logical CPUs in its affinity mask,
    the physical cores among them (distinct (package, core id) pairs in
    /proc/cpuinfo) and the CPU model name."""
    allowed = sorted(os.sched_getaffinity(0))
    model, cores = None, set()
    try:
        phys, core, proc = None, None, None
        for line in open("/proc/cpuinfo"):
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k == "processor":
                proc = int(v)
            elif k == "physical id":
                phys = v
            elif k == "core id":
                core = v
            elif k == "model name" and model is None:
                model = v
            elif not k and proc is not None:
                if proc in allowed:
                    cores.add((phys, core))
                proc = phys = core = None
        if proc is not None and proc in allowed:
            cores.add((phys, core))
    except OSError:
        pass
    return {"nproc": len(allowed), "os_cpu_count": os.cpu_count(),
            "physical_cores": len(cores) if cores else None, "model": model}


def cpu_baseline(state, cells, n, dt, tol, tau, threads, iters_gpu, host, caps=(20, 40, 60)):
    """The oracle (CPU restatement, OpenMP) on steps of the same workload from
    the GPU's starting state with PCG capped at caps[0], caps[1] and caps[2]
    iterations.  A full step at 2048^2 takes minutes, so the per-step time is
    extrapolated: fixed + iterations x per-iteration, fitted on caps[0] and
    caps[2] and checked against the step measured at caps[1]."""
    from oracle import oracle as O
    O.set_threads(threads)
    times = []
    for cap in caps:
        o = O.Oracle(cells, k=0, dt=dt, dx=1.0 / n, mode=O.MODE_SL, level_set=False, gravity=(0.0, 0.0),
                     tau=tau, tol=tol, max_iters=cap, sl_inplace=False, density=True)
        for f, a in zip((O.F_U, O.F_V, O.F_DENS), state):
            o.set(f, a.astype(np.float64))
        t0 = time.perf_counter()
        o.step()
        times.append(time.perf_counter() - t0)
        done = o.stats()["last_iters"]
        o.close()
        if done != cap:
            raise RuntimeError("oracle converged in %d < %d iterations: sample too small" % (done, cap))
    per_iter = (times[2] - times[0]) / (caps[2] - caps[0])
    fixed = times[0] - caps[0] * per_iter
    step = fixed + iters_gpu * per_iter
    check = fixed + caps[1] * per_iter
    return {"value": n * n / step, "unit": "cell-steps/s", "cores": threads, "kind": "port",
            "extrapolated": True,
            "host": dict(host, omp_proc_bind=os.environ.get("OMP_PROC_BIND"),
                         omp_places=os.environ.get("OMP_PLACES")),
            "fit": {"caps": list(caps), "seconds": times, "fixed_s": fixed, "per_iter_s": per_iter,
                    "check_rel_err": (check - times[1]) / times[1]},
            "sample": "oracle (FP64, OpenMP, %d threads) steps of the %dx%d workload from the GPU's starting state "
                      "with PCG capped at %d, %d and %d iterations (%.1f, %.1f, %.1f s); extrapolated: %.3f s fixed "
                      "+ %.4f s/iteration (fitted on %d and %d, off by %.1f%% at %d), scaled to the GPU's %.1f "
                      "iterations per step" % (threads, n, n, caps[0], caps[1], caps[2], times[0], times[1], times[2],
                                               fixed, per_iter, caps[0], caps[2], 100 * (check - times[1]) / times[1],
                                               caps[1], iters_gpu)}


def ensemble_leg(count, steps, warmup, loop_steps, tune_cache=None, graph=True):
    """The batched ensemble: stepped by replaying captured graphs (the
    value) and by uncaptured batch steps, two batches in lockstep from the
    same state (their members stay bit-identical); and the loop over the same
    simulations."""
    import fs2d
    from fs2d import scenes
    members = scenes.ensemble_members(count)
    cells = int(sum(c.shape[0] * c.shape[1] for c, _ in members))
    batch = fs2d.Batch(members)
    gbatch = fs2d.Batch(members) if graph else None
    twins = [batch] + ([gbatch] if graph else [])
    if graph:
        gbatch.set_graph(True)
    for b in twins:
        b.step()
    tune = batch.autotune(iters=24, cache=tune_cache)
    if graph:
        gbatch.set_tuning(tune["pcg_grid"], tune["sweep_prefetch"], tune["fused_pcg"])
    for _ in range(warmup):
        for b in twins:
            b.step()
    tg = None
    if graph:
        t0 = time.perf_counter()
        for _ in range(steps):
            gbatch.step()  # returns after the step's work (fs2d_batch_step synchronizes)
        tg = (time.perf_counter() - t0) / steps
        gbatch.close()
    t0 = time.perf_counter()
    for _ in range(steps):
        batch.step()
    tb = (time.perf_counter() - t0) / steps
    tv = tg if graph else tb
    # per-class breakdown from further steps (events around every launch cost time)
    batch.reset_kernel_times()
    batch.set_timing(True)
    for _ in range(PROFILE_STEPS):
        batch.step()
    kt = batch.kernel_times()
    batch.set_timing(False)
    its = [s.stats()["pcg_iters"] for s in batch.sims]
    batch.close()
    tl = None
    if loop_steps > 0:  # 0: batch only (profiling runs)
        singles = [fs2d.Sim(c, **kw) for c, kw in members]
        for _ in range(warmup):
            for s in singles:
                s.step()
        t0 = time.perf_counter()
        for _ in range(loop_steps):
            for s in singles:
                s.step()
        tl = (time.perf_counter() - t0) / loop_steps
        for s in singles:
            s.close()
    return {
        "metric": "cell-steps/s of the batched ensemble", "value": cells / tv, "unit": "cell-steps/s",
        "ms_per_step": tv * 1e3, "steps": steps, "warmup": warmup, "graph": graph,
        "no_graph_value": cells / tb, "no_graph_ms_per_step": tb * 1e3,
        "tuning": tune,
        "loop_value": cells / tl if tl else None, "loop_ms_per_step": tl * 1e3 if tl else None,
        "loop_steps": loop_steps, "speedup_vs_loop": tl / tv if tl else None,
        "config": {"workload": "%d PIC/FLIP dam breaks, sides 64..512 (seed 2), block fractions U[0.2,0.7], "
                               "k=2, FLIP 0.95, dt 0.002, PCG 1e-6 masked per member" % count,
                   "members": count, "cells": cells, "pcg_iters_max": int(max(its)),
                   "pcg_iters_mean": float(np.mean(its))},
        "kernel_ms_per_step": {k: v[0] / PROFILE_STEPS for k, v in kt.items() if v[1]},
    }


def pic_leg(n=512, steps=20, warmup=2):
    """The 512^2 PIC/FLIP dam break: one step from rest, `warmup` steps, then
    `steps` timed steps replayed from captured graphs (no events); the
    per-class breakdown, and each particle stage's rate, from 2 further steps
    with every class timed."""
    import fs2d
    from fs2d import scenes
    sim = fs2d.Sim(scenes.dam_break(n), k=2, dt=0.002, dx=1.0 / n, mode=fs2d.PICFLIP, alpha=0.05, tau=0.97,
                   tol=1e-6, max_iters=500, seed=0, uniform_seeding=True, rk2=True)
    npart = int(sim.stats()["n_particles"])
    sim.step()  # leave the zero-velocity start
    sim.set_graph(True)
    for _ in range(warmup):
        sim.step()
    iters = []
    sim.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        sim.step()
        iters.append(sim.stats()["pcg_iters"])
    sim.synchronize()
    t = (time.perf_counter() - t0) / steps
    sim.set_graph(False)
    sim.reset_kernel_times()
    sim.set_timing(True)
    for _ in range(PROFILE_STEPS):
        sim.step()
    kt = sim.kernel_times()
    sim.set_timing(False)
    tune = sim.tuning()
    sim.close()
    kms = {k: v[0] / PROFILE_STEPS for k, v in kt.items() if v[1]}
    rates = {k: npart / (kms[k] / 1e3) for k in ("particle_sort", "p2g", "g2p", "particle_advect") if k in kms}
    return {
        "metric": "cell-steps/s of the PIC/FLIP step", "value": n * n / t, "unit": "cell-steps/s",
        "ms_per_step": t * 1e3, "steps": steps, "warmup": warmup, "graph": True,
        "particles": npart, "particle_steps_per_s": npart / t,
        "config": {"workload": "%dx%d PIC/FLIP dam break: fluid block 40%% x 60%%, 4 particles per fluid cell "
                               "seeded uniformly at random (seed 0), FLIP blend 0.95, RK2 particle advection, "
                               "dt 0.002, MIC(0)-PCG tau 0.97 to 1e-6" % (n, n),
                   "grid": [n, n], "pcg_iters_mean": float(np.mean(iters)), "pcg_iters_max": int(max(iters)),
                   "tuning": tune},
        "kernel_ms_per_step": kms,
        "particles_per_s_by_stage": rates,
    }


def adjoint_leg(n=256, steps=64, every=8):
    """Forward rollout and checkpointed reverse pass of the smoke workload."""
    import fs2d
    from fs2d import scenes
    sim = fs2d.Sim(scenes.all_fluid(n), k=0, dt=0.01, dx=1.0 / n, mode=fs2d.SEMILAGRANGIAN, gravity=(0.0, 0.0),
                   tau=0.97, tol=1e-6, max_iters=4000, static_cells=True, density=True)
    u, v, d = scenes.smoke_state(n, 3)
    for f, a in zip((fs2d.F_U, fs2d.F_V, fs2d.F_DENSITY), (u, v, d)):
        sim.set(f, a)
    sim.run_stage(fs2d.STAGE_PROJECT)
    sim.run_stage(fs2d.STAGE_UPDATE_VELOCITY)
    x0 = [sim.get(f) for f in (fs2d.F_U, fs2d.F_V, fs2d.F_DENSITY)]
    target = scenes.target_density(n)
    res = {}
    for ev in (every, 1):
        adj = fs2d.Adjoint(sim, checkpoint_every=ev, max_steps=steps)
        for rep in range(2):  # the first pass warms up
            for f, a in zip((fs2d.F_U, fs2d.F_V, fs2d.F_DENSITY), x0):
                sim.set(f, a)
            sim.synchronize()
            t0 = time.perf_counter()
            adj.forward(steps)
            sim.synchronize()
            tf = time.perf_counter() - t0
            dT = sim.get(fs2d.F_DENSITY)
            adj.set(fs2d.F_U, np.zeros_like(x0[0]))
            adj.set(fs2d.F_V, np.zeros_like(x0[1]))
            adj.set(fs2d.F_DENSITY, 2.0 * (dT - target))
            t0 = time.perf_counter()
            adj.backward()
            sim.synchronize()
            tb = time.perf_counter() - t0
        res[ev] = (tf, tb, float(((dT - target) ** 2).sum()),
                   float(np.linalg.norm(adj.get(fs2d.F_U))) + float(np.linalg.norm(adj.get(fs2d.F_V))))
        adj.close()
    sim.close()
    tf, tb, loss, gnorm = res[every]
    return {
        "metric": "reverse-pass time / forward time per step", "value": tb / tf, "unit": "ratio",
        "higher_is_better": False,
        "fwd_ms_per_step": tf / steps * 1e3, "bwd_ms_per_step": tb / steps * 1e3,
        "ratio_every_state_kept": res[1][1] / res[1][0], "bwd_ms_per_step_every_state_kept": res[1][1] / steps * 1e3,
        "loss": loss, "grad_norm_uv": gnorm,
        "config": {"workload": "%dx%d all-fluid SL smoke rollout, %d steps, dt 0.01, checkpoint every %d, "
                               "loss = |d_T - target|^2" % (n, n, steps, every)},
    }


def grid3d_leg(n=256, npts=1 << 24, reps=10):
    """The 3D MAC grid (MACGrid3D.h): velInterp at npts points of an n^3 grid,
    random and cell-sorted, and velDiv over every cell; kernel times from HIP
    events recorded on the grid's stream."""
    import ctypes
    import fs2d
    from fs2d import grid3d as G
    hip = ctypes.CDLL("libamdhip64.so.7")  # the runtime libfs2d.so is linked to
    vp = ctypes.c_void_p
    hip.hipMalloc.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int]
    hip.hipFree.argtypes = [vp]
    hip.hipEventCreate.argtypes = [ctypes.POINTER(vp)]
    hip.hipEventRecord.argtypes = [vp, vp]
    hip.hipEventSynchronize.argtypes = [vp]
    hip.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), vp, vp]
    hip.hipEventDestroy.argtypes = [vp]

    def dmalloc(nbytes):
        p = vp()
        assert hip.hipMalloc(ctypes.byref(p), nbytes) == 0
        return p

    g = G.MACGrid3D(n, n, n)
    r = np.random.default_rng(11)
    for c in (G.U, G.V, G.W):
        g.set(c, r.standard_normal(g.shape(c), dtype=np.float32))
    pos = r.uniform(0.0, n, (npts, 3)).astype(np.float32)
    key = (pos[:, 2].astype(np.int64) * n + pos[:, 1].astype(np.int64)) * n + pos[:, 0].astype(np.int64)
    sorted_pos = np.ascontiguousarray(pos[np.argsort(key, kind="stable")])
    dpos, dout, ddiv = dmalloc(pos.nbytes), dmalloc(pos.nbytes), dmalloc(4 * n ** 3)
    stream = g.L.fs3d_stream(g.h)
    ev = [vp(), vp()]
    for e in ev:
        assert hip.hipEventCreate(ctypes.byref(e)) == 0

    def timed(launch):
        launch()  # warm
        assert hip.hipEventRecord(ev[0], stream) == 0
        for _ in range(reps):
            launch()
        assert hip.hipEventRecord(ev[1], stream) == 0
        assert hip.hipEventSynchronize(ev[1]) == 0
        ms = ctypes.c_float()
        assert hip.hipEventElapsedTime(ctypes.byref(ms), ev[0], ev[1]) == 0
        return ms.value / reps * 1e-3

    res = {}
    for name, pts in (("random", pos), ("cell_sorted", sorted_pos)):
        assert hip.hipMemcpy(dpos, pts.ctypes.data, pts.nbytes, 1) == 0
        t = timed(lambda: g.L.fs3d_vel_interp_device(g.h, dpos, npts, dout, G.SAMPLE_VELOCITY))
        res["vel_interp_%s" % name] = {"samples_per_s": npts / t, "us_per_launch": t * 1e6}
    assert g.L.fs3d_range_errors(g.h) == 0
    t = timed(lambda: g.L.fs3d_vel_div_device(g.h, ddiv))
    div_bytes = 16 * n ** 3  # u, v, w in and the divergence out, 4 bytes each per cell
    res["vel_div"] = {"cells_per_s": n ** 3 / t, "us_per_launch": t * 1e6, "GB_per_s": div_bytes / t / 1e9,
                      "hbm_frac": div_bytes / t / 1e9 / HBM_PEAK_GBS}
    for e in ev:
        hip.hipEventDestroy(e)
    for p_ in (dpos, dout, ddiv):
        hip.hipFree(p_)
    g.close()
    res["config"] = {"workload": "%d^3 MAC grid, FP32 N(0,1) u, v, w; velInterp (velocity taps) at %d uniform "
                                 "random points, in random and in cell order; velDiv at every cell" % (n, npts)}
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=2048)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--tau", type=float, default=0.97)
    ap.add_argument("--max-iters", type=int, default=1000)
    ap.add_argument("--no-cpu", action="store_true", help="skip the cpu_baseline leg")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="oracle threads (0: the physical cores this process may run on, at most 16)")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the batched-ensemble leg")
    ap.add_argument("--ensemble-members", type=int, default=256)
    ap.add_argument("--ensemble-steps", type=int, default=5)
    ap.add_argument("--loop-steps", type=int, default=2)
    ap.add_argument("--no-pic", action="store_true", help="skip the 512^2 PIC/FLIP leg")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint leg")
    ap.add_argument("--no-grid3d", action="store_true", help="skip the 3D grid leg")
    ap.add_argument("--no-events", action="store_true",
                    help="no HIP events anywhere (for profiler runs); no roofline or per-class breakdown")
    ap.add_argument("--no-graph", action="store_true",
                    help="skip the captured-graph twins of the 2048^2 and ensemble legs (profiling runs)")
    ap.add_argument("--tune-cache", default=os.path.join(ROOT, "fluid-sim_amd", "tune_cache.txt"),
                    help="autotuner record (read, or appended to after timing the candidates)")
    ap.add_argument("--only", choices=["ensemble", "pic", "adjoint", "grid3d"],
                    help="run one leg alone and print its object (profiling runs)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: there is no multi-GPU path "
              "(each GPU would run its own independent grids)", file=sys.stderr)
        sys.exit(2)

    # read before OpenMP starts: with OMP_PROC_BIND its runtime pins this
    # thread to one core, and the affinity mask shrinks to that core
    host = host_cpus()
    # OpenMP reads these when the oracle's runtime starts (the cpu_baseline leg)
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    import fs2d
    if args.only == "ensemble":
        print(json.dumps(ensemble_leg(args.ensemble_members, args.ensemble_steps, args.warmup, args.loop_steps,
                                      args.tune_cache, not args.no_graph)))
        return
    if args.only == "pic":
        print(json.dumps(pic_leg()))
        return
    if args.only == "adjoint":
        print(json.dumps(adjoint_leg()))
        return
    if args.only == "grid3d":
        print(json.dumps(grid3d_leg()))
        return
    n = args.n
    sim, cells, dt = build_gpu_sim(n, args.tol, args.tau, args.max_iters)
    state = [sim.get(f) for f in (fs2d.F_U, fs2d.F_V, fs2d.F_DENSITY)]
    # the same workload stepped by replaying captured graphs, in lockstep
    gsim = None if args.no_graph else build_gpu_sim(n, args.tol, args.tau, args.max_iters)[0]
    twins = [sim] + ([gsim] if gsim else [])
    if gsim:
        gsim.set_graph(True)
    for s_ in twins:
        s_.step()
    tune = sim.autotune(iters=24, cache=args.tune_cache)
    if gsim:
        gsim.set_tuning(pcg_grid=tune["pcg_grid"], fused_pcg=tune["fused_pcg"])
    for _ in range(args.warmup):
        for s_ in twins:
            s_.step()
    # per-class breakdown and the dominant class, with every class timed
    kt_all, dom = {}, None
    if not args.no_events:
        sim.synchronize()
        sim.reset_kernel_times()
        sim.set_timing(True)
        for _ in range(PROFILE_STEPS):
            for s_ in twins:
                s_.step()
        kt_all = sim.kernel_times()
        sim.set_timing(False)
        dom = max((k for k in kt_all if k in BYTES_PER_CELL), key=lambda k: kt_all[k][0])
    # timed region: only the dominant class carries events
    sim.reset_kernel_times()
    if dom:
        sim.set_timing(True, only=[dom])
    iters = []
    capped = 0
    sim.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        sim.step()
        st = sim.stats()
        iters.append(st["pcg_iters"])
        capped += int(st["pcg_hit_max"])
    sim.synchronize()
    elapsed = time.perf_counter() - t0
    kt = sim.kernel_times()
    sim.set_timing(False)
    graph_out = None
    if gsim:
        gsim.synchronize()
        t0 = time.perf_counter()
        giters = []
        for _ in range(args.steps):
            gsim.step()
            giters.append(gsim.stats()["pcg_iters"])
        gsim.synchronize()
        g_elapsed = time.perf_counter() - t0
        gsim.close()
        graph_out = {"value": n * n * args.steps / g_elapsed, "ms_per_step": g_elapsed / args.steps * 1e3,
                     "pcg_iters_mean": float(np.mean(giters)),
                     "note": "the same steps replayed from captured HIP graphs (fs2d_set_graph), no events"}

    value, ms_per_step = n * n * args.steps / elapsed, elapsed / args.steps * 1e3
    no_graph = {"no_graph_value": value, "no_graph_ms_per_step": ms_per_step}
    if graph_out:
        value, ms_per_step = graph_out["value"], graph_out["ms_per_step"]
    roofline = None
    if dom:
        tot_ms, launches = kt[dom]
        if dom in EFFECTIVE:
            # fused_pcg: the forward class also runs the speculative solve of the converged iteration
            extra = EFFECTIVE[dom] + (1 if dom == "pcg_forward_solve" and tune.get("fused_pcg") else 0)
            launches = int(sum(iters)) + extra * args.steps
        avg_s = tot_ms / 1e3 / max(launches, 1)
        label = dom
        if tune.get("fused_pcg") and dom in ("pcg_backward_solve", "pcg_forward_solve"):
            label = dom + ("+spmv" if dom == "pcg_backward_solve" else "+axpy")
        bytes_per_launch = BYTES_PER_CELL[label] * n * n
        achieved = bytes_per_launch / avg_s / 1e9
        roofline = {
            "bound": "hbm", "kernel": label, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(label, n),
            "avg_launch_us": avg_s * 1e6, "launches": launches, "launches_all": kt[dom][1],
            "bytes_per_launch": bytes_per_launch,
        }
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "all-fluid %dx%d semi-Lagrangian step: advect u, v, density; MIC(0)-PCG "
                        "projection to 1e-6; velocity update" % (n, n),
            "grid": [n, n], "pcg_tol": args.tol, "tau": args.tau, "cfl": 1.0, "dt": dt,
            "pcg_iters_mean": float(np.mean(iters)), "pcg_iters_max": int(max(iters)),
            "pcg_max_iters": args.max_iters, "pcg_capped_steps": capped,
            "tuning": tune,
        },
        "graph": graph_out,
        **no_graph,
        "roofline": roofline,
        "kernel_ms_per_step": {k: v[0] / PROFILE_STEPS for k, v in kt_all.items() if v[1]},
    }
    sim.close()
    if not args.no_ensemble:
        out["ensemble"] = ensemble_leg(args.ensemble_members, args.ensemble_steps, 2, args.loop_steps,
                                       args.tune_cache, not args.no_graph)
    if not args.no_pic:
        out["pic"] = pic_leg()
    if not args.no_adjoint:
        out["adjoint"] = adjoint_leg()
    if not args.no_grid3d:
        out["grid3d"] = grid3d_leg()
    if not args.no_cpu:
        threads = args.cpu_threads
        if threads <= 0:
            threads = max(1, min(16, host["physical_cores"] or host["nproc"]))
        out["cpu_baseline"] = cpu_baseline(state, cells, n, dt, args.tol, args.tau, threads,
                                           float(np.mean(iters)), host)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


if __name__ == "__main__":
    main()
