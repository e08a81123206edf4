"""Benchmark: cell-steps/s of the 2D step on one MI355X.

Workload (the single-grid configuration the headline metric is quoted on):
a 2048 x 2048 MAC grid, all interior cells FLUID inside one-cell SOLID walls,
semi-Lagrangian mode.  One step = semi-Lagrangian advection of u, v and a
passive density, then the MIC(0)-preconditioned CG projection to a relative
infinity-norm residual of 1e-6 (tau 0.999, the reference's value) and the
velocity update with extrapolation.  Initial velocity: i.i.d. N(0,1) from
numpy.random.default_rng(1) (u then v), Gaussian-smoothed with sigma = 8 cells
(periodic padding), projected divergence-free on the GPU; density: 16 Gaussian
blobs drawn from the same generator; dt = 1.0 * dx / max|u| (CFL 1) fixed
after the projection.  Gravity is off (a closed box of smoke-like fluid).

Prints ONE JSON line (see DESIGN.md for the byte accounting of `roofline`).
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

# Algorithmic HBM bytes per cell per launch of each kernel class: every array
# the kernel must read or write, counted once (neighbour re-reads are cache
# hits by design).  FP32 fields, uint8 flags, FP64 level set.
BYTES_PER_CELL = {
    "pcg_forward_solve": 20,    # r, P, ex, ey in; q out
    "pcg_backward_solve": 24,   # q, P, ex, ey, r in; z out
    "pcg_spmv_dot": 29,         # z, s, Adiag, Ax, Ay, cell in; s', w out
    "pcg_axpy_norm": 24,        # p, r, s, w in; p, r out
    "mic0_factor": 25,          # cell, Adiag, Ax, Ay in; P, ex, ey out
    "assemble": 41,             # cell, phi(f64), u, v in; Adiag, Ax, Ay, r, p, s out
    "sl_advect": 24,            # u, v, density in and out
    "update_velocity": 37,      # cell, phi(f64), p, u, v in; u', v', masks out
}


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


def cpu_baseline(sim, cells, n, dt, tol, tau, max_iters, threads):
    """The oracle (CPU restatement, OpenMP) on one step of the same workload
    from the GPU's starting state: a bounded sample of the timed work."""
    import fs2d
    from oracle import oracle as O
    O.set_threads(threads)
    o = O.Oracle(cells, k=0, dt=dt, dx=1.0 / n, mode=O.MODE_SL, level_set=False, gravity=(0.0, 0.0),
                 tau=tau, tol=tol, max_iters=max_iters, sl_inplace=False, density=True)
    o.set(O.F_U, sim.get(fs2d.F_U).astype(np.float64))
    o.set(O.F_V, sim.get(fs2d.F_V).astype(np.float64))
    o.set(O.F_DENS, sim.get(fs2d.F_DENSITY).astype(np.float64))
    t0 = time.perf_counter()
    o.step()
    el = time.perf_counter() - t0
    return {"value": n * n / el, "unit": "cell-steps/s", "cores": threads, "kind": "port",
            "sample": "1 step of the %dx%d workload (oracle, FP64, OpenMP; %d PCG iterations; %.1f s)"
                      % (n, n, o.stats()["last_iters"], el)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=2048)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--tau", type=float, default=0.999)
    ap.add_argument("--max-iters", type=int, default=1000)
    ap.add_argument("--no-cpu", action="store_true", help="skip the cpu_baseline leg")
    ap.add_argument("--cpu-threads", type=int, default=16)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(each GPU would run its own independent grids; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import fs2d
    n = args.n
    sim, cells, dt = build_gpu_sim(n, args.tol, args.tau, args.max_iters)
    for _ in range(args.warmup):
        sim.step()
    sim.synchronize()
    sim.reset_kernel_times()
    sim.set_timing(True)
    iters = []
    sim.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        sim.step()
        iters.append(sim.stats()["pcg_iters"])
    sim.synchronize()
    elapsed = time.perf_counter() - t0
    kt = sim.kernel_times()
    sim.set_timing(False)

    value = n * n * args.steps / elapsed
    # dominant kernel class by total time inside the timed region
    dom = max((k for k in kt if k in BYTES_PER_CELL), key=lambda k: kt[k][0])
    tot_ms, launches = kt[dom]
    avg_s = tot_ms / 1e3 / max(launches, 1)
    bytes_per_launch = BYTES_PER_CELL[dom] * n * n
    achieved = bytes_per_launch / avg_s / 1e9
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "all-fluid %dx%d semi-Lagrangian step: advect u, v, density; MIC(0)-PCG "
                        "projection to 1e-6; velocity update" % (n, n),
            "grid": [n, n], "pcg_tol": args.tol, "tau": args.tau, "cfl": 1.0, "dt": dt,
            "pcg_iters_mean": float(np.mean(iters)),
        },
        "roofline": {
            "bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": None,
            "avg_launch_us": avg_s * 1e6, "launches": launches,
            "bytes_per_launch": bytes_per_launch,
        },
        "kernel_ms_per_step": {k: v[0] / args.steps for k, v in kt.items() if v[1]},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(sim, cells, n, dt, args.tol, args.tau, args.max_iters,
                                           args.cpu_threads)
    else:
        out["cpu_baseline"] = None
    sim.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
