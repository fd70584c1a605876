#!/usr/bin/env python3
"""Benchmark: DFSPH particle-updates/s on one MI355X (2M-particle dam break with a sphere obstacle).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config dam_break_2m] [--no-cpu-baseline]

Prints one JSON line. A step is one full DFSPH step (neighbour search, boundary, density/factor, divergence solve,
viscosity solve, time step, pressure solve, integration) of every particle; particles stay resident in HBM and frame
capture is off. `value` = particles * K / seconds over K timed steps (synchronised on both sides). `roofline` is
computed for the kernel with the largest share of GPU time, from its algorithmic bytes per launch (bytes_per_particle
below, with the measured mean neighbour count) and its mean duration measured with HIP events on the simulation's
stream during the timed steps. `cpu_baseline` times the CPU oracle (FP32, OpenMP over the host cores) on the same
scene for a few steps.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, vendor peak


def bytes_per_particle(kernel, nbar, bodies):
    """Compulsory HBM bytes per particle per launch (12 B per vec3, 4 B per scalar, 4 B per list entry + 8 B per
    list header, 16 B per boundary body), following the per-kernel data each kernel must read and write."""
    L = 4.0 * nbar + 8.0
    B = 16.0 * bodies
    table = {
        "build_lists": 24 + 8 + 4 * nbar,
        "boundary": (12 + 16) * bodies,
        "density_factor": 12 + L + B + 4 + 4,
        "div_init": 12 + 12 + L + B + 8 + 4 + 4,
        "div_accel": 12 + 4 + L + B + 12,
        "div_iter": 12 + 12 + L + B + 4 + 4 + 8 + 4 + 4,
        "div_final": 12 + 4 + L + B + 24 + 8 + 12,
        "pres_init": 12 + 12 + 4 + L + B + 8 + 4 + 4,
        "pres_accel": 12 + 4 + L + B + 12,
        "pres_iter": 12 + 12 + L + B + 4 + 4 + 8 + 4 + 4,
        "pres_final": 12 + 4 + L + B + 24 + 12 + 12,
        "visc_precond": 12 + 4 + L + B + 36 + 24 + 12,
        "cg_begin": 12 + 4 + 12 + L + B + 12 + 36 + 24,
        "cg_matvec": 12 + 4 + 12 + L + B + 12,
        "cg_update": 12 * 5 + 36,
        "cg_direction": 36,
        "visc_apply": 60,
        "time_step": 24,
        "velocity": 36,
        "minmax": 12,
    }
    return table.get(kernel)


def cpu_baseline(cfg, steps, snap):
    """The CPU oracle (FP32 mirror) advanced `steps` steps from the GPU's state `snap`, so both sides do the same
    solver iterations; returns its throughput and its positions after the first step."""
    from oracle.oracle import OracleSim, build_oracle
    from vfd_amd import scenes
    if not os.path.exists(os.path.join(ROOT, "oracle", "liboracle.so")):
        build_oracle()
    descs, fluids, bodies = scenes.CONFIGS[cfg]()
    o = OracleSim(descs[0])
    o.set_fluid(snap["x"], snap["v"])
    o.set_field("velocity_difference", snap["dv"])
    o.set_field("delta_final_curvature", snap["dk"])
    o.set_time_step(snap["dt"])
    for b in bodies[0]:
        o.add_boundary(b.to_abi())
    t0 = time.perf_counter()
    its = []
    x1 = None
    for k in range(steps):
        o.step()
        if k == 0:
            x1 = o.field("position")
        d = o.debug_info()
        its.append((d.DivergenceSolverIterationCount, d.PressureSolverIterationCount, d.ViscositySolverIterationCount))
    dt = time.perf_counter() - t0
    n = len(snap["x"])
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": n * steps / dt, "unit": "particle-updates/s", "cores": cores, "kind": "port",
            "sample": "%s, %d particles, %d step(s) from the GPU state after the warm-up steps (setup excluded); "
                      "solver iterations div/pres/visc per step %s" % (cfg, n, steps, its),
            "seconds": dt, "x1": x1, "iterations": its}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", default="dam_break_2m")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=1)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the DFSPH step runs on a single MI355X (multi-GPU sharding is a design, "
              "not a measured path); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    from vfd_amd import DFSPHSimulation, scenes

    descs, fluids, bodies = scenes.CONFIGS[args.config]()
    sim = DFSPHSimulation(descs[0])
    sim.set_fluid(*fluids[0])
    sim.SetRigidBodies(bodies[0])
    sim.set_frame_capture(False)
    n = sim.GetParticleCount()
    print("[bench] %s: %d particles, warmup %d steps" % (args.config, n, args.warmup), file=sys.stderr, flush=True)
    sim.advance(args.warmup)
    # state after warm-up (for the CPU baseline) and one untimed check step
    snap = {"x": sim.field("position"), "v": sim.field("velocity"), "dv": sim.field("velocity_difference"),
            "dk": sim.field("delta_final_curvature"), "dt": sim.GetDebugInfo().TimeStepSize}
    sim.OnUpdate()
    d = sim.GetDebugInfo()
    check_its = (d.DivergenceSolverIterationCount, d.PressureSolverIterationCount, d.ViscositySolverIterationCount)
    x1_gpu = sim.field("position")

    sim.set_profiling(True)
    its = []
    nb = []
    t0 = time.perf_counter()
    for k in range(args.steps):
        sim.OnUpdate()
        d = sim.GetDebugInfo()
        its.append((d.DivergenceSolverIterationCount, d.PressureSolverIterationCount, d.ViscositySolverIterationCount))
        nb.append(d.TotalNeighborCount / max(n, 1))
    elapsed = time.perf_counter() - t0
    prof = sim.profile()
    sim.set_profiling(False)
    # the same steps once more without event instrumentation, for the headline number
    t1 = time.perf_counter()
    sim.advance(args.steps)
    elapsed_plain = time.perf_counter() - t1
    print("[bench] timed %d steps: %.3f s (profiled %.3f s)" % (args.steps, elapsed_plain, elapsed), file=sys.stderr,
          flush=True)

    nbar = float(np.mean(nb))
    nbody = len(bodies[0])
    kernels = {}
    for name, (launches, ms) in prof.items():
        bpp = bytes_per_particle(name, nbar, nbody)
        row = {"launches": launches, "total_ms": round(ms, 4), "mean_us": round(1000 * ms / max(launches, 1), 2)}
        if bpp is not None and launches:
            gbs = bpp * n / (ms / launches * 1e-3) / 1e9
            row.update(bytes_per_launch=int(bpp * n), achieved_GBs=round(gbs, 1))
        kernels[name] = row
    single = {k: v for k, v in kernels.items() if "bytes_per_launch" in v}
    dom = max(single, key=lambda k: single[k]["total_ms"])
    achieved = single[dom]["achieved_GBs"]
    total_gpu_ms = sum(v["total_ms"] for v in kernels.values())
    value = n * args.steps / elapsed_plain
    it = np.array(its, dtype=float).mean(0)
    out = {
        "metric": "DFSPH particle-updates/s at 2M particles",
        "value": value,
        "unit": "particle-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1000 * elapsed_plain / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (FP64 solver reductions)",
        "data": "synthetic: seeded cubic lattice, analytic SDF volume maps",
        "config": {"workload": args.config, "particles": n, "particle_radius": descs[0].ParticleRadius,
                   "boundary_bodies": nbody, "mean_neighbors": round(nbar, 2),
                   "mean_iterations": {"divergence": round(it[0], 2), "pressure": round(it[1], 2),
                                       "viscosity_cg": round(it[2], 2)}},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None},
        "kernels": kernels,
        "gpu_kernel_ms_per_step": round(total_gpu_ms / args.steps, 3),
    }
    if not args.no_cpu_baseline:
        print("[bench] cpu baseline: oracle, %d steps" % args.cpu_steps, file=sys.stderr, flush=True)
        cb = cpu_baseline(args.config, args.cpu_steps, snap)
        out["cpu_baseline"] = {k: cb[k] for k in ("value", "unit", "cores", "kind", "sample")}
        r = descs[0].ParticleRadius
        out["parity_full_size"] = {
            "check": "one step from the same state on GPU and CPU oracle",
            "max_position_difference_in_particle_radii": float(np.abs(x1_gpu - cb["x1"]).max() / r),
            "iterations_gpu": check_its, "iterations_cpu": cb["iterations"][0]}
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
