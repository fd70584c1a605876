#!/usr/bin/env python3
"""Benchmark: DFSPH particle-updates/s on one MI355X (2M-particle dam break with a sphere obstacle).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config dam_break_2m] [--no-cpu-baseline]
    python bench.py --copy-probe                      # plain 512 MiB device copies, for counter calibration
    python bench.py --pmc-summary FETCH WRITE COPY_FETCH COPY_WRITE [TRACE]   # rocprofv3 output dirs -> traffic table
    python bench.py --trace-summary DIR               # rocprofv3 --kernel-trace output -> per-kernel means
    python bench.py --counter-summary LDS_DIR VALU_DIR  # rocprofv3 --pmc LDS / instruction counters -> table

Prints one JSON line. A step is one full DFSPH step (neighbour search, boundary, density/factor, divergence solve,
viscosity solve, time step, pressure solve, integration) of every particle; particles stay resident in HBM and frame
capture is off. `value` = particles * K / seconds over K timed steps (synchronised on both sides). `roofline` is
computed for the kernel with the largest share of GPU time, from its algorithmic bytes per launch (bytes_per_particle
below, with the measured mean neighbour count) and its mean duration measured with HIP events on the simulation's
stream during a replay of the timed steps from a snapshot of the state (the replay is deterministic and is checked to
end in the identical state, so both passes do the same solver iterations). `roofline.traffic` is the HBM traffic per
launch of that kernel from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary, when one exists for it.
`cpu_baseline` times the CPU oracle (FP32, OpenMP over the host cores) from the same snapshot for as many of the
timed steps as fit in --cpu-seconds, and reports the GPU rate over exactly those steps beside it; the GPU's first
step from the snapshot is compared with the oracle's for a full-size parity check. The default K = 95 after W = 5
warm-up steps covers the 100-step job. `batched_gain` (second figure of merit) runs config C4, 64 uneven scenes, as
one batch and as one simulation per scene on the same GPU: value = batched throughput / one-scene-per-launch
throughput over the same steps. `adjoint_over_forward` (third) runs config C5, 100k particles over 60 steps with
checkpoints every 10: value = time of the backward pass (segment re-runs with recording, plus the reverse sweeps)
/ time of the forward rollout. `adjoint_2m` times the same reversal on three steps of the headline workload.
"""
import argparse
import csv
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, vendor peak
COPY_PROBE_BYTES = 512 * 1024 * 1024


def bytes_per_particle(kernel, nbar, bodies):
    """Compulsory HBM bytes per particle per launch (12 B per vec3, 4 B per scalar, 4 B per list entry + 8 B per
    list header, 16 B per boundary body), following the per-kernel data each kernel must read and write. Neighbour
    data re-read through the caches is not counted."""
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


def cpu_baseline(cfg, snap, max_steps, budget_s):
    """The CPU oracle (FP32 mirror) advanced from the GPU's state `snap` until `budget_s` seconds have passed or
    `max_steps` steps are done, so it meets the same solver iterations as the GPU's first timed steps; returns its
    throughput and its positions after the first step."""
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
    while len(its) < max_steps:
        o.step()
        if not its:
            x1 = o.field("position")
        d = o.debug_info()
        its.append((d.DivergenceSolverIterationCount, d.PressureSolverIterationCount, d.ViscositySolverIterationCount))
        if time.perf_counter() - t0 >= budget_s:
            break
    dt = time.perf_counter() - t0
    n = len(snap["x"])
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": n * len(its) / dt, "unit": "particle-updates/s", "cores": cores, "kind": "port",
            "sample": "%s, %d particles, the first %d timed step(s) from the GPU's state after the warm-up (setup "
                      "excluded); solver iterations div/pres/visc per step %s" % (cfg, n, len(its), its),
            "seconds": dt, "x1": x1, "iterations": its, "steps": len(its)}


def cpu_model():
    """The host CPU's model name as /proc/cpuinfo reports it."""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _oracle_scene(descs, fluids, bodies, s=0, fp64=False):
    from oracle.oracle import OracleSim
    o = OracleSim(descs[s], fp64=fp64)
    o.set_fluid(*fluids[s])
    for b in bodies[s]:
        o.add_boundary(b.to_abi())
    return o


def _iters(d):
    return (d.DivergenceSolverIterationCount, d.PressureSolverIterationCount, d.ViscositySolverIterationCount)


def cpu_baselines_other(c5_steps=60):
    """The CPU oracle (FP32 mirror, OpenMP over the host cores) on the other benchmark configurations, each beside the
    GPU over exactly the same steps from the same initial state (setup excluded on both):
      C1 dam_break_8k: all 200 steps;  C2 viscous_coil_250k: the first 5 steps;
      C4 batched_uneven_64: the first 3 steps of every scene, stepped one scene after another on the CPU and as one
         batch on the GPU;
      C5 falling_viscous_block_100k: the first `c5_steps` steps of its 60-step forward, plus the FP64 oracle
         adjoint of a 2,028-particle version (3 recorded steps, the oracle's reference gradient), timed forward and
         backward.
    Solver iteration counts are reported per config so that equal work can be checked."""
    import numpy as np
    from oracle.oracle import build_oracle
    from vfd_amd import DFSPHBatch, DFSPHSimulation, scenes
    if not os.path.exists(os.path.join(ROOT, "oracle", "liboracle.so")):
        build_oracle()
    out = {"cores": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)), "cpu_model": cpu_model(),
           "kind": "port"}

    def gpu_single(descs, fluids, bodies, steps):
        sim = DFSPHSimulation(descs[0])
        sim.set_fluid(*fluids[0])
        sim.SetRigidBodies(bodies[0])
        sim.set_frame_capture(False)
        sim.field("position")  # state allocated before the clock starts, as the oracle's setup is excluded
        its = []
        t0 = time.perf_counter()
        for _ in range(steps):
            sim.OnUpdate()
            its.append(_iters(sim.GetDebugInfo()))
        return time.perf_counter() - t0, its

    for name, steps in (("dam_break_8k", 200), ("viscous_coil_250k", 5)):
        descs, fluids, bodies = scenes.CONFIGS[name]()
        n = len(fluids[0][0])
        o = _oracle_scene(descs, fluids, bodies)
        its = []
        t0 = time.perf_counter()
        for _ in range(steps):
            o.step()
            its.append(_iters(o.debug_info()))
        cpu_s = time.perf_counter() - t0
        del o
        gpu_s, gits = gpu_single(descs, fluids, bodies, steps)
        out[name] = {"particles": n, "steps": steps, "cpu_value": n * steps / cpu_s, "gpu_value": n * steps / gpu_s,
                     "unit": "particle-updates/s", "same_iterations": its == gits,
                     "mean_iterations_cpu": [round(float(v), 2) for v in np.mean(its, 0)]}
        print("[bench] cpu baseline %s: %.3g (GPU %.3g) updates/s" % (name, out[name]["cpu_value"],
                                                                       out[name]["gpu_value"]),
              file=sys.stderr, flush=True)

    descs, fluids, bodies = scenes.batched_uneven()
    steps = 3
    total = sum(len(f[0]) for f in fluids)
    cpu_s, its = 0.0, []
    for s in range(len(descs)):
        o = _oracle_scene(descs, fluids, bodies, s)
        t0 = time.perf_counter()
        for _ in range(steps):
            o.step()
        cpu_s += time.perf_counter() - t0
        its.append(_iters(o.debug_info()))
        del o
    b = DFSPHBatch(descs)
    for s in range(len(descs)):
        b.scenes[s].set_fluid(*fluids[s])
        b.scenes[s].SetRigidBodies(bodies[s])
    b.set_frame_capture(False)
    b.scenes[0].field("position")
    t0 = time.perf_counter()
    for _ in range(steps):
        b.step()
    gpu_s = time.perf_counter() - t0
    gits = [_iters(b.scenes[s].GetDebugInfo()) for s in range(len(descs))]
    out["batched_uneven_64"] = {"particles": total, "scenes": len(descs), "steps": steps,
                                "cpu_value": total * steps / cpu_s, "gpu_value": total * steps / gpu_s,
                                "unit": "particle-updates/s", "same_iterations": its == gits,
                                "note": "CPU: scenes one after another; GPU: one batch"}
    del b
    print("[bench] cpu baseline batched_uneven_64: %.3g (GPU %.3g) updates/s" % (
        out["batched_uneven_64"]["cpu_value"], out["batched_uneven_64"]["gpu_value"]), file=sys.stderr, flush=True)

    descs, fluids, bodies = scenes.falling_viscous_block()
    n = len(fluids[0][0])
    o = _oracle_scene(descs, fluids, bodies)
    t0 = time.perf_counter()
    for _ in range(c5_steps):
        o.step()
    cpu_s = time.perf_counter() - t0
    del o
    gpu_s, _ = gpu_single(descs, fluids, bodies, c5_steps)
    d2, f2, b2 = scenes.falling_viscous_block(n_side=(13, 12, 13), differentiable=True)
    o = _oracle_scene(d2, f2, b2, fp64=True)
    o.tape(True)
    t0 = time.perf_counter()
    for _ in range(3):
        o.step()
    fwd_s = time.perf_counter() - t0
    rng = np.random.default_rng(0)
    t0 = time.perf_counter()
    o.adjoint(rng.normal(size=(len(f2[0][0]), 3)), rng.normal(size=(len(f2[0][0]), 3)))
    adj_s = time.perf_counter() - t0
    del o
    out["falling_viscous_block_100k"] = {
        "particles": n, "steps": c5_steps, "cpu_value": n * c5_steps / cpu_s, "gpu_value": n * c5_steps / gpu_s,
        "unit": "particle-updates/s",
        "fp64_adjoint_2k": {"particles": len(f2[0][0]), "steps": 3, "forward_s": round(fwd_s, 4),
                            "adjoint_s": round(adj_s, 4), "adjoint_over_forward": round(adj_s / fwd_s, 3)}}
    print("[bench] cpu baseline falling_viscous_block_100k: %.3g (GPU %.3g) updates/s" % (
        out["falling_viscous_block_100k"]["cpu_value"], out["falling_viscous_block_100k"]["gpu_value"]),
        file=sys.stderr, flush=True)
    return out


LIVE_US = 20.0  # loop-kernel dispatches shorter than this exited at once (issued after convergence)
LOOP_KERNELS = {"k_div_accel", "k_div_iter", "k_pres_accel", "k_pres_iter", "k_cg_matvec", "k_cg_update",
                "k_cg_direction"}


def _find(path, suffix):
    if os.path.isdir(path):
        return [os.path.join(dp, f) for dp, _, fs in os.walk(path) for f in fs if f.endswith(suffix)][0]
    return path


def _kname(raw):
    return re.sub(r"^(?:void )?(?:vfd::)?", "", raw).split("(")[0]


def _pmc_table(path):
    """{kernel: {counter: mean value}} from a rocprofv3 --pmc counter_collection.csv (or the directory holding it),
    over the live dispatches (>= LIVE_US) of the solver-loop kernels and every dispatch of the others."""
    acc = {}
    for r in csv.DictReader(open(_find(path, "counter_collection.csv"))):
        us = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1000.0
        k = _kname(r["Kernel_Name"])
        if us < LIVE_US and k in LOOP_KERNELS:
            continue
        d = acc.setdefault(k, {}).setdefault(r["Counter_Name"], [])
        d.append(float(r["Counter_Value"]))
    return {k: {c: sum(v) / len(v) for c, v in cs.items()} for k, cs in acc.items()}


def _trace_rows(path):
    """(kernel, grid_x, microseconds) of every dispatch in a rocprofv3 --kernel-trace output: kernel_trace.csv
    (--output-format csv) or the default SQLite *_results.db (its `kernels` view)."""
    try:
        for r in csv.DictReader(open(_find(path, "kernel_trace.csv"))):
            yield (_kname(r["Kernel_Name"]), int(r.get("Grid_Size_X", r.get("Grid_Size", 0)) or 0),
                   (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1000.0)
    except (IndexError, FileNotFoundError):
        import sqlite3
        db = sqlite3.connect(_find(path, "_results.db"))
        for name, grid, start, end in db.execute("select name, grid_x, start, end from kernels"):
            yield _kname(name), int(grid), (end - start) / 1000.0


def trace_summary(path):
    """Markdown table from a rocprofv3 --kernel-trace output: every dispatch, and the live ones."""
    acc = {}
    for k, _, us in _trace_rows(path):
        acc.setdefault(k, []).append(us)
    lines = ["| kernel | dispatches | mean us (all) | live dispatches (>= %g us) | mean us (live) | total ms |" % LIVE_US,
             "|---|---:|---:|---:|---:|---:|"]
    for k, v in sorted(acc.items(), key=lambda kv: -sum(kv[1])):
        live = [x for x in v if x >= LIVE_US]
        lines.append("| %s | %d | %.2f | %d | %s | %.3f |" % (k, len(v), sum(v) / len(v), len(live),
                                                           "%.2f" % (sum(live) / len(live)) if live else "-",
                                                           sum(v) / 1000.0))
    return "\n".join(lines)


def pmc_summary(fetch_dir, write_dir, copy_fetch_dir, copy_write_dir, trace_dir=None):
    """Markdown table of HBM traffic per launch. FETCH_SIZE/WRITE_SIZE (KiB per dispatch) are scaled by what the
    same counters read for a device-to-device copy of COPY_PROBE_BYTES (`--copy-probe`). With a --kernel-trace output
    of the same command, each kernel's measured traffic over its mean duration there (live dispatches of the
    solver-loop kernels) gives the HBM bandwidth it achieves and its share of the MI355X's 8 TB/s peak."""
    dur = {}
    if trace_dir:
        acc = {}
        for k, _, us in _trace_rows(trace_dir):
            if us >= LIVE_US or k not in LOOP_KERNELS:
                acc.setdefault(k, []).append(us)
        dur = {k: sum(v) / len(v) for k, v in acc.items()}
    cf, cw = _pmc_table(copy_fetch_dir), _pmc_table(copy_write_dir)
    copy_k = [k for k in cf if "copyBuffer" in k][0]
    f_scale = COPY_PROBE_BYTES / (cf[copy_k]["FETCH_SIZE"] * 1024.0)
    w_scale = COPY_PROBE_BYTES / (cw[copy_k]["WRITE_SIZE"] * 1024.0)
    fe, wr = _pmc_table(fetch_dir), _pmc_table(write_dir)
    lines = ["calibration: %s of %d bytes reads FETCH_SIZE %.1f KiB, WRITE_SIZE %.1f KiB -> scale fetch x%.4f, "
             "write x%.4f" % (copy_k, COPY_PROBE_BYTES, cf[copy_k]["FETCH_SIZE"], cw[copy_k]["WRITE_SIZE"], f_scale,
                              w_scale), "",
             "| kernel | FETCH_SIZE KiB | WRITE_SIZE KiB | traffic bytes per launch (scaled) | mean us | HBM GB/s | "
             "of peak |", "|---|---:|---:|---:|---:|---:|---:|"]
    for k in sorted(fe):
        if k in wr and k.startswith("k_"):
            t = fe[k]["FETCH_SIZE"] * 1024.0 * f_scale + wr[k]["WRITE_SIZE"] * 1024.0 * w_scale
            us = dur.get(k)
            rate = "| %.1f | %.0f | %.3f |" % (us, t / (us * 1e3), t / (us * 1e3) / HBM_PEAK_GBS) if us else "| - | - | - |"
            lines.append("| %s | %.1f | %.1f | %d %s" % (k[2:], fe[k]["FETCH_SIZE"], wr[k]["WRITE_SIZE"], t, rate))
    return "\n".join(lines)


SHADER_ENGINES = 32  # MI355X: 8 XCDs x 4 shader engines; SQ_BUSY_CYCLES sums over them
SIMDS = 1024         # 256 CUs x 4 SIMDs; a wave64 VALU instruction occupies a 16-lane SIMD for 4 cycles


def counter_summary(lds_dir, valu_dir, kernels=("div_iter", "div_accel", "pres_iter", "pres_accel", "cg_matvec",
                                                   "density_factor", "visc_precond", "build_lists", "build_halo")):
    """Markdown table of the LDS and instruction counters of the neighbour-loop kernels (means per dispatch, summed
    over the chip as rocprofv3 reports them): LDS busy = SQ_LDS_IDX_ACTIVE per CU over the kernel's duration
    (SQ_BUSY_CYCLES / shader engines); VALU busy = 4 cycles per SQ_INSTS_VALU per SIMD over the same duration."""
    lds, valu = _pmc_table(lds_dir), _pmc_table(valu_dir)
    lines = ["| kernel | duration cycles | LDS active | bank-conflict share | LDS busy | LDS insts | LDS issue wait | "
             "VALU insts | VALU busy | VMEM reads | wait any (issue) |", "|---|" + "---:|" * 10]
    for k in kernels:
        a, b = lds.get("k_" + k), valu.get("k_" + k)
        if not a or not b:
            continue
        dur = a["SQ_BUSY_CYCLES"] / SHADER_ENGINES
        lines.append("| %s | %.0f | %.3g | %.0f%% | %.0f%% | %.3g | %.3g | %.3g | %.0f%% | %.3g | %.3g |" % (
            k, dur, a["SQ_LDS_IDX_ACTIVE"], 100.0 * a["SQ_LDS_BANK_CONFLICT"] / max(a["SQ_LDS_IDX_ACTIVE"], 1.0),
            100.0 * a["SQ_LDS_IDX_ACTIVE"] / (SIMDS / 4) / max(dur, 1.0), a["SQ_INSTS_LDS"], a["SQ_WAIT_INST_LDS"],
            b["SQ_INSTS_VALU"], 100.0 * 4.0 * b["SQ_INSTS_VALU"] / SIMDS / max(dur, 1.0), b["SQ_INSTS_VMEM_RD"],
            b["SQ_WAIT_INST_ANY"]))
    return "\n".join(lines)


CUS = 256  # MI355X compute units; the TA and TD counters' _sum adds their per-CU busy cycles


def memory_unit_summary(path, kernels=("build_lists", "div_iter", "div_accel", "cg_matvec", "build_halo")):
    """Markdown table of the vector memory units from one rocprofv3 --pmc run of TA_TA_BUSY_sum TD_TD_BUSY_sum
    SQ_BUSY_CYCLES SQ_INSTS_VMEM_RD: per CU busy share = counter / CUS over the duration SQ_BUSY_CYCLES / shader
    engines (the texture addresser TA takes the load addresses, the texture data unit TD returns the data)."""
    t = _pmc_table(path)
    lines = ["| kernel | duration cycles | TA busy | TD busy | VMEM reads |", "|---|---:|---:|---:|---:|"]
    for k in kernels:
        a = t.get("k_" + k)
        if not a:
            continue
        dur = a["SQ_BUSY_CYCLES"] / SHADER_ENGINES
        lines.append("| %s | %.0f | %.0f%% | %.0f%% | %.3g |" % (
            k, dur, 100.0 * a["TA_TA_BUSY_sum"] / CUS / max(dur, 1.0), 100.0 * a["TD_TD_BUSY_sum"] / CUS / max(dur, 1.0),
            a["SQ_INSTS_VMEM_RD"]))
    return "\n".join(lines)


def committed_traffic(config, kernel):
    """Traffic bytes per launch of `kernel` from profiles/<config>_traffic.md (written from pmc_summary)."""
    path = os.path.join(ROOT, "profiles", "%s_traffic.md" % config)
    if not os.path.exists(path):
        return None
    for line in open(path):
        m = re.match(r"\|\s*%s\s*\|[^|]*\|[^|]*\|\s*(\d+)\s*\|" % re.escape(kernel), line)
        if m:
            return int(m.group(1))
    return None


def batched_gain(steps, warmup, config="batched_uneven_64"):
    """Config C4 both ways on the same GPU: one DFSPHBatch advancing every scene with one set of launches, and one
    DFSPHSimulation per scene stepped in turn (one scene per launch). Both run the same `steps` steps after the same
    warm-up and must end in the same state, bit for bit."""
    import numpy as np
    from vfd_amd import DFSPHBatch, DFSPHSimulation, scenes
    descs, fluids, bodies = scenes.CONFIGS[config]()
    n = sum(len(f[0]) for f in fluids)
    batch = DFSPHBatch(descs)
    for s, (f, b) in enumerate(zip(fluids, bodies)):
        batch.scenes[s].set_fluid(*f)
        batch.scenes[s].SetRigidBodies(b)
    batch.set_frame_capture(False)
    batch.advance(warmup)
    t0 = time.perf_counter()
    batch.advance(steps)
    tb = time.perf_counter() - t0
    singles = []
    for d, f, b in zip(descs, fluids, bodies):
        sim = DFSPHSimulation(d)
        sim.set_fluid(*f)
        sim.SetRigidBodies(b)
        sim.set_frame_capture(False)
        sim.advance(warmup)
        singles.append(sim)
    t0 = time.perf_counter()
    for _ in range(steps):
        for sim in singles:
            sim.OnUpdate()
    ts = time.perf_counter() - t0
    same = all(np.array_equal(batch.scenes[s].field("position"), singles[s].field("position"))
               for s in range(len(singles)))
    return {"scenes": len(descs), "particles": n, "steps": steps, "warmup": warmup,
            "batched_updates_per_s": n * steps / tb, "one_scene_per_launch_updates_per_s": n * steps / ts,
            "value": ts / tb, "identical_states": bool(same)}


def adjoint_ratio(reps=3):
    """Config C5, the viscosity / initial-velocity inversion: 100k particles, 60 steps, checkpoints every 10. Times
    the forward rollout (with its checkpoints) and the backward pass (re-running each segment with recording on, then
    vfd_adjoint) through the torch.autograd.Function, after one untimed pass; reports the median of `reps`."""
    import numpy as np
    import torch
    from vfd_amd import adjoint
    ro, v0, loss = adjoint.viscosity_inversion()
    v = v0.clone().requires_grad_(True)
    mu = torch.tensor(5.0, dtype=torch.float64, requires_grad=True)
    rows = []
    for rep in range(reps + 1):
        v.grad, mu.grad = None, None
        x, _ = ro(v, mu)
        loss(x).backward()
        if rep:
            rows.append(dict(ro.timings))
    fwd = float(np.median([r["forward_s"] for r in rows]))
    bwd = float(np.median([r["backward_s"] for r in rows]))
    rev = float(np.median([r["reverse_s"] for r in rows]))
    return {"workload": "falling_viscous_block_100k", "particles": len(v0), "steps": ro.steps,
            "checkpoint_every": ro.every, "forward_ms": round(1000 * fwd, 2), "backward_ms": round(1000 * bwd, 2),
            "backward_reverse_ms": round(1000 * rev, 2), "value": round(bwd / fwd, 3),
            "dloss_dmu": float(mu.grad)}


def adjoint_headline(sim, snap, steps=3):
    """The adjoint on the headline workload: `steps` steps of the 2M dam break from the snapshot after the warm-up
    (boundary contact, push-out, boundary viscosity), timed plain, then the same steps again with recording on (the
    replay is deterministic and must end in the same state), then vfd_adjoint of a random cotangent with its
    transposed neighbour loops per neighbour group (the default at this size) and per particle, each after one
    untimed pass. reverse_over_forward = adjoint time / plain forward time of the same steps; backward_over_forward
    adds the recorded re-run, as a checkpointed rollout pays it."""
    import numpy as np
    n = sim.GetParticleCount()
    sim.restore(snap)
    t0 = time.perf_counter()
    for _ in range(steps):
        sim.OnUpdate()
    fwd = time.perf_counter() - t0
    x_plain = sim.field("position")
    sim.restore(snap)
    sim.tape(True)
    t0 = time.perf_counter()
    for _ in range(steps):
        sim.OnUpdate()
    taped = time.perf_counter() - t0
    sim.tape(False)
    same = bool(np.array_equal(x_plain, sim.field("position")))
    rng = np.random.default_rng(1)
    wx, wv = rng.normal(size=(n, 3)).astype(np.float32), rng.normal(size=(n, 3)).astype(np.float32)
    keep = sim.tuning()
    rev, grads = {}, {}
    for name, mode in (("group_loops", 1), ("particle_loops", 0)):
        sim.set_tuning(AdjointGroupLoops=mode)
        sim.adjoint(wx, wv)
        t0 = time.perf_counter()
        grads[name] = sim.adjoint(wx, wv)
        rev[name] = time.perf_counter() - t0
    sim.set_tuning(**keep)
    g, p = grads["group_loops"], grads["particle_loops"]
    return {"workload": "dam_break_2m", "particles": n, "steps": steps, "forward_ms": round(1000 * fwd, 2),
            "recorded_forward_ms": round(1000 * taped, 2), "recorded_replay_identical": same,
            "adjoint_ms": {k: round(1000 * v, 2) for k, v in rev.items()},
            "loop_forms_identical": bool(np.array_equal(g[0], p[0]) and np.array_equal(g[1], p[1])),
            "reverse_over_forward": round(rev["group_loops"] / fwd, 3),
            "backward_over_forward": round((taped + rev["group_loops"]) / fwd, 3)}


def frame_capture_cost(sim, snap, steps=10):
    """Frame output on the headline workload (a frame is due every step there): the same `steps` steps from `snap`
    with capture off and on. With capture on, each frame is packed in original order on the GPU and copied to pinned
    host memory while the next steps run; `readback_ms_per_frame` is the host copy-out of a finished frame."""
    d = sim.GetDescription()
    d.FrameCount = sim.GetDebugInfo().FrameIndex + 4 * steps  # room for a frame in every measured step
    sim.SetDescription(d)

    def run(on):
        sim.restore(snap)
        sim.set_frame_capture(on)
        d0 = sim.GetDebugInfo().FrameIndex
        t0 = time.perf_counter()
        for _ in range(steps):
            sim.OnUpdate()
        dt = time.perf_counter() - t0
        return dt, sim.GetDebugInfo().FrameIndex - d0
    off, _ = run(False)
    on, due = run(True)
    frames = sim.GetParticleFrameBuffer()  # waits for the copies, then copies every frame out
    t0 = time.perf_counter()
    frames = sim.GetParticleFrameBuffer()
    back = time.perf_counter() - t0
    sim.set_frame_capture(False)
    n = sim.GetParticleCount()
    return {"steps": steps, "frames": len(frames), "frames_due": int(due), "frame_bytes": 36 * n,
            "ms_per_step_capture_off": round(1000 * off / steps, 3), "ms_per_step_capture_on": round(1000 * on / steps, 3),
            "readback_ms_per_frame": round(1000 * back / max(len(frames), 1), 3)}


def mesh_inputs(map_res=48, sample_radius=0.005):
    """The mesh inputs at setup: the volume map of a mesh body (the reference's Torus.obj, 1,152 triangles; signed
    distance of every map node on the GPU, then the boundary-volume field) and the mesh sampler filling the example
    scene's Cone.obj fluid object at a fine particle radius. Wall-clock of the synchronous calls, after one untimed
    call of each."""
    import numpy as np
    from vfd_amd import DFSPHSimulation, DFSPHSimulationDescription, RigidBodyDescription, TriangleMesh
    from vfd_amd.simulation import sample_mesh_volume
    models = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden", "models")
    torus = TriangleMesh.load_obj(os.path.join(models, "Torus.obj"))
    cone = TriangleMesh.load_obj(os.path.join(models, "Cone.obj"))
    sim = DFSPHSimulation(DFSPHSimulationDescription(ParticleRadius=0.01))
    sim.set_fluid(np.zeros((1, 3), np.float32))
    body = RigidBodyDescription("mesh", mesh=torus, transform=np.diag([1.5, 1.5, 1.5, 1.0]),
                                resolution=(map_res,) * 3)
    ms = []
    for _ in range(3):
        t0 = time.perf_counter()
        sim.SetRigidBodies([body])
        ms.append(1000 * (time.perf_counter() - t0))
    nodes = len(sim.boundary_map(0, 0))
    # a large mesh: an icosphere of 81,920 triangles with a bumpy surface, same map resolution
    big = TriangleMesh.icosphere(6, 1.5, 0.2, 1)
    big_body = RigidBodyDescription("mesh", mesh=big, resolution=(map_res,) * 3)
    bs = []
    for _ in range(3):
        t0 = time.perf_counter()
        sim.SetRigidBodies([big_body])
        bs.append(1000 * (time.perf_counter() - t0))
    big_nodes = len(sim.boundary_map(0, 0))
    flip = np.array([[1, 0, 0, 0], [0, -1, 0, 3], [0, 0, -1, 0], [0, 0, 0, 1]], np.float32)
    ss = []
    for _ in range(3):
        t0 = time.perf_counter()
        x = sample_mesh_volume(cone, sample_radius, (40, 40, 40), False, 1, flip)
        ss.append(1000 * (time.perf_counter() - t0))
    return {"map": {"mesh": "Torus.obj", "triangles": len(torus.triangles), "resolution": map_res, "nodes": nodes,
                    "build_ms": round(float(np.median(ms[1:])), 2),
                    "node_triangle_pairs": nodes * len(torus.triangles)},
            "map_large": {"mesh": "icosphere level 6, bumps 0.2", "triangles": len(big.triangles),
                          "resolution": map_res, "nodes": big_nodes,
                          "build_ms": round(float(np.median(bs[1:])), 2),
                          "node_triangle_pairs": big_nodes * len(big.triangles)},
            "sampler": {"mesh": "Cone.obj", "radius": sample_radius, "samples": len(x),
                        "sample_ms": round(float(np.median(ss[1:])), 2)}}


def other_configs(steps=20, warmup=3):
    """Throughput of the two smaller benchmark configurations: the 8k dam break (C1) and the 250k viscous column with
    surface tension, a ground plane and walls (C2), each `steps` plain steps after `warmup`, then the same steps again
    with per-kernel HIP-event timing for the kernels that only C2 runs (surface tension) and its viscosity solve."""
    import numpy as np
    from vfd_amd import DFSPHSimulation, scenes
    out = {}
    for name in ("dam_break_8k", "viscous_coil_250k"):
        descs, fluids, bodies = scenes.CONFIGS[name]()
        sim = DFSPHSimulation(descs[0])
        sim.set_fluid(*fluids[0])
        sim.SetRigidBodies(bodies[0])
        sim.set_frame_capture(False)
        sim.advance(warmup)
        snap = sim.checkpoint()
        n = sim.GetParticleCount()
        t0 = time.perf_counter()
        for _ in range(steps):
            sim.OnUpdate()
        el = time.perf_counter() - t0
        sim.restore(snap)
        sim.set_profiling(True)
        its = []
        for _ in range(steps):
            sim.OnUpdate()
            d = sim.GetDebugInfo()
            its.append((d.DivergenceSolverIterationCount, d.PressureSolverIterationCount,
                        d.ViscositySolverIterationCount))
        prof = sim.profile()
        sim.set_profiling(False)
        keep = ("surface_tension", "visc_precond", "cg_matvec", "density_factor", "div_iter", "build_lists")
        out[name] = {"particles": n, "steps": steps, "warmup": warmup, "ms_per_step": round(1000 * el / steps, 3),
                     "particle_updates_per_s": n * steps / el,
                     "mean_iterations": [round(float(x), 2) for x in np.asarray(its, float).mean(0)],
                     "kernels_mean_us": {k: round(1000 * ms / max(l, 1), 2) for k, (l, ms) in prof.items()
                                         if k in keep}}
        del sim
    return out


def lattice_variants(steps, warmup, names=("dam_break_2m_cubic", "dam_break_2m_max")):
    """The 2M dam break seeded with the other lattices (cubic spacing 2r; the reference's MaxDensity pattern), timed
    like the headline: `warmup` steps, then `steps` plain steps; with the mean neighbour count and iterations."""
    import numpy as np
    from vfd_amd import DFSPHSimulation, scenes
    out = {}
    for name in names:
        descs, fluids, bodies = scenes.CONFIGS[name]()
        sim = DFSPHSimulation(descs[0])
        sim.set_fluid(*fluids[0])
        sim.SetRigidBodies(bodies[0])
        sim.set_frame_capture(False)
        n = sim.GetParticleCount()
        sim.advance(warmup)
        its, nb, el = [], [], 0.0
        for _ in range(steps):
            t0 = time.perf_counter()
            sim.OnUpdate()
            el += time.perf_counter() - t0
            d = sim.GetDebugInfo()
            its.append(_iters(d))
            nb.append(d.TotalNeighborCount / max(n, 1))
        out[name] = {"particles": n, "steps": steps, "warmup": warmup, "value": n * steps / el,
                     "unit": "particle-updates/s", "ms_per_step": round(1000 * el / steps, 3),
                     "mean_neighbors": round(float(np.mean(nb)), 2),
                     "mean_iterations": [round(float(v), 2) for v in np.mean(its, 0)]}
        print("[bench] %s: %.3g updates/s, mean neighbours %.2f" % (name, out[name]["value"],
                                                                      out[name]["mean_neighbors"]),
              file=sys.stderr, flush=True)
        del sim
    return out


def copy_probe():
    import torch
    n = COPY_PROBE_BYTES // 4
    x = torch.ones(n, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print("copy probe: 3 device-to-device copies of %d bytes" % COPY_PROBE_BYTES, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=95)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", default="dam_break_2m")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-cpu-other", action="store_true",
                    help="skip the CPU oracle on configs C1, C2, C4 and C5 (cpu_baselines)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--copy-probe", action="store_true")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched-gain measurement (config C4)")
    ap.add_argument("--batched-steps", type=int, default=20)
    ap.add_argument("--no-adjoint", action="store_true",
                    help="skip the adjoint/forward measurements (config C5 and the headline workload)")
    ap.add_argument("--no-c5", action="store_true", help="skip the adjoint/forward measurement of config C5 only")
    ap.add_argument("--no-frames", action="store_true", help="skip the frame-output measurement")
    ap.add_argument("--frame-steps", type=int, default=10)
    ap.add_argument("--no-mesh", action="store_true", help="skip the mesh-input (map build, sampler) measurement")
    ap.add_argument("--no-other-configs", action="store_true", help="skip the 8k and 250k configuration throughputs")
    ap.add_argument("--no-lattices", action="store_true",
                    help="skip the 2M dam break on the cubic and MaxDensity lattices (lattice_variants)")
    ap.add_argument("--mesh-only", action="store_true", help="only measure the mesh inputs")
    ap.add_argument("--pmc-summary", nargs="+", metavar="DIR",
                    help="FETCH WRITE COPY_FETCH COPY_WRITE [KERNEL_TRACE]: rocprofv3 output dirs -> traffic table")
    ap.add_argument("--trace-summary", metavar="KERNEL_TRACE_DIR")
    ap.add_argument("--memory-unit-summary", metavar="TA_TD_COUNTERS_DIR")
    ap.add_argument("--counter-summary", nargs=2, metavar=("LDS_COUNTERS_DIR", "VALU_COUNTERS_DIR"))
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the DFSPH step runs on a single MI355X (multi-GPU sharding is a design, "
              "not a measured path); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.pmc_summary:
        print(pmc_summary(*args.pmc_summary))
        return
    if args.trace_summary:
        print(trace_summary(args.trace_summary))
        return
    if args.memory_unit_summary:
        print(memory_unit_summary(args.memory_unit_summary))
        return
    if args.counter_summary:
        print(counter_summary(*args.counter_summary))
        return
    if args.copy_probe:
        copy_probe()
        return

    if args.mesh_only:
        print(json.dumps({"mesh_inputs": mesh_inputs()}), flush=True)
        return

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
    snap = sim.checkpoint()

    # headline: K plain steps (no per-kernel instrumentation); every step ends with a host synchronisation
    step_s = []
    t1 = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        sim.OnUpdate()
        step_s.append(time.perf_counter() - ts)
    elapsed_plain = time.perf_counter() - t1
    x_end = sim.field("position")
    print("[bench] timed %d steps: %.3f s" % (args.steps, elapsed_plain), file=sys.stderr, flush=True)

    # the same K steps again from the same state, with HIP-event timing around every kernel launch
    sim.restore(snap)
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
    replay_identical = bool(np.array_equal(x_end, sim.field("position")))
    print("[bench] profiled replay %.3f s, identical end state: %s" % (elapsed, replay_identical), file=sys.stderr,
          flush=True)

    # one untimed check step from the snapshot, compared with the CPU oracle's step from the same state below
    sim.restore(snap)
    sim.OnUpdate()
    d = sim.GetDebugInfo()
    check_its = (d.DivergenceSolverIterationCount, d.PressureSolverIterationCount, d.ViscositySolverIterationCount)
    x1_gpu = sim.field("position")

    fc = None
    if not args.no_frames:
        print("[bench] frame output: %d steps with capture off and on" % args.frame_steps, file=sys.stderr, flush=True)
        fc = frame_capture_cost(sim, snap, args.frame_steps)

    adj2m = None
    if not args.no_adjoint and args.config == "dam_break_2m":
        print("[bench] adjoint of the headline workload: 3 steps", file=sys.stderr, flush=True)
        adj2m = adjoint_headline(sim, snap)

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
        "data": "synthetic: seeded lattice (%s), analytic SDF volume maps" % (
            "the reference sampler's MediumDensity pattern" if args.config == "dam_break_2m" else args.config),
        "config": {"workload": args.config, "particles": n, "particle_radius": descs[0].ParticleRadius,
                   "boundary_bodies": nbody, "mean_neighbors": round(nbar, 2),
                   "mean_iterations": {"divergence": round(it[0], 2), "pressure": round(it[1], 2),
                                       "viscosity_cg": round(it[2], 2)}},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": committed_traffic(args.config, dom),
                     "bytes_per_launch": single[dom]["bytes_per_launch"], "mean_us": single[dom]["mean_us"]},
        "kernels": kernels,
        "profiled_replay_identical": replay_identical,
        "gpu_kernel_ms_per_step": round(total_gpu_ms / args.steps, 3),
        "iterations_per_step": its,
        "frame_output": fc,
        "adjoint_2m": adj2m,
    }
    if not args.no_cpu_baseline:
        print("[bench] cpu baseline: oracle, up to %.0f s" % args.cpu_seconds, file=sys.stderr, flush=True)
        cb = cpu_baseline(args.config, snap, args.steps, args.cpu_seconds)
        out["cpu_baseline"] = {k: cb[k] for k in ("value", "unit", "cores", "kind", "sample")}
        out["cpu_baseline"]["cpu_model"] = cpu_model()
        out["cpu_baseline"]["gpu_value_same_steps"] = n * cb["steps"] / sum(step_s[:cb["steps"]])
        r = descs[0].ParticleRadius
        out["parity_full_size"] = {
            "check": "one step from the same state on GPU and CPU oracle",
            "max_position_difference_in_particle_radii": float(np.abs(x1_gpu - cb["x1"]).max() / r),
            "iterations_gpu": check_its, "iterations_cpu": cb["iterations"][0]}
    else:
        out["cpu_baseline"] = None
    if not args.no_lattices and args.config == "dam_break_2m":
        print("[bench] the 2M dam break on the cubic and MaxDensity lattices", file=sys.stderr, flush=True)
        out["lattice_variants"] = lattice_variants(args.steps, args.warmup)
    if not args.no_cpu_baseline and not args.no_cpu_other:
        print("[bench] cpu baselines of the other configurations", file=sys.stderr, flush=True)
        out["cpu_baselines"] = cpu_baselines_other()
    if not args.no_batched:
        del sim
        print("[bench] batched gain: C4, %d steps" % args.batched_steps, file=sys.stderr, flush=True)
        out["batched_gain"] = batched_gain(args.batched_steps, 3)
    if not args.no_adjoint and not args.no_c5:
        print("[bench] adjoint/forward: C5", file=sys.stderr, flush=True)
        out["adjoint_over_forward"] = adjoint_ratio()
    if not args.no_mesh:
        print("[bench] mesh inputs: map build and sampler", file=sys.stderr, flush=True)
        out["mesh_inputs"] = mesh_inputs()
    if not args.no_other_configs:
        print("[bench] other configurations: 8k dam break, 250k viscous column", file=sys.stderr, flush=True)
        out["other_configs"] = other_configs()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
