"""Autotuner for the launch shapes and solver control of the DFSPH step.

Compile-time shapes (each candidate is a separate build of the library, vfd_amd.build with defines):
  VFD_LIST_CHUNK     halo-slot list entries fetched per batch in the one-record staged neighbour loops
  VFD_LIST_PREFETCH  the next batch is requested before the current one is consumed (1) or not (0)
  VFD_LIST_CHUNK2, VFD_LIST_PREFETCH2  the same for the two-record loops
  VFD_PACK_RECORDS   halo record halves read as three floats stored in 12 bytes (1), as 16-byte float4 (0), or a
                     record of two three-float halves as 16 + 8 bytes (2)
  VFD_GROUP_TILES    tiles per neighbour group (4: 1024-thread blocks, one per CU; 2: 512-thread blocks, two per
                     CU, halos of at most 2048 particles)
  VFD_HALO_MAX       largest halo a 4-tile group stages in LDS
  VFD_SEARCH_CELLS   neighbour lists built per 64-slot window with a lane per candidate of a cell (1) or one
                     particle per thread (0)
  VFD_SEARCH_UNROLL  candidate positions loaded per batch in the per-particle neighbour search
  VFD_STAGE_LISTS    neighbour lists staged in LDS and written row by row (1) or stored as found (0)
Run-time knobs (vfd_set_tuning; their defaults are compiled in as VFD_DEFAULT_*):
  SolverCheckInterval  Jacobi / CG iterations enqueued between two convergence checks

    python -m vfd_amd.autotune build                     # CPU: compile the candidates into vfd_amd/_tune/
    python -m vfd_amd.autotune run --config dam_break_2m # GPU: time them, write vfd_amd/tuning.json

The candidates are the defaults and every shape that differs from them in one axis (a star around the defaults; the
full grid would be hundreds of builds). `run` times each of them, keeps the fastest whose end state is identical, and
then sweeps the check interval on it. Every candidate runs in a process of its own (VFD_LIBRARY points it at its
build) over the same deterministic steps; its final positions must equal the default build's bit for bit, or the
candidate is rejected. tuning.json records every timing and the choice; vfd_amd.build compiles the chosen values in.
"""
import argparse
import hashlib
import json
import os
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
TUNE_DIR = os.path.join(HERE, "_tune")
TUNING_JSON = os.path.join(HERE, "tuning.json")

AXES = {
    "VFD_LIST_CHUNK": [4, 8],
    "VFD_LIST_PREFETCH": [0, 1, 2],
    "VFD_LIST_CHUNK2": [2, 4, 8],
    "VFD_LIST_PREFETCH2": [1, 2, 3],
    "VFD_SEARCH_UNROLL": [4, 8],
    "VFD_STAGE_LISTS": [0, 1],
    "VFD_PACK_RECORDS": [0, 1, 2],
    "VFD_GROUP_TILES": [2, 4],
    "VFD_HALO_MAX": [3072, 3808],
    "VFD_SEARCH_CELLS": [0, 1],
}
DEFAULTS = {"VFD_LIST_CHUNK": 4, "VFD_LIST_PREFETCH": 1, "VFD_LIST_CHUNK2": 2, "VFD_LIST_PREFETCH2": 2,
            "VFD_SEARCH_UNROLL": 8, "VFD_STAGE_LISTS": 0, "VFD_PACK_RECORDS": 2,
            "VFD_GROUP_TILES": 4, "VFD_HALO_MAX": 3808, "VFD_SEARCH_CELLS": 0}
# shapes that only make sense together: 2-tile groups take at most 2048-particle halos, two blocks per CU
LINKED = {("VFD_GROUP_TILES", 2): {"VFD_HALO_MAX": 2048, "VFD_GROUPS_PER_CU": 2}}
INTERVALS = [2, 4, 8, 16]
DEFAULT_INTERVAL = 4


def shape_defines(shape):
    """The defines of a shape, with the values LINKED to its choices (they override the shape's own)."""
    d = dict(shape)
    for (k, v), extra in LINKED.items():
        if d.get(k) == v:
            d.update(extra)
    return ["%s=%d" % kv for kv in sorted(d.items())]


def chosen_defines():
    """Compile-time values for vfd_amd.build: the tuned choice when tuning.json exists, else the defaults."""
    shape, interval = dict(DEFAULTS), DEFAULT_INTERVAL
    if os.path.exists(TUNING_JSON):
        t = json.load(open(TUNING_JSON))
        shape.update({k: int(v) for k, v in t.get("shape", {}).items() if k in AXES})
        interval = int(t.get("SolverCheckInterval", interval))
    return shape_defines(shape) + ["VFD_DEFAULT_CHECK_INTERVAL=%d" % interval]


def _name(shape):
    return "_".join("%s%d" % (k.replace("VFD_", "").lower(), v) for k, v in sorted(shape.items()))


def _candidates():
    """The defaults and every shape one axis away from them."""
    shapes = [dict(DEFAULTS)]
    for k, vals in AXES.items():
        shapes += [dict(DEFAULTS, **{k: v}) for v in vals if v != DEFAULTS[k]]
    return shapes


def _build_one(shape):
    from vfd_amd import build as vb
    out = os.path.join(TUNE_DIR, _name(shape), "libvfd_hip.so")
    os.makedirs(os.path.dirname(out), exist_ok=True)
    vb.build(out=out, defines=shape_defines(shape))
    import shutil
    shutil.rmtree(os.path.join(os.path.dirname(out), "_build"), ignore_errors=True)
    return out


def build_all(jobs=4):
    from concurrent.futures import ProcessPoolExecutor
    with ProcessPoolExecutor(jobs) as ex:
        for out in ex.map(_build_one, _candidates()):
            print("built", out, flush=True)


def trial(config, steps, warmup, interval):
    """One candidate in this process: warm up, time `steps` steps; returns ms/step and a digest of the positions."""
    sys.path.insert(0, ROOT)
    from vfd_amd import DFSPHSimulation, scenes
    descs, fluids, bodies = scenes.CONFIGS[config]()
    sim = DFSPHSimulation(descs[0])
    sim.set_fluid(*fluids[0])
    sim.SetRigidBodies(bodies[0])
    sim.set_frame_capture(False)
    sim.set_tuning(SolverCheckInterval=interval)
    sim.advance(warmup)
    t0 = time.perf_counter()
    sim.advance(steps)
    dt = time.perf_counter() - t0
    x = sim.field("position")
    return {"ms_per_step": 1000.0 * dt / steps, "digest": hashlib.sha1(x.tobytes()).hexdigest()}


def run(config, steps, warmup, timeout, margin=0.003):
    """Times the candidates; a shape or interval replaces the current choice only when it is faster by more than
    `margin` (relative): smaller differences are within the run-to-run noise of a 10-30 step timing, and chasing them
    flips the choice from run to run."""
    def measure(shape, interval):
        lib = os.path.join(TUNE_DIR, _name(shape), "libvfd_hip.so")
        if not os.path.exists(lib):
            raise RuntimeError("%s missing: run `python -m vfd_amd.autotune build` first" % lib)
        env = dict(os.environ, VFD_LIBRARY=lib)
        cmd = [sys.executable, "-m", "vfd_amd.autotune", "trial", "--config", config, "--steps", str(steps),
               "--warmup", str(warmup), "--interval", str(interval)]
        p = subprocess.run(cmd, cwd=ROOT, env=env, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                           timeout=timeout)
        if p.returncode != 0:
            raise RuntimeError("trial %s interval %d failed (%d): %s" % (_name(shape), interval, p.returncode,
                                                                         p.stderr[-2000:]))
        r = json.loads(p.stdout.strip().splitlines()[-1])
        print("[autotune] %-40s interval %2d: %.3f ms/step" % (_name(shape), interval, r["ms_per_step"]), flush=True)
        return r

    log = []
    best_shape, best_interval = dict(DEFAULTS), DEFAULT_INTERVAL
    ref = measure(best_shape, best_interval)
    best_ms = ref["ms_per_step"]
    log.append({"shape": dict(best_shape), "interval": best_interval, **ref})
    for s in _candidates()[1:]:
        r = measure(s, DEFAULT_INTERVAL)
        ok = r["digest"] == ref["digest"]
        log.append({"shape": s, "interval": DEFAULT_INTERVAL, "identical": ok, **r})
        if ok and r["ms_per_step"] < best_ms * (1.0 - margin):
            best_shape, best_ms = s, r["ms_per_step"]
    for iv in INTERVALS:
        if iv == best_interval:
            continue
        r = measure(best_shape, iv)
        ok = r["digest"] == ref["digest"]
        log.append({"shape": dict(best_shape), "interval": iv, "identical": ok, **r})
        if ok and r["ms_per_step"] < best_ms * (1.0 - margin):
            best_interval, best_ms = iv, r["ms_per_step"]
    out = {"config": config, "steps": steps, "warmup": warmup, "shape": best_shape,
           "SolverCheckInterval": best_interval, "ms_per_step": best_ms, "default_ms_per_step": ref["ms_per_step"],
           "trials": log}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("mode", choices=["build", "run", "trial"])
    ap.add_argument("--config", default="dam_break_2m")
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--margin", type=float, default=0.003, help="relative gain a candidate needs to be chosen")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--interval", type=int, default=DEFAULT_INTERVAL)
    ap.add_argument("--timeout", type=float, default=240.0)
    ap.add_argument("--out", default=TUNING_JSON)
    args = ap.parse_args()
    if args.mode == "build":
        build_all()
    elif args.mode == "trial":
        print(json.dumps(trial(args.config, args.steps, args.warmup, args.interval)), flush=True)
    else:
        res = run(args.config, args.steps, args.warmup, args.timeout, args.margin)
        with open(args.out, "w") as f:
            json.dump(res, f, indent=1)
        print(json.dumps({k: res[k] for k in ("shape", "SolverCheckInterval", "ms_per_step",
                                              "default_ms_per_step")}), flush=True)


if __name__ == "__main__":
    main()
