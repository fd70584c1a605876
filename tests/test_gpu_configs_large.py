"""The two large benchmark configurations against the CPU oracle (FP32 mirror), through the C ABI.

- config 3, the 2M-particle dam break with a sphere obstacle: the first two full-size steps from the initial state,
  with the north_star bars (neighbour sets bit-exact, density, factor and viscosity solution within 1e-5 relative,
  positions within 1e-4 r) and equal solver iteration counts;
- config 4, 64 uneven scenes advanced as one batch: every scene of the batch against the oracle running that scene
  alone (DFSPHImplementation.cu:63-170 per scene), three steps, the same bars and iteration counts per scene.
"""
import time

import numpy as np
import pytest

from conftest import rel_inf
from scene_util import make_gpu, make_oracle
from vfd_amd import scenes

pytestmark = pytest.mark.gpu

REL = 1e-5


def _iterations(d):
    return (d.DivergenceSolverIterationCount, d.PressureSolverIterationCount, d.ViscositySolverIterationCount)


def _say(*a):
    print(*a, flush=True)


def test_config3_dam_break_2m_two_steps():
    t0 = time.perf_counter()
    scene = scenes.dam_break_2m()
    r = scene[0][0].ParticleRadius
    gpu = make_gpu(*scene)
    gpu.set_frame_capture(False)
    _say("config 3: GPU set up (%d particles), %.1f s" % (gpu.GetParticleCount(), time.perf_counter() - t0))
    orc = make_oracle(*scene)
    _say("config 3: oracle set up, %.1f s" % (time.perf_counter() - t0))
    for k in range(2):
        gpu.OnUpdate()
        orc.step()
        gi, oi = _iterations(gpu.GetDebugInfo()), _iterations(orc.debug_info())
        _say("config 3 step %d: iterations gpu %s oracle %s, %.1f s" % (k + 1, gi, oi, time.perf_counter() - t0))
        assert gi == oi, k
        if k == 0:
            gc, go, gids = gpu.neighbors()
            oc, oo, oids = orc.neighbors()
            assert np.array_equal(gc, oc) and np.array_equal(go, oo) and np.array_equal(gids, oids)
            _say("config 3 step 1: %d neighbour entries identical (mean %.2f per particle)" % (len(oids), oc.mean()))
            del gids, oids
        e = {f: rel_inf(gpu.field(f), orc.field(f)) for f in ("density", "factor", "viscosity_solution")}
        dx = float(np.abs(gpu.field("position") - orc.field("position")).max()) / r
        _say("config 3 step %d:" % (k + 1), {f: "%.2e" % v for f, v in e.items()}, "position %.2e r" % dx)
        assert dx <= 1e-4
        assert max(e.values()) <= REL


def test_config4_every_batched_scene_matches_the_oracle_alone():
    t0 = time.perf_counter()
    descs, fluids, bodies = scenes.batched_uneven()
    batch = make_gpu(descs, fluids, bodies)
    batch.set_frame_capture(False)
    steps = 3
    its = []
    for _ in range(steps):
        batch.step()
        its.append([_iterations(batch.scenes[s].GetDebugInfo()) for s in range(len(descs))])
    _say("config 4: %d scenes, %d particles, %d batched steps, %.1f s" % (
        len(descs), sum(len(f[0]) for f in fluids), steps, time.perf_counter() - t0))
    worst = {"position": 0.0, "density": 0.0, "factor": 0.0, "viscosity_solution": 0.0}
    for s in range(len(descs)):
        orc = make_oracle(descs, fluids, bodies, scene=s)
        oits = []
        for _ in range(steps):
            orc.step()
            oits.append(_iterations(orc.debug_info()))
        # iteration counts of the last step per scene, and of every step through the recorded batch counts
        assert [its[k][s] for k in range(steps)] == oits, s
        sc = batch.scenes[s]
        r = descs[s].ParticleRadius
        dx = float(np.abs(sc.field("position") - orc.field("position")).max()) / r
        worst["position"] = max(worst["position"], dx)
        for f in ("density", "factor", "viscosity_solution"):
            worst[f] = max(worst[f], rel_inf(sc.field(f), orc.field(f)))
        assert dx <= 1e-4, s
        assert max(worst[f] for f in ("density", "factor", "viscosity_solution")) <= REL, s
        if s % 16 == 15:
            _say("config 4: scenes 0-%d match, %.1f s" % (s, time.perf_counter() - t0))
        del orc
    _say("config 4, worst over 64 scenes after %d steps:" % steps, {f: "%.2e" % v for f, v in worst.items()})
