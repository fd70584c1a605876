"""vfd_step_no_alloc against vfd_step, through the C ABI.

The allocation-free step takes the neighbour grid's cell arrays and the surface-tension curvature tables from the
library's static scratch and must give the same state as vfd_step, bit for bit: two no-alloc calls in a row, the
second on a smaller particle set (the scratch then holds the first call's larger grid), on a single scene with
viscosity and surface tension and on a batch of scenes; plus the refusals that keep it from allocating."""
import numpy as np
import pytest

from scene_util import make_gpu, small_scene
from vfd_amd import _abi

pytestmark = pytest.mark.gpu

VFD_ERROR_STATE, VFD_ERROR_NOT_SUPPORTED = 3, 5  # include/vfd.h status codes

FIELDS = ("position", "velocity", "density", "factor", "viscosity_solution")


def _say(*a):
    print(*a, flush=True)


def _same(a, b, what):
    for f in FIELDS:
        x, y = a.field(f), b.field(f)
        assert x.shape == y.shape and np.array_equal(x, y), (what, f, float(np.abs(x - y).max()))
    ca, oa, ia = a.neighbors()
    cb, ob, ib = b.neighbors()
    assert np.array_equal(ca, cb) and np.array_equal(oa, ob) and np.array_equal(ia, ib), what
    assert a.GetDebugInfo().PressureSolverIterationCount == b.GetDebugInfo().PressureSolverIterationCount, what


def _no_alloc(sim):
    return sim._lib.vfd_step_no_alloc(sim._h)


def _smaller(fluid, keep):
    x, v = fluid
    return np.ascontiguousarray(x[:keep]), np.ascontiguousarray(v[:keep])


@pytest.mark.parametrize("prepare", ["get_particles", "step"])
def test_no_alloc_step_matches_step_twice_then_smaller(prepare):
    descs, fluids, bodies = small_scene("st")
    ref = make_gpu(descs, fluids, bodies)
    na = make_gpu(descs, fluids, bodies)
    for s in (ref, na):
        s.set_frame_capture(False)
    if prepare == "step":  # buffers (and iteration graphs) made by an ordinary step on both
        ref.OnUpdate()
        na.OnUpdate()
    else:  # buffers made by a read, no step taken
        na.field("position")
    for k in range(2):
        assert ref._lib.vfd_step(ref._h) == 0
        rc = _no_alloc(na)
        assert rc == 0, _abi.load_library().vfd_last_error()
        _same(ref, na, "call %d (%d particles)" % (k + 1, len(fluids[0][0])))
        _say(prepare, "no-alloc call %d matches vfd_step" % (k + 1))
    # a smaller input: the same scene with a third of the particles removed
    keep = 2 * len(fluids[0][0]) // 3
    for s in (ref, na):
        s.set_fluid(*_smaller(fluids[0], keep))
        s.SetRigidBodies(bodies[0])
    assert ref._lib.vfd_step(ref._h) == 0
    assert _no_alloc(na) == 0, _abi.load_library().vfd_last_error()
    _same(ref, na, "smaller input (%d particles)" % keep)
    _say(prepare, "no-alloc call on %d particles matches vfd_step" % keep)
    # an ordinary step afterwards (graphs dropped during the no-alloc call are rebuilt) still agrees
    for _ in range(2):
        ref.OnUpdate()
        na.OnUpdate()
    _same(ref, na, "vfd_step after no-alloc steps")


def test_no_alloc_batch_matches_step():
    scenes = [small_scene(k, seed=3 + i) for i, k in enumerate(("dam", "visc", "st"))]
    descs = [s[0][0] for s in scenes]
    fluids = [s[1][0] for s in scenes]
    bodies = [s[2][0] for s in scenes]
    r0 = min(d.ParticleRadius for d in descs)
    for d in descs:  # one kernel table per batch
        d.ParticleRadius = r0
    ref = make_gpu(descs, fluids, bodies)
    na = make_gpu(descs, fluids, bodies)
    for b in (ref, na):
        b.set_frame_capture(False)
    ref.step()
    na.step()
    for k in range(2):
        if k == 1:  # second call: every scene smaller
            for b in (ref, na):
                for s in range(len(descs)):
                    b.scenes[s].set_fluid(*_smaller(fluids[s], len(fluids[s][0]) // 2))
                    b.scenes[s].SetRigidBodies(bodies[s])
        ref.step()
        assert na._lib.vfd_step_no_alloc(na._h) == 0, _abi.load_library().vfd_last_error()
        for s in range(len(descs)):
            _same(ref.scenes[s], na.scenes[s], "batch call %d scene %d" % (k + 1, s))
        _say("batch no-alloc call %d matches vfd_step" % (k + 1))


def test_no_alloc_refusals_leave_the_state_alone():
    descs, fluids, bodies = small_scene("visc")
    sim = make_gpu(descs, fluids, bodies)
    lib = sim._lib
    # frame capture is on by default: refused
    assert _no_alloc(sim) == VFD_ERROR_STATE and b"frame capture" in lib.vfd_last_error()
    sim.set_frame_capture(False)
    # fresh simulation: its buffers would have to be allocated
    assert _no_alloc(sim) == VFD_ERROR_STATE and b"allocated" in lib.vfd_last_error()
    # the refused calls changed nothing: the same run as a simulation that never saw them
    ref = make_gpu(descs, fluids, bodies)
    ref.set_frame_capture(False)
    ref.OnUpdate()
    sim.OnUpdate()
    _same(ref, sim, "after refused calls")
    # a grid beyond the static scratch (one particle about 40 m away): refused before the step changes the state
    x, v = fluids[0]
    x = x.copy()
    x[3] = 40.0
    for s in (ref, sim):
        s.set_fluid(x, v)
        s.SetRigidBodies(bodies[0])
        s.field("position")
    assert _no_alloc(sim) == VFD_ERROR_NOT_SUPPORTED and b"cells" in lib.vfd_last_error()
    assert np.array_equal(sim.field("position"), ref.field("position"))
    assert np.array_equal(sim.field("velocity"), ref.field("velocity"))
