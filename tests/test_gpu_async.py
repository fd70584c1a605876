"""Simulate on a worker thread while another thread polls, as VFD's editor does: DFSPHSimulation::Simulate pushes the
frame loop onto a ThreadPool worker (DFSPHSimulator.cpp:156-166) and the UI thread reads GetSimulationState, the
debug info and the particle count every frame (ComponentPanel.cpp:446-705), then the frames once Ready
(Scene.cpp:323-399)."""
import threading
import time

import numpy as np
import pytest

from scene_util import make_gpu, small_scene
from vfd_amd import VfdError, _abi

pytestmark = pytest.mark.gpu


def _scene(frames=120):
    descs, fluids, bodies = small_scene("visc")
    descs[0].FrameLength = 0.0
    descs[0].FrameCount = frames
    return descs, fluids, bodies


def _poll_until_done(sim, is_done):
    seen, progress = set(), []
    while True:
        st = sim.GetSimulationState()
        seen.add(st)
        d = sim.GetDebugInfo()
        progress.append(d.FrameIndex)
        assert sim.GetParticleCount() > 0 and sim.GetDescription().FrameCount > 0
        if is_done():
            return seen, progress
        time.sleep(0.0005)


def test_simulate_async_polled_from_the_calling_thread():
    scene = _scene()
    ref = make_gpu(*scene)
    assert ref.GetSimulationState() == _abi.STATE_NONE
    ref.Simulate()
    assert ref.GetSimulationState() == _abi.STATE_READY
    want = [f.ParticleData.copy() for f in ref.GetParticleFrameBuffer()]

    sim = make_gpu(*scene)
    sim.Simulate(asynchronous=True)
    # a driving call while the worker runs is refused, not raced
    with pytest.raises(VfdError) as e:
        sim.OnUpdate()
    assert e.value.code == 3
    seen, progress = _poll_until_done(sim, lambda: sim.GetSimulationState() != _abi.STATE_SIMULATING)
    sim.Wait()
    assert _abi.STATE_SIMULATING in seen
    assert sim.GetSimulationState() == _abi.STATE_READY
    assert all(b >= a for a, b in zip(progress, progress[1:]))  # FrameIndex only grows while it is polled
    assert len(set(progress)) > 3  # the poll saw the run progress
    got = [f.ParticleData for f in sim.GetParticleFrameBuffer()]
    assert len(got) == len(want) == 120
    assert all(np.array_equal(a, b) for a, b in zip(got, want))
    # setting new fluid makes the state None again (SetFluidObjects, DFSPHImplementation.cu:177)
    sim.set_fluid(*scene[1][0])
    assert sim.GetSimulationState() == _abi.STATE_NONE


def test_synchronous_simulate_on_a_second_thread_with_polling():
    """The host application's own worker thread calls the synchronous Simulate while the main thread polls the
    state, debug info and frames captured so far (ctypes releases the interpreter lock for the call)."""
    scene = _scene(frames=80)
    sim = make_gpu(*scene)
    box = {}

    def run():
        try:
            sim.Simulate()
        except Exception as exc:  # reported below
            box["error"] = exc

    t = threading.Thread(target=run)
    t.start()
    counts = []
    seen, progress = _poll_until_done(sim, lambda: not t.is_alive())
    t.join()
    assert "error" not in box, box
    assert sim.GetSimulationState() == _abi.STATE_READY
    assert all(b >= a for a, b in zip(progress, progress[1:]))
    frames = sim.GetParticleFrameBuffer()
    assert len(frames) == 80
    # a second run reproduces the frames; polling the frame list while it grows stays consistent
    first = [f.ParticleData.copy() for f in frames]
    sim.Simulate(asynchronous=True)
    while sim.GetSimulationState() == _abi.STATE_SIMULATING:
        n = len(sim.GetParticleFrameBuffer())
        counts.append(n)
    sim.Wait()
    assert all(b >= a for a, b in zip(counts, counts[1:]))
    assert all(np.array_equal(a, f.ParticleData) for a, f in zip(first, sim.GetParticleFrameBuffer()))


def test_failed_async_run_reports_through_wait():
    """A run that fails returns its status and message from Wait and leaves the state None. Here one particle sits
    5 km away in every axis, so the first step's neighbour grid would need more than 2^30 cells and the search
    refuses it on the host (VFD_ERROR_NOT_SUPPORTED) before any kernel reads the grid."""
    descs, fluids, bodies = _scene(frames=10)
    x, v = fluids[0]
    x = x.copy()
    x[5] = 5000.0
    sim = make_gpu(descs, [(x, v)], bodies)
    sim.Simulate(asynchronous=True)
    with pytest.raises(VfdError) as e:
        sim.Wait()
    assert e.value.code == 5 and "too sparse" in str(e.value)
    assert sim.GetSimulationState() == _abi.STATE_NONE
    # the simulation stays usable: a far but representable particle steps normally
    x[5] = 20.0
    sim.set_fluid(x, v)
    sim.OnUpdate()
    assert np.isfinite(sim.field("position")).all()
