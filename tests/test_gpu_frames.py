"""Frame output (DFSPHImplementation.cu:148-167, DFSPHParticleBuffer): frames are packed in original order on the
GPU and copied asynchronously into pinned host memory while the next steps run.

Each captured frame must hold exactly the {x, v, a} state of the step that produced it, in original particle order;
frames are due every FrameLength of simulated time, at most FrameCount of them; Simulate restarts from the baked
particles and captures the same frames again."""
import numpy as np
import pytest

from scene_util import make_gpu, make_oracle, small_scene

pytestmark = pytest.mark.gpu


def _frames(sim):
    return sim.GetParticleFrameBuffer()


def test_frames_hold_the_state_of_their_step():
    descs, fluids, bodies = small_scene("visc")
    descs[0].FrameLength = 0.009  # about every second step at dt = 5 ms
    sim = make_gpu(descs, fluids, bodies)
    expected = []
    for _ in range(12):
        before = sim.GetDebugInfo().FrameIndex
        sim.OnUpdate()
        d = sim.GetDebugInfo()
        if d.FrameIndex > before:
            x, v, a = sim.particles()
            expected.append((np.concatenate([x, v, a], axis=1), d.MaxVelocityMagnitude, d.TimeStepSize))
    frames = _frames(sim)
    assert 2 <= len(frames) == len(expected) < 12  # some steps end without a frame
    for fr, (xva, maxv, dt) in zip(frames, expected):
        assert fr.ParticleData.shape == xva.shape
        assert np.array_equal(fr.ParticleData, xva)
        assert fr.MaxVelocityMagnitude == np.float32(maxv) and fr.CurrentTimeStep == np.float32(dt)


def test_frame_count_cap_and_simulate_restart():
    descs, fluids, bodies = small_scene("dam")
    descs[0].FrameLength = 0.0
    descs[0].FrameCount = 5
    sim = make_gpu(descs, fluids, bodies)
    sim.Simulate()  # runs until FrameCount frames are captured (DFSPHImplementation.cu:33-61)
    first = [f.ParticleData.copy() for f in _frames(sim)]
    assert len(first) == 5
    sim.Simulate()
    second = [f.ParticleData for f in _frames(sim)]
    assert len(second) == 5 and all(np.array_equal(a, b) for a, b in zip(first, second))


def test_capture_off_keeps_frame_timing():
    descs, fluids, bodies = small_scene("dam")
    descs[0].FrameLength = 0.0
    sim = make_gpu(descs, fluids, bodies)
    sim.set_frame_capture(False)
    for _ in range(4):
        sim.OnUpdate()
    assert sim.GetDebugInfo().FrameIndex == 4 and len(_frames(sim)) == 0
    sim.set_frame_capture(True)
    sim.OnUpdate()
    assert len(_frames(sim)) == 1


def test_frames_match_the_oracle():
    """Frames against the CPU oracle stepped alongside (DFSPHImplementation.cu:148-167): a frame is due at the same
    steps (FrameTime accumulates dt until it reaches FrameLength), and each frame's {x, v, a} in original order, its
    MaxVelocityMagnitude and CurrentTimeStep equal the oracle's state after that step."""
    descs, fluids, bodies = small_scene("st")
    descs[0].FrameLength = 0.007
    descs[0].FrameCount = 6
    sim = make_gpu(descs, fluids, bodies)
    orc = make_oracle(descs, fluids, bodies)
    r = descs[0].ParticleRadius
    expected, steps = [], []
    for k in range(16):
        before = orc.debug_info().FrameIndex
        sim.OnUpdate()
        orc.step()
        d = orc.debug_info()
        assert sim.GetDebugInfo().FrameIndex == d.FrameIndex, k
        if d.FrameIndex > before:
            xva = np.concatenate([orc.field("position"), orc.field("velocity"), orc.field("acceleration")], axis=1)
            expected.append((xva, d.MaxVelocityMagnitude, d.TimeStepSize))
            steps.append(k)
    frames = sim.GetParticleFrameBuffer()
    print("frames due after steps", steps, flush=True)
    assert len(frames) == len(expected) == 6 and steps[-1] < 15  # the FrameCount cap stops capture
    for fr, (xva, maxv, dt) in zip(frames, expected):
        d = np.abs(fr.ParticleData.astype(np.float64) - xva)
        assert d[:, :3].max() <= 1e-4 * r
        assert d[:, 3:].max() <= 1e-5 * max(np.abs(xva[:, 3:]).max(), 1.0)
        assert fr.MaxVelocityMagnitude == pytest.approx(maxv, rel=1e-6)
        assert fr.CurrentTimeStep == pytest.approx(dt, rel=1e-6)
