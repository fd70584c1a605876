"""Edge cases of the step (the failure modes VFD's own code paths handle or expose), HIP product against the oracle.

- an empty scene: OnUpdate returns at once (DFSPHImplementation.cu:65-67), alone and inside a batch;
- a zero right-hand side of the viscosity solve (all velocities zero, no gravity: :654-659);
- solvers stopped by their iteration caps before converging (Max below what the tolerance needs);
- a particle whose velocity is NaN: the build's per-step check reports VFD_ERROR_NONFINITE (the reference lets NaN
  propagate).
Neighbour overflow beyond 70 and particles pushed out of an SDF are covered in test_gpu_parity.py."""
import numpy as np
import pytest

from conftest import rel_inf
from scene_util import make_gpu, make_oracle, small_scene
from vfd_amd import VfdError, scenes
from vfd_amd.simulation import DFSPHSimulation, DFSPHSimulationDescription, RigidBodyDescription

pytestmark = pytest.mark.gpu


def test_empty_scene_steps():
    d = DFSPHSimulationDescription(ParticleRadius=0.02)
    sim = DFSPHSimulation(d)
    sim.set_fluid(np.zeros((0, 3), np.float32))
    sim.OnUpdate()
    sim.OnUpdate()
    assert sim.GetParticleCount() == 0
    assert sim.particles()[0].shape == (0, 3)


def test_empty_scene_inside_a_batch():
    from vfd_amd import DFSPHBatch
    descs, fluids, bodies = scenes.dam_break_8k()
    batch = DFSPHBatch([descs[0], descs[0], descs[0]])
    batch.scenes[0].set_fluid(*fluids[0])
    batch.scenes[0].SetRigidBodies(bodies[0])
    batch.scenes[1].set_fluid(np.zeros((0, 3), np.float32))
    batch.scenes[2].set_fluid(fluids[0][0][:300], fluids[0][1][:300])
    batch.scenes[2].SetRigidBodies(bodies[0])
    single = make_gpu(descs, fluids, bodies)
    small = DFSPHSimulation(descs[0])
    small.set_fluid(fluids[0][0][:300], fluids[0][1][:300])
    small.SetRigidBodies(bodies[0])
    for _ in range(3):
        batch.step()
        single.OnUpdate()
        small.OnUpdate()
    assert batch.scenes[1].GetParticleCount() == 0
    assert np.array_equal(batch.scenes[0].field("position"), single.field("position"))
    assert np.array_equal(batch.scenes[2].field("position"), small.field("position"))


def test_zero_viscosity_right_hand_side():
    r = 0.02
    rng = np.random.default_rng(3)
    x = scenes.lattice_block((-0.2, -0.2, -0.2), (10, 10, 10), 2 * r, 0.01 * r, rng)
    d = DFSPHSimulationDescription(ParticleRadius=r, Gravity=(0.0, 0.0, 0.0), Viscosity=0.05, BoundaryViscosity=0.05,
                                   EnableSurfaceTensionSolver=0, MinDivergenceSolverIterations=2,
                                   MinPressureSolverIterations=2)
    box = RigidBodyDescription("box", half_extents=(1, 1, 1), inverted=True, resolution=(12, 12, 12))
    scene = ([d], [(x, np.zeros_like(x))], [[box]])
    gpu, orc = make_gpu(*scene), make_oracle(*scene)
    gpu.OnUpdate()
    orc.step()
    assert gpu.GetDebugInfo().ViscositySolverIterationCount == orc.debug_info().ViscositySolverIterationCount
    e = {k: rel_inf(gpu.field(k), orc.field(k)) for k in ("viscosity_solution", "velocity", "position")}
    print("zero rhs", e, flush=True)
    assert np.array_equal(gpu.field("position"), orc.field("position").astype(np.float32))
    assert np.abs(gpu.field("velocity")).max() == np.abs(orc.field("velocity")).max()


@pytest.mark.parametrize("caps", [(1, 1, 1), (3, 2, 2)])
def test_solvers_stopped_by_iteration_caps(caps):
    div, pres, visc = caps
    descs, fluids, bodies = small_scene("visc")
    d = descs[0]
    d.MaxDivergenceSolverError, d.MaxPressureSolverError, d.MaxViscositySolverError = 1e-6, 1e-6, 1e-6
    d.MinDivergenceSolverIterations = d.MaxDivergenceSolverIterations = div
    d.MinPressureSolverIterations = d.MaxPressureSolverIterations = pres
    d.MaxViscositySolverIterations = visc
    gpu, orc = make_gpu(descs, fluids, bodies), make_oracle(descs, fluids, bodies)
    for _ in range(2):
        gpu.OnUpdate()
        orc.step()
        g, o = gpu.GetDebugInfo(), orc.debug_info()
        its = (g.DivergenceSolverIterationCount, g.PressureSolverIterationCount, g.ViscositySolverIterationCount)
        assert its == (o.DivergenceSolverIterationCount, o.PressureSolverIterationCount,
                       o.ViscositySolverIterationCount)
        assert its[0] == div and its[1] == pres and its[2] <= visc
    r = d.ParticleRadius
    err = np.abs(gpu.field("position") - orc.field("position")).max() / r
    print("caps", caps, "iterations", its, "position error / r", err, flush=True)
    assert err < 1e-4


def test_nonfinite_particle_is_reported():
    scene = small_scene("dam")
    sim = make_gpu(*scene)
    sim.OnUpdate()
    v = sim.field("velocity")
    v[17] = np.nan
    sim.set_field("velocity", v)
    with pytest.raises(VfdError, match="NONFINITE") as e:
        sim.OnUpdate()
    assert e.value.code == 6
    assert not np.isfinite(sim.field("velocity")).all()
