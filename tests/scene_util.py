"""Helpers that build the same scene on the HIP product and on the CPU oracle (tests only)."""
import numpy as np

from vfd_amd import scenes
from vfd_amd.simulation import DFSPHSimulationDescription, RigidBodyDescription


def small_scene(kind="dam", seed=7):
    """Small scenes (<= ~8k particles) that exercise every solver: kind = dam (C1 settings), visc (viscosity +
    boundary viscosity), st (surface tension + viscosity), all (everything, temporal smoothing, 2 blend passes),
    push (visc with the obstacle inside the block, so the boundary moves particles)."""
    if kind == "dam":
        return scenes.dam_break_8k(seed=seed)
    rng = np.random.default_rng(seed)
    r = 0.02
    x = scenes.lattice_block((-0.3, -0.84, -0.2), (16, 14, 16), 2 * r, 0.01 * r, rng)
    v = np.zeros_like(x)
    v[:, 1] = -0.5
    v += rng.uniform(-0.05, 0.05, size=v.shape).astype(np.float32)
    kw = dict(ParticleRadius=r, MinDivergenceSolverIterations=2, MinPressureSolverIterations=2,
              MaxDivergenceSolverError=0.1, MaxPressureSolverError=0.01, Viscosity=0.05, BoundaryViscosity=0.05,
              MaxViscositySolverError=0.1, EnableViscositySolver=1, EnableSurfaceTensionSolver=0)
    if kind in ("st", "all"):
        kw.update(EnableSurfaceTensionSolver=1, SurfaceTension=0.05, CSD=10000)
    if kind == "all":
        kw.update(TemporalSmoothing=1, SurfaceTensionSmoothPassCount=2, Viscosity=0.5, BoundaryViscosity=0.2)
    desc = DFSPHSimulationDescription(**kw)
    box = RigidBodyDescription("box", center=(0, 0, 0), half_extents=(1, 1, 1), inverted=True, resolution=(16, 16, 16))
    ball = RigidBodyDescription("sphere", center=(0.5, -0.8, 0.3), radius=0.15, padding=2 * r, resolution=(10, 10, 10))
    if kind == "push":
        # the obstacle sits inside the fluid block: the boundary pass pushes the particles it contains out before
        # the density of the first step
        ball = RigidBodyDescription("sphere", center=(0.02, -0.56, 0.12), radius=0.09, padding=2 * r,
                                    resolution=(10, 10, 10))
    return [desc], [(x, v.astype(np.float32))], [[box, ball]]


def contact_scene(kind="floor"):
    """Scenes for the adjoint through boundary contact: fixed 4 ms steps, 3 + 3 Jacobi iterations, viscosity 20 with
    no boundary viscosity, r = 0.01, velocities -0.5 m/s downwards plus a smooth field. Each has one extra particle
    starting inside the solid (pushed out by 0.1 r in the first step, velocity zeroed), placed off the map's cell
    faces (where the interpolant's gradient jumps).
    floor: a 16 x 8 x 16 block (2,048 particles) whose two lowest layers sample the floor of a 1 m box (flat wall),
    spaced 2.1 r so that no lattice pair sits at the support radius (a difference quotient would see the lists change).
    sphere: an 8 x 5 x 8 block on the top of a sphere of radius 0.06 (curved wall: the normal's derivative matters).
    Returns (desc, x (float64), v (float64), bodies)."""
    rng = np.random.default_rng(1)
    r = 0.01
    if kind == "floor":
        x = scenes.lattice_block((0.3, 0.025, 0.3), (16, 8, 16), 2.1 * r, 0.01 * r, rng).astype(np.float64)
        x = np.vstack([x, [[0.311, 0.006, 0.309]]])
        bodies = [RigidBodyDescription("box", center=(0.5, 0.5, 0.5), half_extents=(0.5, 0.5, 0.5), inverted=True,
                                       resolution=(20, 20, 20))]
    else:
        x = scenes.lattice_block((0.23, 0.362, 0.23), (8, 5, 8), 2 * r, 0.01 * r, rng).astype(np.float64)
        x = np.vstack([x, [[0.303, 0.348, 0.298]]])
        bodies = [RigidBodyDescription("sphere", center=(0.3, 0.3, 0.3), radius=0.06, resolution=(16, 16, 16))]
    v = 0.2 * np.sin(7 * x[:, [1, 2, 0]]) + 0.05 * rng.normal(size=x.shape)
    v[:, 1] -= 0.5
    d = DFSPHSimulationDescription(
        ParticleRadius=r, Viscosity=0.02, BoundaryViscosity=0.0, EnableSurfaceTensionSolver=0, EnableViscositySolver=1,
        MaxViscositySolverError=1e-5, MaxViscositySolverIterations=200, MinDivergenceSolverIterations=3,
        MaxDivergenceSolverIterations=3, MinPressureSolverIterations=3, MaxPressureSolverIterations=3,
        MinTimeStepSize=0.004, MaxTimeStepSize=0.004, TimeStepSize=0.004)
    return d, x, v, bodies


def st_scene(n_side=(8, 8, 8)):
    """A free viscous block for the adjoint through surface tension: spaced 2.1 r (no lattice pair at the support
    radius), surface tension 0.1, two blend passes, 16 Monte-Carlo samples per particle (CSDFix), fixed 4 ms steps,
    3 + 3 Jacobi iterations. Returns (desc, x (float64), v (float64))."""
    rng = np.random.default_rng(1)
    r = 0.01
    x = scenes.lattice_block((0.3, 0.3, 0.3), n_side, 2.1 * r, 0.01 * r, rng).astype(np.float64)
    v = 0.3 * np.sin(7 * x[:, [1, 2, 0]]) + 0.1 * rng.normal(size=x.shape)
    d = DFSPHSimulationDescription(
        ParticleRadius=r, Viscosity=0.02, BoundaryViscosity=0.0, EnableSurfaceTensionSolver=1, SurfaceTension=0.1,
        SurfaceTensionSmoothPassCount=2, CSDFix=16, EnableViscositySolver=1, MaxViscositySolverError=1e-5,
        MaxViscositySolverIterations=200, MinDivergenceSolverIterations=3, MaxDivergenceSolverIterations=3,
        MinPressureSolverIterations=3, MaxPressureSolverIterations=3, MinTimeStepSize=0.004, MaxTimeStepSize=0.004,
        TimeStepSize=0.004)
    return d, x, v


def make_gpu(descs, fluids, bodies):
    from vfd_amd import DFSPHBatch, DFSPHSimulation
    if len(descs) == 1:
        sim = DFSPHSimulation(descs[0])
        sim.set_fluid(*fluids[0])
        sim.SetRigidBodies(bodies[0])
        return sim
    b = DFSPHBatch(descs)
    for s, (f, bl) in enumerate(zip(fluids, bodies)):
        b.scenes[s].set_fluid(*f)
        b.scenes[s].SetRigidBodies(bl)
    return b


def make_oracle(descs, fluids, bodies, fp64=False, scene=0):
    from oracle.oracle import OracleSim
    o = OracleSim(descs[scene], fp64=fp64)
    o.set_fluid(*fluids[scene])
    for b in bodies[scene]:
        if getattr(b, "is_mesh", False):
            o.add_boundary_mesh(b.to_abi(), b.mesh.vertices, b.mesh.triangles)
        else:
            o.add_boundary(b.to_abi())
    return o
