"""Synthetic scenes for the five benchmark configurations (seeded, reproducible).

The reference samples fluid volumes from meshes (ParticleSampler::SampleMeshVolume, Utility/Sampler/
ParticleSampler.cpp:7-91) and builds boundaries from meshes (RigidBody.cu:10-73). These scenes use cubic lattices
and analytic boundary shapes instead; every choice the configuration text leaves open is fixed here:

  - Jacobi solves run at least MinPressure/MinDivergenceSolverIterations = 2 (with the reference default 0 both
    loops run zero iterations, DFSPHImplementation.cu:458-461, 529-532);
  - "dynamic viscosity mu" maps to Viscosity = mu / rho0 (the code multiplies by rho0, :347) and
    BoundaryViscosity = Viscosity;
  - a CG tolerance of 1e-3 is MaxViscositySolverError = 0.1 (threshold tol^2 * 1e-4 * |b|^2, :661);
  - Jacobi tolerances are the description's own units (eta = tol * 1e-4 * rho0, :455, :526). Config 1 states
    them (divergence 0.1, density 0.01); the other configurations state none and use the reference defaults (10 and
    10, i.e. a mean density error below 0.1 %);
  - fluid blocks start one support radius (4r) away from walls: the reference's single-point boundary term
    V_b W(x - x_b) (DFSPHKernels.cu:176-184) gives a lattice layer at distance 2r from a wall about 2.5 rho0, which
    the first pressure solve turns into velocities of metres per second.

Each builder returns (descs, fluids, bodies): one vfd_desc per scene, one (positions, velocities) pair per scene and
one list of RigidBodyDescription per scene.
"""
import math

import numpy as np

from .simulation import DFSPHSimulationDescription, RigidBodyDescription

RHO0 = 1000.0


def lattice_block(origin, counts, spacing, jitter, rng):
    """Cubic lattice of counts[0] x counts[1] x counts[2] points; x fastest. Uniform jitter in [-jitter, jitter]."""
    nx, ny, nz = counts
    i = np.arange(nx, dtype=np.float64)
    j = np.arange(ny, dtype=np.float64)
    k = np.arange(nz, dtype=np.float64)
    Z, Y, X = np.meshgrid(k, j, i, indexing="ij")
    p = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1) * spacing + np.asarray(origin, np.float64)
    if jitter:
        p = p + rng.uniform(-jitter, jitter, size=p.shape)
    return p.astype(np.float32)


def sampler_lattice(mode, origin, counts, r, jitter, rng):
    """The lattice patterns of ParticleSampler::SampleMeshVolume (Utility/Sampler/ParticleSampler.cpp:19-73) over
    counts[0] x counts[1] x counts[2] points (x fastest, then y, then z, the reference's loop order), starting at
    `origin`, with uniform jitter in [-jitter, jitter]:
      "min"    MinDensity: cubic, spacing 2r, points at the cell centres (+r);
      "medium" MediumDensity: x step 2r, y step sqrt(3) r, z step 2r; even rows at (x, y + r, z + r), odd rows at
               (x + r, y + r, z) - 6.93 r^3 per particle, 34 neighbours within 4r in the interior;
      "max"    MaxDensity: x step sqrt(3) r, y step 2 sqrt(6) r / 3, z step 2r; at (x, y + r, z + r) shifted by
               +-r in z on odd columns (sign by row parity) and by (sqrt(3) r / 2, 0, r) on odd rows - 5.66 r^3 per
               particle, 50 neighbours within 4r in the interior.
    The point of column i, row j, layer k uses the loop variables i * shiftX, j * shiftY, k * 2r (the reference
    accumulates them in float; here they are products, a difference of a few ulps in a synthetic input)."""
    nx, ny, nz = counts
    d = 2.0 * r
    sx, sy = {"min": (d, d), "medium": (d, math.sqrt(3.0) * r),
              "max": (math.sqrt(3.0) * r, math.sqrt(6.0) * d / 3.0)}[mode]
    K, J, I = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    I, J, K = I.ravel(), J.ravel(), K.ravel()
    x, y, z = I * sx, J * sy, K * d
    odd_row, odd_col = (J % 2).astype(np.float64), (I % 2).astype(np.float64)
    if mode == "min":
        p = np.stack([x + r, y + r, z + r], 1)
    elif mode == "medium":
        p = np.stack([x + odd_row * r, y + r, z + (1.0 - odd_row) * r], 1)
    else:
        dz = odd_col * np.where(J % 2 == 1, -r, r) + odd_row * r
        p = np.stack([x + odd_row * (sx / 2.0), y + r, z + r + dz], 1)
    p = p + np.asarray(origin, np.float64)
    if jitter:
        p = p + rng.uniform(-jitter, jitter, size=p.shape)
    return p.astype(np.float32)


def dam_break_8k(seed=1):
    """Config 1: 20^3 = 8,000 particles, r = 0.025, spacing 2r, jitter +-0.01r, in the corner of a 2 m box
    ([-1, 1]^3, the span of the reference's Cube.obj) given as an inverted-box volume map, the block one support radius (4r) off the walls; no viscosity, no surface
    tension; divergence tolerance 0.1 %, density tolerance 0.01 %; CFL 0.4, dt <= 5 ms."""
    rng = np.random.default_rng(seed)
    r = 0.025
    desc = DFSPHSimulationDescription(
        ParticleRadius=r, TimeStepSize=0.001, MaxTimeStepSize=0.005, MinTimeStepSize=0.0001,
        EnableViscositySolver=0, EnableSurfaceTensionSolver=0, Viscosity=0.0, BoundaryViscosity=0.0,
        MaxDivergenceSolverError=0.1, MaxPressureSolverError=0.01,
        MinDivergenceSolverIterations=2, MinPressureSolverIterations=2, FrameCount=200)
    origin = (-1.0 + 4 * r,) * 3  # one support radius from the walls
    x = lattice_block(origin, (20, 20, 20), 2 * r, 0.01 * r, rng)
    v = np.zeros_like(x)
    box = RigidBodyDescription("box", center=(0, 0, 0), half_extents=(1, 1, 1), inverted=True, resolution=(20, 20, 20))
    return [desc], [(x, v)], [[box]]


def viscous_coil_250k(seed=2, n=250_000):
    """Config 2: a 250k-particle column (radius 0.1 m, r = 0.005) whose bottom starts 0.5 m above a ground plane,
    falling at 1 m/s plus seeded jitter; mu = 50 (nu = 0.05), CG tolerance 1e-3, surface tension 0.05; ground plane
    and walls (an inverted box) as volume maps."""
    rng = np.random.default_rng(seed)
    r = 0.005
    d = 2 * r
    rc = 0.1
    m = int(math.ceil(rc / d)) + 1
    g = np.arange(-m, m + 1) * d
    gx, gz = np.meshgrid(g, g, indexing="ij")
    disc = np.stack([gx.ravel(), gz.ravel()], axis=1)
    disc = disc[(disc ** 2).sum(1) <= (rc - r) ** 2]
    layers = int(math.ceil(n / len(disc)))
    ys = 0.5 + r + np.arange(layers) * d
    pts = np.concatenate([np.stack([disc[:, 0], np.full(len(disc), y), disc[:, 1]], 1) for y in ys])[:n]
    pts = pts + rng.uniform(-0.01 * r, 0.01 * r, size=pts.shape)
    x = pts.astype(np.float32)
    v = np.zeros_like(x)
    v[:, 1] = -1.0
    v += rng.uniform(-0.01, 0.01, size=v.shape).astype(np.float32)
    top = float(x[:, 1].max()) + 0.5
    desc = DFSPHSimulationDescription(
        ParticleRadius=r, Viscosity=50.0 / RHO0, BoundaryViscosity=50.0 / RHO0, MaxViscositySolverError=0.1,
        EnableSurfaceTensionSolver=1, SurfaceTension=0.05, CSD=10000,
        MinDivergenceSolverIterations=2, MinPressureSolverIterations=2,
        MaxDivergenceSolverError=10.0, MaxPressureSolverError=10.0, FrameCount=500)
    plane = RigidBodyDescription("plane", center=(0, 0, 0), normal=(0, 1, 0), resolution=(24, 6, 24),
                                 domain=((-0.6, -0.12, -0.6), (0.6, 0.12, 0.6)))
    walls = RigidBodyDescription("box", center=(0, top / 2 - 0.05, 0), half_extents=(0.5, top / 2 + 0.05, 0.5),
                                 inverted=True, resolution=(16, max(16, int(top / 0.06)), 16))
    return [desc], [(x, v.astype(np.float32))], [[plane, walls]]


def dam_break_2m(seed=3, lattice="medium"):
    """Config 3: 2,000,000 particles (r = 0.004, jitter +-0.01r) in a 2 x 1 x 1 m tank (inverted box) with a sphere
    obstacle of radius 0.15; viscosity 0.01. The block is 200 x 100 x 100 points of a sampler lattice:
      "medium" (default): the reference's MediumDensity pattern, the one its DFSPH example scene uses, with about 35
               neighbours per particle as the configuration states; the block is 1.6 x 0.69 x 0.8 m, 44 % of the
               tank (40 % holds 1.8M particles at this density);
      "cubic": a cubic lattice of spacing 2r (28.5 neighbours), 1.6 x 0.8 x 0.8 m, 51 % of the tank;
      "max":   the MaxDensity pattern (about 48 neighbours, 1.13 rest densities), 1.39 x 0.65 x 0.8 m, 36 %."""
    rng = np.random.default_rng(seed)
    r = 0.004
    desc = DFSPHSimulationDescription(
        ParticleRadius=r, Viscosity=0.01, BoundaryViscosity=0.01, MaxViscositySolverError=0.1,
        EnableSurfaceTensionSolver=0, MaxDivergenceSolverError=10.0, MaxPressureSolverError=10.0,
        MinDivergenceSolverIterations=2, MinPressureSolverIterations=2, FrameCount=100)
    if lattice == "cubic":
        x = lattice_block((4 * r, 4 * r, 4 * r), (200, 100, 100), 2 * r, 0.01 * r, rng)
    else:
        origin = (4 * r, 3 * r, 4 * r) if lattice == "medium" else (4 * r, 3 * r, 3 * r)  # lowest points 4r in
        x = sampler_lattice(lattice, origin, (200, 100, 100), r, 0.01 * r, rng)
    v = np.zeros_like(x)
    tank = RigidBodyDescription("box", center=(1.0, 0.5, 0.5), half_extents=(1.0, 0.5, 0.5), inverted=True,
                                resolution=(40, 24, 24))
    # Padding 2r keeps particle centres r outside the sphere (with Padding 0 a solid admits them r deep,
    # RigidBody.cu:29,39-42)
    sphere = RigidBodyDescription("sphere", center=(1.8, 0.15, 0.5), radius=0.15, padding=2 * r,
                                  resolution=(16, 16, 16))
    return [desc], [(x, v)], [[tank, sphere]]


def dam_break_16m(seed=6):
    """The 2M dam break at eight times the particles (a scale point, not a benchmark configuration of the reference):
    16,000,000 particles, a 400 x 200 x 200 lattice at r = 0.002 in the same 2 x 1 x 1 m tank with the same sphere."""
    rng = np.random.default_rng(seed)
    r = 0.002
    desc = DFSPHSimulationDescription(
        ParticleRadius=r, Viscosity=0.01, BoundaryViscosity=0.01, MaxViscositySolverError=0.1,
        EnableSurfaceTensionSolver=0, MaxDivergenceSolverError=10.0, MaxPressureSolverError=10.0,
        MinDivergenceSolverIterations=2, MinPressureSolverIterations=2, FrameCount=100)
    x = lattice_block((4 * r, 4 * r, 4 * r), (400, 200, 200), 2 * r, 0.01 * r, rng)
    v = np.zeros_like(x)
    tank = RigidBodyDescription("box", center=(1.0, 0.5, 0.5), half_extents=(1.0, 0.5, 0.5), inverted=True,
                                resolution=(80, 48, 48))
    sphere = RigidBodyDescription("sphere", center=(1.8, 0.15, 0.5), radius=0.15, padding=2 * r,
                                  resolution=(32, 32, 32))
    return [desc], [(x, v)], [[tank, sphere]]


def batched_uneven(seed=4, scenes=64, r=0.01):
    """Config 4: 64 dam breaks with N_s = round(exp U(ln 5e3, ln 6e4)) and mu_s = exp U(ln 0.1, ln 100), one seed;
    every scene has its own box."""
    rng = np.random.default_rng(seed)
    ns = np.rint(np.exp(rng.uniform(math.log(5e3), math.log(6e4), size=scenes))).astype(int)
    mus = np.exp(rng.uniform(math.log(0.1), math.log(100.0), size=scenes))
    descs, fluids, bodies = [], [], []
    d = 2 * r
    for s in range(scenes):
        n = int(ns[s])
        side = int(math.ceil(n ** (1.0 / 3.0)))
        counts = (side, side + 1, side)
        x = lattice_block((2 * d, 2 * d, 2 * d), counts, d, 0.01 * r, rng)[:n]
        descs.append(DFSPHSimulationDescription(
            ParticleRadius=r, Viscosity=float(mus[s]) / RHO0, BoundaryViscosity=float(mus[s]) / RHO0,
            EnableSurfaceTensionSolver=0, MaxViscositySolverError=0.1, MaxDivergenceSolverError=10.0,
            MaxPressureSolverError=10.0, MinDivergenceSolverIterations=2, MinPressureSolverIterations=2,
            FrameCount=200))
        fluids.append((x, np.zeros_like(x)))
        w = side * d + 4 * d
        bodies.append([RigidBodyDescription("box", center=(w, 0.75 * w, w / 2), half_extents=(w, 0.75 * w, w / 2),
                                            inverted=True, resolution=(16, 12, 8))])
    return descs, fluids, bodies


def fourier_field(x, rng, modes=8, amplitude=0.5, wavelength=1.0):
    """Smooth vector field at the points x: a sum of `modes` plane waves sin(k.x + phase) a with wave vectors
    k ~ N(0, (2 pi / wavelength)^2 I), uniform phases and a ~ N(0, I), scaled so that max |component| = amplitude."""
    k = rng.normal(size=(modes, 3)) * 2 * math.pi / wavelength
    ph = rng.uniform(0, 2 * math.pi, size=modes)
    amp = rng.normal(size=(modes, 3))
    xd = np.asarray(x, np.float64)
    v = np.zeros_like(xd)
    for m in range(modes):
        v += np.sin(xd @ k[m] + ph[m])[:, None] * amp[m]
    return v * (amplitude / max(np.abs(v).max(), 1e-12))


def falling_viscous_block(seed=5, n_side=(50, 40, 50), mu=20.0, differentiable=False):
    """Config 5: 100k particles in a falling viscous block with a smooth seeded initial velocity (8 random Fourier
    modes, amplitude 0.5 m/s); target mu = 20. n_side=(16, 8, 16) gives the 2,048-particle version.

    differentiable=True fixes what the adjoint holds constant, so that the 60-step map is smooth in (v0, mu) and
    central differences see the same map the adjoint differentiates: a fixed time step of 2 ms (the block falls
    about 7 cm in 60 steps and never reaches the walls), fixed Jacobi iteration counts (divergence 2, pressure 3) and
    a viscosity solve converged to a relative residual of 1e-5 (MaxViscositySolverError = 1e-3)."""
    rng = np.random.default_rng(seed)
    r = 0.01
    x = lattice_block((0.3, 0.3, 0.3), n_side, 2 * r, 0.01 * r, rng)
    v = fourier_field(x, rng)
    fields = dict(
        ParticleRadius=r, Viscosity=mu / RHO0, BoundaryViscosity=mu / RHO0, EnableSurfaceTensionSolver=0,
        MaxViscositySolverError=0.1, MaxDivergenceSolverError=10.0, MaxPressureSolverError=10.0,
        MinDivergenceSolverIterations=2, MinPressureSolverIterations=2, FrameCount=60)
    if differentiable:
        fields.update(TimeStepSize=0.002, MinTimeStepSize=0.002, MaxTimeStepSize=0.002,
                      MaxDivergenceSolverIterations=2, MinPressureSolverIterations=3, MaxPressureSolverIterations=3,
                      MaxViscositySolverError=1e-3)
    desc = DFSPHSimulationDescription(**fields)
    box = RigidBodyDescription("box", center=(1, 1, 1), half_extents=(1, 1, 1), inverted=True, resolution=(20, 20, 20))
    return [desc], [(x, v.astype(np.float32))], [[box]]


CONFIGS = {
    "dam_break_8k": dam_break_8k,
    "viscous_coil_250k": viscous_coil_250k,
    "dam_break_2m": dam_break_2m,
    "dam_break_2m_cubic": lambda seed=3: dam_break_2m(seed, lattice="cubic"),
    "dam_break_2m_max": lambda seed=3: dam_break_2m(seed, lattice="max"),
    "dam_break_16m": dam_break_16m,
    "batched_uneven_64": batched_uneven,
    "falling_viscous_block_100k": falling_viscous_block,
}


def load_scene_file(path, resource_root):
    """The DFSPH inputs of a VFD scene file (an entt snapshot in cereal JSON written by Scene::Save,
    Scene/Scene.cpp:10-40), as the editor's "Bake Objects" builds them (Editor/Panels/ComponentPanel.cpp:608-656):
    the first simulation description, one FluidObjectDescription per entity holding a fluid-object and a mesh
    component, one mesh RigidBodyDescription per entity holding a rigid-body and a mesh component.

    Components follow their entity id in the snapshot; the component objects are recognised by their keys
    ("transform", "meshSource", {"inverted", "padding", "collisionMapResolution"}, {"inverted", "resolution",
    "sampleMode", "velocity"}, "description"). Mesh paths ("Resources/Models/Cube.obj") are resolved under
    `resource_root` (its last component may stand for "Resources/Models", as for tests/golden/models). Transforms are
    the stored column-major matrices of entities without parents; the reference decomposes a loaded matrix into
    translation, Euler angles and scale and recomposes it (TransformComponent.h:33-38), which may move entries by
    a few ulps. Returns (description, [FluidObjectDescription], [RigidBodyDescription])."""
    import json
    import os

    from .mesh import TriangleMesh
    from .simulation import FluidObjectDescription, read_descriptions

    with open(path, "r") as f:
        data = json.load(f)
    values = list(data.values())
    comps = {}
    for k in range(1, len(values)):
        v, ent = values[k], values[k - 1]
        if not isinstance(v, dict) or isinstance(ent, (dict, list, bool)) or not isinstance(ent, int):
            continue
        keys = set(v)
        if "transform" in keys:
            cols = [v["transform"]["value%d" % c] for c in range(4)]
            m = np.array([[cols[c][a] for c in range(4)] for a in "xyzw"], dtype=np.float32)
            comps.setdefault(ent, {})["transform"] = m
        elif "meshSource" in keys:
            comps.setdefault(ent, {})["mesh"] = v["meshSource"]
        elif {"inverted", "padding", "collisionMapResolution"} <= keys:
            comps.setdefault(ent, {})["rigid"] = v
        elif {"inverted", "resolution", "sampleMode", "velocity"} <= keys:
            comps.setdefault(ent, {})["fluid"] = v
    descs = read_descriptions(path)
    if not descs:
        raise ValueError("the scene holds no DFSPH simulation description")

    def mesh_of(src):
        rel = src.replace("\\", "/")
        for cand in (os.path.join(resource_root, rel), os.path.join(resource_root, os.path.basename(rel))):
            if os.path.exists(cand):
                return TriangleMesh.load_obj(cand)
        raise FileNotFoundError(src)

    fluids, bodies = [], []
    for ent in sorted(comps):
        c = comps[ent]
        if "mesh" not in c:
            continue
        tr = c.get("transform", np.eye(4, dtype=np.float32))
        if "fluid" in c:
            fo = c["fluid"]
            fluids.append(FluidObjectDescription(
                mesh_of(c["mesh"]), transform=tr, inverted=bool(fo["inverted"]),
                resolution=tuple(int(fo["resolution"][a]) for a in "xyz"), sample_mode=int(fo["sampleMode"]),
                velocity=tuple(float(fo["velocity"][a]) for a in "xyz")))
        if "rigid" in c:
            rb = c["rigid"]
            bodies.append(RigidBodyDescription(
                "mesh", mesh=mesh_of(c["mesh"]), transform=tr, inverted=bool(rb["inverted"]),
                padding=float(rb["padding"]), resolution=tuple(int(rb["collisionMapResolution"][a]) for a in "xyz")))
    return descs[0], fluids, bodies
