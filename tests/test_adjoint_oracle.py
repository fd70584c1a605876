"""The FP64 reference adjoint of the oracle (the result the HIP adjoint is compared with), pinned by central finite
differences of a forward whose kernel is the spline itself (oracle smooth mode), where the adjoint is the exact
derivative.

Setup: a free 512-particle block with viscosity 20, fixed time step and fixed Jacobi iteration counts (the adjoint
holds time steps, iteration counts, clamp masks and neighbour lists fixed, as the forward's derivative does between
discrete events); the loss is a random linear functional of the positions and velocities after 3 steps."""
import copy

import numpy as np
import pytest

from oracle.oracle import OracleSim
from vfd_amd import scenes
from vfd_amd.simulation import DFSPHSimulationDescription

STEPS = 3


def _setup():
    rng = np.random.default_rng(1)
    r = 0.01
    x = scenes.lattice_block((0.3, 0.3, 0.3), (8, 8, 8), 2 * r, 0.01 * r, rng).astype(np.float64)
    v = 0.3 * np.sin(7 * x[:, [1, 2, 0]]) + 0.1 * rng.normal(size=x.shape)
    d = DFSPHSimulationDescription(
        ParticleRadius=r, Viscosity=0.02, BoundaryViscosity=0.0, EnableSurfaceTensionSolver=0, EnableViscositySolver=1,
        MaxViscositySolverError=1e-5, MaxViscositySolverIterations=200, MinDivergenceSolverIterations=3,
        MaxDivergenceSolverIterations=3, MinPressureSolverIterations=3, MaxPressureSolverIterations=3,
        MinTimeStepSize=0.004, MaxTimeStepSize=0.004, TimeStepSize=0.004)
    return d, x, v


def _run(d, x, v, tape=False):
    o = OracleSim(d, fp64=True)
    o.set_smooth_kernel(True)
    o.set_fluid(x.astype(np.float32), v.astype(np.float32))
    o.set_field("position", x)
    o.set_field("velocity", v)
    if tape:
        o.tape(True)
    for _ in range(STEPS):
        o.step()
    return o


def test_reference_adjoint_matches_finite_differences():
    d, x, v = _setup()
    n = len(x)
    rng = np.random.default_rng(7)
    wx, wv = rng.normal(size=(n, 3)), rng.normal(size=(n, 3))

    def loss(o):
        return float((o.field("position") * wx).sum() + (o.field("velocity") * wv).sum())

    o = _run(d, x, v, tape=True)
    lx0, lv0, lmu = o.adjoint(wx, wv)
    dv, dx = rng.normal(size=(n, 3)), rng.normal(size=(n, 3))
    fd_v = (loss(_run(d, x, v + 1e-5 * dv)) - loss(_run(d, x, v - 1e-5 * dv))) / 2e-5
    fd_x = (loss(_run(d, x + 1e-7 * dx, v)) - loss(_run(d, x - 1e-7 * dx, v))) / 2e-7
    dp, dm = copy.copy(d), copy.copy(d)
    dp.Viscosity, dm.Viscosity = d.Viscosity * 1.0001, d.Viscosity * 0.9999
    mu_p, mu_m = 1000.0 * float(np.float32(dp.Viscosity)), 1000.0 * float(np.float32(dm.Viscosity))
    fd_mu = (loss(_run(dp, x, v)) - loss(_run(dm, x, v))) / (mu_p - mu_m)
    errs = {"v0": abs(float((lv0 * dv).sum()) - fd_v) / abs(fd_v), "x0": abs(float((lx0 * dx).sum()) - fd_x) / abs(fd_x),
            "mu": abs(lmu - fd_mu) / abs(fd_mu)}
    print("adjoint vs central differences (relative):", errs)
    for k, e in errs.items():
        assert e < 1e-4, (k, e)


def test_tape_rejects_boundary_contact():
    scene = scenes.dam_break_8k()
    o = OracleSim(scene[0][0], fp64=True)
    o.set_fluid(*scene[1][0])
    for b in scene[2][0]:
        o.add_boundary(b.to_abi())
    o.tape(True)
    o.step()
    with pytest.raises(RuntimeError, match="boundary"):
        o.adjoint(np.zeros((o.n, 3)))
