"""The HIP adjoint (vfd_tape / vfd_adjoint) against the oracle's reference adjoint on a 2,048-particle viscous block,
free and landing on a floor, and on a block landing on a sphere.

The oracle runs the same forward in its FP32 mirror (identical lists, masks and iteration counts, so the two adjoints
differ only in summation order) and in its FP64 twin (the FP64 reference gradient). The differentiated map and its
scope are those of vfd_adjoint: fixed time step, fixed Jacobi iteration counts, boundary contact (with and without
boundary viscosity), surface tension (with and without temporal smoothing)."""
import numpy as np
import pytest

from oracle.oracle import OracleSim
from scene_util import contact_scene, st_scene
from vfd_amd import DFSPHSimulation, scenes
from vfd_amd.simulation import DFSPHSimulationDescription

pytestmark = pytest.mark.gpu
STEPS = 3


def _setup(n_side=(16, 8, 16)):
    rng = np.random.default_rng(3)
    r = 0.01
    x = scenes.lattice_block((0.3, 0.3, 0.3), n_side, 2 * r, 0.01 * r, rng)
    v = (0.3 * np.sin(7 * x[:, [1, 2, 0]]) + 0.1 * rng.normal(size=x.shape)).astype(np.float32)
    d = DFSPHSimulationDescription(
        ParticleRadius=r, Viscosity=0.02, BoundaryViscosity=0.0, EnableSurfaceTensionSolver=0, EnableViscositySolver=1,
        MaxViscositySolverError=1e-4, MaxViscositySolverIterations=200, MinDivergenceSolverIterations=3,
        MaxDivergenceSolverIterations=3, MinPressureSolverIterations=3, MaxPressureSolverIterations=3,
        MinTimeStepSize=0.004, MaxTimeStepSize=0.004, TimeStepSize=0.004)
    return d, x, v


def _rel(a, b):
    return abs(a - b) / max(abs(b), 1e-30)


def test_adjoint_matches_oracle_reference():
    d, x, v = _setup()
    n = len(x)
    rng = np.random.default_rng(11)
    wx, wv = rng.normal(size=(n, 3)), rng.normal(size=(n, 3))
    gpu = DFSPHSimulation(d)
    gpu.set_fluid(x, v)
    gpu.tape(True)
    for _ in range(STEPS):
        gpu.OnUpdate()
    gx, gv, gmu = gpu.adjoint(wx, wv)
    assert gpu.taped_steps() == STEPS
    dirs = rng.normal(size=(2, n, 3))
    out = {}
    for fp64 in (False, True):
        o = OracleSim(d, fp64=fp64)
        o.set_fluid(x, v)
        o.tape(True)
        for _ in range(STEPS):
            o.step()
        if not fp64:
            assert np.array_equal(o.field("position").astype(np.float32), gpu.field("position"))
        ox, ov, omu = o.adjoint(wx, wv)
        e = {"x0": _rel(float((gx * dirs[0]).sum()), float((ox * dirs[0]).sum())),
             "v0": _rel(float((gv * dirs[1]).sum()), float((ov * dirs[1]).sum())),
             "mu": _rel(gmu, omu),
             "x0_inf": float(np.abs(gx - ox).max() / np.abs(ox).max()),
             "v0_inf": float(np.abs(gv - ov).max() / np.abs(ov).max())}
        out["fp64" if fp64 else "fp32"] = e
        print("gpu adjoint vs oracle", "fp64" if fp64 else "fp32", {k: "%.2e" % v for k, v in e.items()}, flush=True)
    for tag, e in out.items():
        for k in ("x0", "v0", "mu"):
            assert e[k] < 1e-3, (tag, k, e[k])


@pytest.mark.parametrize("kind,contacts,bvisc", [("floor", 512, 0.0), ("sphere", 44, 0.0), ("floor", 512, 0.02),
                                                 ("sphere", 44, 0.05)])
def test_adjoint_through_boundary_contact_matches_oracle(kind, contacts, bvisc):
    """Contact samples (density, factor, divergence and pressure terms through the volume map, and with a boundary
    viscosity its tangential samples in the viscosity matrix) and the push-out of one particle starting inside the
    solid. Directions over all particles, over the contact particles of the first
    step and on the pushed particle; the pushed particle's entry velocity has no effect. Against the FP32 mirror
    (same forward, bit for bit) the bar is 1e-3; against the FP64 twin it is 1e-3 or twice the mirror's own distance
    from the twin, whichever is larger: the FP32 and FP64 forwards part slightly (on the sphere the mirror's mu
    derivative is 8e-4 from the twin's), and the GPU cannot be closer to the twin than the forward it differentiates."""
    d, x, v, bodies = contact_scene(kind)
    d.BoundaryViscosity = bvisc
    x32, v32 = x.astype(np.float32), v.astype(np.float32)
    n = len(x)
    rng = np.random.default_rng(11)
    wx, wv = rng.normal(size=(n, 3)), rng.normal(size=(n, 3))
    print("contact adjoint", kind, n, "particles, boundary viscosity", bvisc, flush=True)
    gpu = DFSPHSimulation(d)
    gpu.set_fluid(x32, v32)
    gpu.SetRigidBodies(bodies)
    gpu.tape(True)
    for _ in range(STEPS):
        gpu.OnUpdate()
    gx, gv, gmu = gpu.adjoint(wx, wv)
    probe = OracleSim(d)
    probe.set_fluid(x32, v32)
    for b in bodies:
        probe.add_boundary(b.to_abi())
    probe.step_until("boundary")
    contact = probe.field("boundary_volume") > 0
    assert contact.sum() == contacts
    push = np.zeros((n, 3))
    push[-1] = rng.normal(size=3)
    dirs = {"x0": rng.normal(size=(n, 3)), "x0_contact": rng.normal(size=(n, 3)) * contact[:, None], "x0_push": push,
            "v0": rng.normal(size=(n, 3)), "v0_contact": rng.normal(size=(n, 3)) * contact[:, None]}
    assert not gv[-1].any()
    out, ref = {}, {}
    for fp64 in (False, True):
        o = OracleSim(d, fp64=fp64)
        o.set_fluid(x32, v32)
        for b in bodies:
            o.add_boundary(b.to_abi())
        o.tape(True)
        for _ in range(STEPS):
            o.step()
        if not fp64:
            assert np.array_equal(o.field("position").astype(np.float32), gpu.field("position"))
        ref[fp64] = o.adjoint(wx, wv)
        ox, ov, omu = ref[fp64]
        e = {k: _rel(float(((gx if k[0] == "x" else gv) * dv).sum()), float(((ox if k[0] == "x" else ov) * dv).sum()))
             for k, dv in dirs.items()}
        e["mu"] = _rel(gmu, omu)
        e["x0_inf"] = float(np.abs(gx - ox).max() / np.abs(ox).max())
        out["fp64" if fp64 else "fp32"] = e
        print("gpu contact adjoint vs oracle", kind, "fp64" if fp64 else "fp32", {k: "%.2e" % v for k, v in e.items()},
              flush=True)
    (mx, mv, mmu), (tx, tv, tmu) = ref[False], ref[True]
    mirror = {k: _rel(float(((mx if k[0] == "x" else mv) * dv).sum()), float(((tx if k[0] == "x" else tv) * dv).sum()))
              for k, dv in dirs.items()}
    mirror["mu"] = _rel(mmu, tmu)
    for k in list(dirs) + ["mu"]:
        assert out["fp32"][k] < 1e-3, ("fp32", k, out["fp32"][k])
        assert out["fp64"][k] < max(1e-3, 2 * mirror[k]), ("fp64", k, out["fp64"][k], mirror[k])


@pytest.mark.parametrize("n_side,temporal", [((8, 8, 8), 0), ((16, 8, 16), 0), ((8, 8, 8), 1), ((16, 8, 16), 1)])
def test_adjoint_through_surface_tension_matches_oracle(n_side, temporal):
    """Surface tension (two blend passes, 16 Monte-Carlo samples per particle) with its Monte-Carlo normals and
    curvatures held, against the oracle's FP32 mirror and FP64 twin (bars as for contact). With temporal smoothing
    the blended curvature carried from step to step has a cotangent of its own across the three recorded steps."""
    d, x, v = st_scene(n_side)
    d.TemporalSmoothing = temporal
    x32, v32 = x.astype(np.float32), v.astype(np.float32)
    n = len(x)
    rng = np.random.default_rng(11)
    wx, wv = rng.normal(size=(n, 3)), rng.normal(size=(n, 3))
    print("surface tension adjoint", n, "particles, temporal smoothing", temporal, flush=True)
    gpu = DFSPHSimulation(d)
    gpu.set_fluid(x32, v32)
    gpu.tape(True)
    for _ in range(STEPS):
        gpu.OnUpdate()
    gx, gv, gmu = gpu.adjoint(wx, wv)
    dirs = {"x0": rng.normal(size=(n, 3)), "v0": rng.normal(size=(n, 3))}
    out, ref = {}, {}
    for fp64 in (False, True):
        o = OracleSim(d, fp64=fp64)
        o.set_fluid(x32, v32)
        o.tape(True)
        for _ in range(STEPS):
            o.step()
        if not fp64:
            assert np.array_equal(o.field("position").astype(np.float32), gpu.field("position"))
            assert (np.abs(o.field("mc_surface_normal")).sum(1) > 0).sum() > n // 5
        ref[fp64] = o.adjoint(wx, wv)
        ox, ov, omu = ref[fp64]
        e = {k: _rel(float(((gx if k[0] == "x" else gv) * dv).sum()), float(((ox if k[0] == "x" else ov) * dv).sum()))
             for k, dv in dirs.items()}
        e["mu"] = _rel(gmu, omu)
        e["x0_inf"] = float(np.abs(gx - ox).max() / np.abs(ox).max())
        out["fp64" if fp64 else "fp32"] = e
        print("gpu st adjoint vs oracle", "fp64" if fp64 else "fp32", {k: "%.2e" % v for k, v in e.items()}, flush=True)
    (mx, mv, mmu), (tx, tv, tmu) = ref[False], ref[True]
    mirror = {k: _rel(float(((mx if k[0] == "x" else mv) * dv).sum()), float(((tx if k[0] == "x" else tv) * dv).sum()))
              for k, dv in dirs.items()}
    mirror["mu"] = _rel(mmu, tmu)
    for k in list(dirs) + ["mu"]:
        assert out["fp32"][k] < 1e-3, ("fp32", k, out["fp32"][k])
        assert out["fp64"][k] < max(1e-3, 2 * mirror[k]), ("fp64", k, out["fp64"][k], mirror[k])


@pytest.mark.parametrize("scene", ["floor_contact", "falling_block_100k"])
def test_adjoint_group_loops_give_the_same_bits(scene):
    """The transposed neighbour loops (pressure solves, viscosity operator) per neighbour group (the step's halos
    rebuilt from its recorded lists and staged
    in LDS; with HaloLimit 0 every group gathers from HBM inside the same loop) against one thread per particle: the
    same terms in the same order, so the cotangents and d/dmu agree bit for bit. The floor scene has contact samples
    with boundary viscosity and a pushed particle; the 100k block spans 98 groups over the XCDs."""
    if scene == "floor_contact":
        d, x, v, bodies = contact_scene("floor")
        d.BoundaryViscosity = 0.02
        steps = STEPS
    else:
        descs, fluids, bl = scenes.falling_viscous_block(differentiable=True)
        d, (x, v), bodies = descs[0], fluids[0], bl[0]
        steps = 2
    x32, v32 = x.astype(np.float32), v.astype(np.float32)
    n = len(x)
    rng = np.random.default_rng(5)
    wx, wv = rng.normal(size=(n, 3)), rng.normal(size=(n, 3))
    gpu = DFSPHSimulation(d)
    gpu.set_fluid(x32, v32)
    gpu.SetRigidBodies(bodies)
    gpu.tape(True)
    for _ in range(steps):
        gpu.OnUpdate()
    out = {}
    for name, kw in [("particles", dict(AdjointGroupLoops=0)), ("groups", dict(AdjointGroupLoops=1)),
                     ("groups_gathered", dict(AdjointGroupLoops=1, HaloLimit=0))]:
        gpu.set_tuning(**kw)
        print("adjoint", scene, n, "particles,", name, flush=True)
        out[name] = gpu.adjoint(wx, wv)
    ref = out["particles"]
    assert np.isfinite(ref[0]).all() and np.abs(ref[0]).max() > 0
    for name in ("groups", "groups_gathered"):
        gx, gv, gmu = out[name]
        assert np.array_equal(gx, ref[0]) and np.array_equal(gv, ref[1]), name
        assert gmu == ref[2], (name, gmu, ref[2])


def test_batch_adjoint_matches_each_scene_alone():
    """Three scenes recorded in one batch and reversed one at a time (vfd_adjoint for scene 0, vfd_batch_adjoint for
    the others): a free viscous block; a block landing on a floor with boundary viscosity and a pushed particle; a
    block with surface tension and temporal smoothing and its own viscosity. Each scene has its own bodies, switches
    and iteration counts; its cotangents and d/dmu equal those of the scene recorded and reversed alone, bit for bit
    (the single-scene adjoint is the one compared with the oracle above)."""
    from vfd_amd import DFSPHBatch
    d0, x0, v0 = _setup((8, 8, 8))
    d1, x1, v1, b1 = contact_scene("floor")
    d1.BoundaryViscosity = 0.02
    d2, x2, v2 = st_scene()
    d2.TemporalSmoothing = 1
    d2.Viscosity = 0.05
    parts = [(d0, x0, v0, []), (d1, x1, v1, b1), (d2, x2, v2, [])]
    batch = DFSPHBatch([p[0] for p in parts])
    for s, (d, x, v, b) in enumerate(parts):
        batch.scenes[s].set_fluid(x.astype(np.float32), v.astype(np.float32))
        batch.scenes[s].SetRigidBodies(b)
    batch.tape(True)
    for _ in range(STEPS):
        batch.step()
    assert batch.taped_steps() == STEPS
    rng = np.random.default_rng(13)
    for s, (d, x, v, b) in enumerate(parts):
        n = len(x)
        wx, wv = rng.normal(size=(n, 3)), rng.normal(size=(n, 3))
        print("batch adjoint, scene", s, n, "particles", flush=True)
        gx, gv, gmu = batch.adjoint(s, wx, wv)
        one = DFSPHSimulation(d)
        one.set_fluid(x.astype(np.float32), v.astype(np.float32))
        one.SetRigidBodies(b)
        one.tape(True)
        for _ in range(STEPS):
            one.OnUpdate()
        assert np.array_equal(one.field("position"), batch.scenes[s].field("position")), s
        ox, ov, omu = one.adjoint(wx, wv)
        assert np.isfinite(gx).all() and np.abs(gx).max() > 0, s
        assert np.array_equal(gx, ox) and np.array_equal(gv, ov), s
        assert gmu == omu, (s, gmu, omu)


def test_checkpointed_rollout_matches_oracle_fp64():
    """torch.autograd through Rollout (60 steps, checkpoints every 10) on the 2,048-particle falling viscous block
    against the oracle's FP64 reference adjoint of the same 60 steps taken in one piece; the loss is the 64^3
    density-splat error against a run at mu = 20, estimate mu = 5."""
    import torch
    from vfd_amd import adjoint

    ro, v0, loss = adjoint.viscosity_inversion(n_side=(16, 8, 16), steps=60, every=10)
    v = v0.clone().requires_grad_(True)
    mu = torch.tensor(5.0, dtype=torch.float64, requires_grad=True)
    x, _ = ro(v, mu)
    xl = x.detach().clone().requires_grad_(True)
    L = loss(xl)
    L.backward()
    gx = xl.grad.numpy().astype(np.float64)
    x2, _ = ro(v, mu)
    assert torch.equal(x2, x)  # the forward is deterministic, so the re-runs from checkpoints repeat it
    loss(x2).backward()
    gv, gmu = v.grad.numpy().astype(np.float64), float(mu.grad)
    assert len(ro.checkpoints) == 6 and ro.sim.taped_steps() == 10

    d = ro.sim.GetDescription()
    o = OracleSim(d, fp64=True)
    o.set_fluid(ro.x0, v0.numpy())
    for b in scenes.falling_viscous_block(n_side=(16, 8, 16), differentiable=True)[2][0]:
        o.add_boundary(b.to_abi())
    o.tape(True)
    for _ in range(60):
        o.step()
    _, ov, omu = o.adjoint(gx, np.zeros_like(gx))
    rng = np.random.default_rng(5)
    dv = rng.normal(size=gv.shape)
    e = {"v0": _rel(float((gv * dv).sum()), float((ov * dv).sum())), "mu": _rel(gmu, omu),
         "v0_inf": float(np.abs(gv - ov).max() / np.abs(ov).max())}
    print("rollout adjoint vs oracle fp64 (60 steps)", {k: "%.2e" % v for k, v in e.items()}, ro.timings, flush=True)
    assert e["v0"] < 1e-3 and e["mu"] < 1e-3 and e["v0_inf"] < 1e-2


def test_second_forward_keeps_the_first_gradient():
    """Two forwards before one backward (as gradcheck, a line search or a loss over two rollouts do): the first call's
    gradient must come from its own checkpoints and viscosity, bit for bit the gradient of a lone call."""
    import torch
    from vfd_amd import adjoint

    ro, v0, loss = adjoint.viscosity_inversion(n_side=(16, 8, 16), steps=10, every=5)

    def grads(second):
        v = v0.clone().requires_grad_(True)
        mu = torch.tensor(5.0, dtype=torch.float64, requires_grad=True)
        x, _ = ro(v, mu)
        if second:
            ro(v0.clone() * 0.5, torch.tensor(12.0, dtype=torch.float64))  # other checkpoints, other viscosity
        loss(x).backward()
        return v.grad.clone(), float(mu.grad)

    g1, m1 = grads(False)
    g2, m2 = grads(True)
    assert torch.equal(g1, g2) and m1 == m2


def test_c5_gradient_against_differences_and_fp64_reference():
    """The 100k-particle inversion workload (60 steps, checkpoints every 10, mu = 5 against a target at mu = 20).

    (a) Central differences with eps = 1e-3 relative on the GPU forward, along mu and along v0 itself (scaling the
    initial field). The step is only piecewise smooth: a source is zeroed below 20 neighbours and gated by an epsilon,
    the Jacobi sources are clamped one-sided (max(rho_adv, 0) near rest density), and the kernel is read from a
    10,000-entry table, so a difference quotient also counts the threshold events its perturbation crosses. Along
    v0 they stay within 1e-3; mu moves every particle a little and meets about 1e-3 of them.
    (b) The oracle's FP64 reference adjoint of the same 60 steps (one recording, same kernel table): mu and the
    smooth directions within 1e-3. Per-particle values are printed, not asserted: over 60 steps the FP32 and FP64
    trajectories part at some threshold events, and the frozen-mask derivative differs locally where they do."""
    import copy

    import torch
    from vfd_amd import adjoint

    ro, v0, loss = adjoint.viscosity_inversion()
    v = v0.clone().requires_grad_(True)
    mu = torch.tensor(5.0, dtype=torch.float64, requires_grad=True)
    x, _ = ro(v, mu)
    loss(x).backward()
    gv, gmu = v.grad.numpy().astype(np.float64), float(mu.grad)
    print("c5 timings", ro.timings, "adjoint/forward %.2f" % (ro.timings["backward_s"] / ro.timings["forward_s"]),
          flush=True)
    v0n = v0.numpy()

    def gpu_loss(muv, vv):
        xx, _ = ro.forward_np(ro.x0, vv, muv)
        return float(loss(torch.as_tensor(xx))), ro.mu

    eps = 1e-3
    (lp, mp), (lm, mm) = gpu_loss(5.0 * (1 + eps), v0n), gpu_loss(5.0 * (1 - eps), v0n)
    vp, vm = (v0n * (1 + eps)).astype(np.float32), (v0n * (1 - eps)).astype(np.float32)
    fd = {"mu": (gmu, (lp - lm) / (mp - mm)),
          "v0_scale": (float((gv * (vp.astype(np.float64) - vm)).sum()), gpu_loss(5.0, vp)[0] - gpu_loss(5.0, vm)[0])}
    err_fd = {k: _rel(a, b) for k, (a, b) in fd.items()}
    print("c5 central differences (adjoint, fd):", fd, {k: "%.2e" % e for k, e in err_fd.items()}, flush=True)

    d = copy.copy(ro.sim.GetDescription())
    o = OracleSim(d, fp64=True)
    o.set_fluid(ro.x0, v0n)
    for b in scenes.falling_viscous_block(differentiable=True)[2][0]:
        o.add_boundary(b.to_abi())
    o.tape(True)
    for _ in range(60):
        o.step()
    xo = torch.as_tensor(o.field("position")).requires_grad_(True)
    loss(xo).backward()
    _, ov, omu = o.adjoint(xo.grad.numpy(), np.zeros((len(v0n), 3)))
    dirs = {"v0_scale": v0n.astype(np.float64), "v0_fourier": scenes.fourier_field(ro.x0, np.random.default_rng(100)),
            "v0_random": np.random.default_rng(0).normal(size=v0n.shape)}
    err_ref = {"mu": _rel(gmu, omu)}
    err_ref.update({k: _rel(float((gv * dv).sum()), float((ov * dv).sum())) for k, dv in dirs.items()})
    err_ref["v0_l2"] = float(np.linalg.norm(gv - ov) / np.linalg.norm(ov))
    print("c5 gpu vs fp64 reference adjoint", {k: "%.2e" % e for k, e in err_ref.items()}, flush=True)
    assert err_fd["v0_scale"] < 1e-3 and err_fd["mu"] < 5e-3, err_fd
    assert err_ref["mu"] < 1e-3 and err_ref["v0_scale"] < 1e-3 and err_ref["v0_fourier"] < 1e-2, err_ref
