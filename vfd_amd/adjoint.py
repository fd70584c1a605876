"""Differentiable DFSPH rollouts: a torch.autograd.Function over vfd_tape / vfd_adjoint (include/vfd.h), with the
state checkpointed every few steps (no reference counterpart).

`Rollout(desc, x0, bodies, steps, every)` is the map (x0, v0, mu) -> (x, v) after `steps` steps of one scene, mu being
the dynamic viscosity (desc.Viscosity = mu / rho0, rho0 = 1000 as in DFSPHImplementation.cu:17). The forward pass
keeps a checkpoint (DFSPHSimulation.checkpoint) at the start of every segment of `every` steps. The backward pass
takes the segments last first: it restores the segment's checkpoint, runs the segment again with recording on (the
re-run repeats the forward bit for bit) and reverses the recorded steps on the GPU (vfd_adjoint), which carries the
cotangents of positions and velocities back through the segment and adds up d/dmu.

What is differentiated is the step as run: neighbour lists, time steps, solver iteration counts and clamp masks are
held at their recorded values; the viscosity solve is differentiated as if solved exactly. The recorded steps must
stay clear of boundaries and surface tension (vfd_adjoint refuses them).

`splat_density` is the loss-side helper of the viscosity-inversion workload: particle mass spread onto a regular grid.
"""
import itertools
import time

import numpy as np
import torch

from .simulation import DFSPHSimulation

RHO0 = 1000.0


def splat_density(x, mass, lo, hi, res=64):
    """Density on a res^3 grid over the box [lo, hi]: each particle's mass spread over the 8 cells whose centres
    surround it with trilinear (cloud-in-cell) weights, divided by the cell volume. Differentiable in x (an (n, 3)
    tensor); cells outside the grid are dropped. Returns a (res, res, res) tensor indexed [ix, iy, iz]."""
    lo = torch.as_tensor(lo, dtype=x.dtype, device=x.device)
    hi = torch.as_tensor(hi, dtype=x.dtype, device=x.device)
    h = (hi - lo) / res
    g = (x - lo) / h - 0.5
    base = torch.floor(g.detach())
    f = g - base
    base = base.long()
    out = torch.zeros(res ** 3, dtype=x.dtype, device=x.device)
    for corner in itertools.product((0, 1), repeat=3):
        c = torch.tensor(corner, device=x.device)
        idx = base + c
        w = torch.where(c.bool(), f, 1.0 - f).prod(dim=1) * mass
        ok = ((idx >= 0) & (idx < res)).all(dim=1)
        flat = (idx[:, 0] * res + idx[:, 1]) * res + idx[:, 2]
        out = out.index_add(0, flat[ok], w[ok])
    return out.view(res, res, res) / h.prod()


class Rollout:
    """`steps` DFSPH steps of one scene from (x0, v0) at dynamic viscosity mu, as a differentiable function.

    Call it with torch tensors: `x, v = rollout(v0, mu)` (x0 fixed to the constructor's) or `rollout(v0, mu, x0)`;
    gradients flow to v0, mu and x0. The simulation runs on the GPU through the HIP library; tensors may live on any
    device (they cross the C ABI as host arrays). `timings` holds the wall time of the last forward and backward."""

    def __init__(self, desc, x0, bodies=(), steps=60, every=10):
        if steps < 1 or every < 1:
            raise ValueError("steps and every must be positive")
        self.desc = desc
        self.x0 = np.ascontiguousarray(x0, np.float32).reshape(-1, 3)
        self.steps = int(steps)
        self.every = int(every)
        self.sim = DFSPHSimulation(desc)
        self.sim.SetRigidBodies(list(bodies))
        self.checkpoints = []
        self._last = None
        self.mu = None
        self.timings = {}

    def __call__(self, v0, mu, x0=None):
        if x0 is None:
            x0 = torch.as_tensor(self.x0, device=v0.device)
        mu = torch.as_tensor(mu, dtype=torch.float64, device=v0.device)
        return _RolloutFn.apply(x0, v0, mu, self)

    def _set_mu(self, mu):
        d = self.sim.GetDescription()
        d.Viscosity = float(mu) / RHO0
        d.BoundaryViscosity = float(mu) / RHO0
        d.TimeStepSize = self.desc.TimeStepSize
        self.sim.SetDescription(d)
        self.mu = float(np.float32(d.Viscosity)) * RHO0  # the viscosity the steps actually use

    def forward_np(self, x0, v0, mu):
        """Run the rollout (host arrays in, host arrays out); backward_np() without a record then differentiates
        this run."""
        x, v, _ = self.run_np(x0, v0, mu)
        return x, v

    def run_np(self, x0, v0, mu):
        """Run the rollout and return (x, v, record). `record` holds what the backward pass of this run needs: its
        checkpoints and its viscosity. The autograd function keeps the record of each call, so that a second forward
        before the first backward (gradcheck, a line search, a loss over two rollouts) leaves the first call's
        gradient unchanged."""
        t0 = time.perf_counter()
        self._set_mu(mu)
        self.sim.set_fluid(x0, v0)
        checkpoints = []
        for k in range(self.steps):
            if k % self.every == 0:
                checkpoints.append(self.sim.checkpoint())
            self.sim.OnUpdate()
        x, v = self.sim.field("position"), self.sim.field("velocity")
        self.timings["forward_s"] = time.perf_counter() - t0
        record = {"checkpoints": checkpoints, "mu": float(mu)}
        self.checkpoints = checkpoints
        self._last = record
        return x, v, record

    def backward_np(self, lx, lv, record=None):
        """Cotangents of the final (x, v) of the run that produced `record` (default: the last run) -> those of
        (x0, v0) and d/dmu, segment by segment from the last. The segments are re-run at that run's viscosity."""
        t0 = time.perf_counter()
        record = record if record is not None else getattr(self, "_last", None)
        if not record or not record["checkpoints"]:
            raise RuntimeError("backward before forward")
        checkpoints = record["checkpoints"]
        self._set_mu(record["mu"])
        lx = np.ascontiguousarray(lx, np.float32)
        lv = np.ascontiguousarray(lv, np.float32)
        lmu, t_rev = 0.0, 0.0
        for s in reversed(range(len(checkpoints))):
            self.sim.restore(checkpoints[s])
            self.sim.tape(True)
            for _ in range(min(self.every, self.steps - s * self.every)):
                self.sim.OnUpdate()
            self.sim.tape(False)
            t1 = time.perf_counter()
            lx, lv, m = self.sim.adjoint(lx, lv)
            t_rev += time.perf_counter() - t1
            lmu += m
        self.timings["backward_s"] = time.perf_counter() - t0
        self.timings["reverse_s"] = t_rev  # the part of backward_s spent in vfd_adjoint (the rest re-runs segments)
        return lx, lv, lmu


class _RolloutFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x0, v0, mu, rollout):
        x, v, record = rollout.run_np(x0.detach().cpu().numpy(), v0.detach().cpu().numpy(), float(mu))
        ctx.rollout = rollout
        ctx.record = record  # this call's checkpoints and viscosity
        ctx.meta = (x0.device, x0.dtype, v0.device, v0.dtype, mu.device, mu.dtype)
        return torch.as_tensor(x, device=v0.device), torch.as_tensor(v, device=v0.device)

    @staticmethod
    def backward(ctx, gx, gv):
        xd, xt, vd, vt, md, mt = ctx.meta
        lx, lv, lmu = ctx.rollout.backward_np(gx.detach().cpu().numpy(), gv.detach().cpu().numpy(), ctx.record)
        return (torch.as_tensor(lx, device=xd, dtype=xt), torch.as_tensor(lv, device=vd, dtype=vt),
                torch.as_tensor(lmu, device=md, dtype=mt), None)


class DensityLoss:
    """Squared error of the splatted density field against a target field: sum over cells of (rho - rho*)^2."""

    def __init__(self, target_x, mass, lo=(0.0, 0.0, 0.0), hi=(2.0, 2.0, 2.0), res=64):
        self.mass, self.lo, self.hi, self.res = float(mass), lo, hi, res
        t = torch.as_tensor(np.asarray(target_x, np.float64))
        self.target = splat_density(t, self.mass, lo, hi, res)

    def __call__(self, x):
        rho = splat_density(x.to(torch.float64).cpu(), self.mass, self.lo, self.hi, self.res)
        return ((rho - self.target) ** 2).sum()


def viscosity_inversion(n_side=(50, 40, 50), steps=60, every=10, mu_target=20.0, seed=5):
    """The viscosity / initial-velocity inversion workload: a falling viscous block (the differentiable form of
    scenes.falling_viscous_block), a target run at mu_target, and the loss of a run against it. Returns
    (rollout, v0, loss): `loss(rollout(v0, mu)[0])` is the density-field error after `steps` steps."""
    from . import scenes
    descs, fluids, bodies = scenes.falling_viscous_block(seed=seed, n_side=n_side, mu=mu_target, differentiable=True)
    x0, v0 = fluids[0]
    ro = Rollout(descs[0], x0, bodies[0], steps=steps, every=every)
    xt, _ = ro.forward_np(x0, v0, mu_target)
    loss = DensityLoss(xt, ro.sim.GetInfo().ParticleMass)
    return ro, torch.as_tensor(v0), loss
