"""Host-side mirror of VFD's DFSPH simulation component, on top of the C ABI (include/vfd.h).

The reference's interface is the C++ class DFSPHSimulation (VFD/Source/Simulation/DFSPH/DFSPHSimulator.h:10-56),
driven by the editor (Editor/Panels/ComponentPanel.cpp:608-667): construct it from a DFSPHSimulationDescription,
SetFluidObjects, SetRigidBodies, then Simulate (or OnUpdate per step) and read the particle frame buffer and debug
info. This module keeps those names and their meaning; errors raise VfdError instead of exiting the process.
"""
import numpy as np

from . import _abi
from ._abi import check, ptr


class FluidObject:
    """FluidObject (FluidObject/FluidObject.h:21-31): sampled positions and one initial velocity.

    `velocity` may also be an (n, 3) array of per-particle velocities (an extension)."""

    def __init__(self, positions, velocity=(0.0, 0.0, 0.0)):
        self.positions = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1, 3)
        v = np.asarray(velocity, dtype=np.float32)
        if v.ndim == 1:
            v = np.broadcast_to(v, self.positions.shape)
        self.velocities = np.ascontiguousarray(v, dtype=np.float32).reshape(-1, 3)
        if self.velocities.shape != self.positions.shape:
            raise ValueError("velocity must be a vec3 or one vec3 per position")

    @classmethod
    def sample(cls, desc, particle_radius):
        """FluidObject(desc, info) (FluidObject.cpp:6-24): the description's mesh sampled on the GPU."""
        x = sample_mesh_volume(desc.mesh, particle_radius, desc.resolution, desc.inverted, desc.sample_mode,
                               desc.transform)
        return cls(x, desc.velocity)

    def GetPositions(self):
        return self.positions

    def GetPositionCount(self):
        return len(self.positions)

    def GetVelocity(self):
        return self.velocities


class RigidBodyDescription:
    """RigidBodyDescription (RigidBody/RigidBody.cuh:15-24): Inverted, Padding, CollisionMapResolution and either a
    triangle mesh with its Transform (shape "mesh", `mesh` a TriangleMesh, `transform` a 4x4 matrix acting on
    (x, y, z, 1)) or an analytic shape ("box", "sphere", "plane", "cylinder") standing in for the mesh."""

    def __init__(self, shape, center=(0, 0, 0), half_extents=(0, 0, 0), radius=0.0, normal=(0, 1, 0),
                 inverted=False, padding=0.0, resolution=(20, 20, 20), domain=None, mesh=None, transform=None):
        self.shape = shape
        self.center = center
        self.half_extents = half_extents
        self.radius = radius
        self.normal = normal
        self.inverted = inverted
        self.padding = padding
        self.resolution = resolution
        self.domain = domain
        self.mesh = mesh
        self.transform = np.eye(4, dtype=np.float32) if transform is None else np.asarray(transform, np.float32)
        if shape == "mesh" and mesh is None:
            raise ValueError("a mesh body needs a mesh")

    @property
    def is_mesh(self):
        return self.shape == "mesh"

    def to_abi(self):
        if self.is_mesh:
            from .mesh import glm_transform
            b = _abi.MeshBoundaryDesc()
            b.Transform[:] = [float(v) for v in glm_transform(self.transform)]
            b.Inverted = 1 if self.inverted else 0
            b.Padding = self.padding
            b.CollisionMapResolution[:] = [int(r) for r in self.resolution]
            return b
        return _abi.make_boundary(self.shape, self.center, self.half_extents, self.radius, self.normal,
                                  self.inverted, self.padding, self.resolution, self.domain)


class FluidObjectDescription:
    """FluidObjectDescription (FluidObject/FluidObject.h:10-19): a triangle mesh with its Transform, Inverted,
    Resolution of the sampling SDF, SampleMode (0 min, 1 medium, 2 max density) and one initial Velocity."""

    def __init__(self, mesh, transform=None, inverted=False, resolution=(10, 10, 10), sample_mode=0,
                 velocity=(0.0, 0.0, 0.0)):
        self.mesh = mesh
        self.transform = np.eye(4, dtype=np.float32) if transform is None else np.asarray(transform, np.float32)
        self.inverted = inverted
        self.resolution = resolution
        self.sample_mode = sample_mode
        self.velocity = velocity


def sample_mesh_volume(mesh, radius, resolution, inverted=False, sample_mode=0, transform=None):
    """ParticleSampler::SampleMeshVolume (Utility/Sampler/ParticleSampler.cpp:7-91) of the mesh transformed by
    `transform` (FluidObject.cpp:9-14), on the GPU: (n, 3) float32 positions in the reference's loop order."""
    from .mesh import glm_transform
    lib = _abi.load_library()
    m = mesh.to_abi()
    tr = None if transform is None else glm_transform(transform)
    res = np.ascontiguousarray(resolution, dtype=np.uint32)
    n = _abi.ctypes.c_uint32()
    # room for every lattice candidate of the bounding box (which holds the origin), so that one call suffices; a
    # second call with the exact room covers an estimate that falls short
    from .mesh import transform_vertices
    v = mesh.vertices if transform is None else transform_vertices(transform, mesh.vertices)
    lo = np.minimum(v.min(axis=0), 0.0) if len(v) else np.zeros(3)
    hi = np.maximum(v.max(axis=0), 0.0) if len(v) else np.zeros(3)
    step = 2.0 * float(radius) * np.array([np.sqrt(3.0) / 2.0 if sample_mode == 2 else 1.0,
                                           np.sqrt(3.0) / 2.0 if sample_mode == 1 else
                                           (np.sqrt(6.0) / 3.0 if sample_mode == 2 else 1.0), 1.0])
    est = np.prod(np.floor((hi - lo) / step) + 2.0) if radius > 0 else 1.0
    cap = int(min(max(est, 1.0), float(1 << 24)))
    while True:
        out = np.zeros((cap, 3), dtype=np.float32)
        check(lib.vfd_sample_mesh_volume(_abi.ctypes.byref(m), ptr(tr), float(radius), ptr(res), int(bool(inverted)),
                                         int(sample_mode), ptr(out), cap, _abi.ctypes.byref(n)),
              "vfd_sample_mesh_volume")
        if n.value <= cap:
            return out[:n.value].copy()
        cap = n.value


def DFSPHSimulationDescription(**fields):
    """DFSPHSimulationDescription with the reference defaults (DFSPHSimulationDescription.h:11-49)."""
    return _abi.make_desc(**fields)


def read_descriptions(source):
    """Every DFSPH description in a VFD scene file (cereal JSON, DFSPHSimulationComponent.h:7-39, 57-70): `source`
    is a path or the JSON text. Fields a description does not name keep the reference defaults."""
    import os
    text = source
    if isinstance(source, (str, os.PathLike)) and os.path.exists(str(source)):
        with open(source, "rb") as f:
            text = f.read()
    data = text.encode() if isinstance(text, str) else bytes(text)
    lib = _abi.load_library()
    n = _abi.ctypes.c_uint32()
    check(lib.vfd_read_descriptions(data, len(data), None, 0, _abi.ctypes.byref(n)), "vfd_read_descriptions")
    arr = (_abi.Desc * max(n.value, 1))()
    check(lib.vfd_read_descriptions(data, len(data), arr, n.value, _abi.ctypes.byref(n)), "vfd_read_descriptions")
    return [arr[k] for k in range(n.value)]


def write_description(desc):
    """The description as a scene file stores it: {"description": {...}} (cereal names)."""
    lib = _abi.load_library()
    size = _abi.ctypes.c_size_t()
    check(lib.vfd_write_description(_abi.ctypes.byref(desc), None, 0, _abi.ctypes.byref(size)),
          "vfd_write_description")
    buf = _abi.ctypes.create_string_buffer(size.value + 1)
    check(lib.vfd_write_description(_abi.ctypes.byref(desc), buf, size.value + 1, _abi.ctypes.byref(size)),
          "vfd_write_description")
    return buf.value.decode()


class DFSPHParticleFrame:
    """DFSPHParticleFrame (ParticleBuffer/DFSPHParticleBuffer.h:13-19)."""

    def __init__(self, xva, max_velocity, time_step):
        self.ParticleData = xva  # (n, 9): position, velocity, acceleration (DFSPHParticleSimple)
        self.MaxVelocityMagnitude = max_velocity
        self.CurrentTimeStep = time_step


class DFSPHSimulation:
    """DFSPHSimulation (DFSPHSimulator.h:10-56) backed by the HIP library. `scene` selects one scene of a batch."""

    @classmethod
    def from_scene_file(cls, path, resource_root):
        """A VFD scene file baked as the editor's "Bake Objects" does (Editor/Panels/ComponentPanel.cpp:608-656): the
        simulation from the file's description, every fluid object sampled on the GPU (SetFluidObjects), then every
        mesh rigid body (SetRigidBodies). Mesh paths are resolved under `resource_root` (see
        vfd_amd.scenes.load_scene_file)."""
        from .scenes import load_scene_file
        desc, fluid_descs, bodies = load_scene_file(path, resource_root)
        sim = cls(desc)
        sim.SetFluidObjects([FluidObject.sample(fd, desc.ParticleRadius) for fd in fluid_descs])
        sim.SetRigidBodies(bodies)
        return sim

    def __init__(self, desc=None, _handle=None, _scene=0, _owner=None):
        self._lib = _abi.load_library()
        self._scene = _scene
        self._owner = _owner
        if _handle is not None:
            self._h = _handle
            return
        if desc is None:
            desc = DFSPHSimulationDescription()
        h = _abi.ctypes.c_void_p()
        check(self._lib.vfd_create(_abi.ctypes.byref(desc), _abi.ctypes.byref(h)), "vfd_create")
        self._h = h

    def close(self):
        if self._owner is None and getattr(self, "_h", None):
            self._lib.vfd_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- description / info ------------------------------------------------------------------------------------
    def SetDescription(self, desc):
        check(self._lib.vfd_batch_set_description(self._h, self._scene, _abi.ctypes.byref(desc)),
              "vfd_set_description")

    def GetDescription(self):
        d = _abi.Desc()
        check(self._lib.vfd_get_description(self._h, _abi.ctypes.byref(d)), "vfd_get_description")
        return d

    def GetInfo(self):
        i = _abi.Info()
        check(self._lib.vfd_batch_get_info(self._h, self._scene, _abi.ctypes.byref(i)), "vfd_get_info")
        return i

    # ---- set-up -------------------------------------------------------------------------------------------------
    def SetFluidObjects(self, fluid_objects):
        xs = [f.positions for f in fluid_objects]
        vs = [f.velocities for f in fluid_objects]
        x = np.ascontiguousarray(np.concatenate(xs) if xs else np.zeros((0, 3), np.float32), dtype=np.float32)
        v = np.ascontiguousarray(np.concatenate(vs) if vs else np.zeros((0, 3), np.float32), dtype=np.float32)
        self.set_fluid(x, v)

    def set_fluid(self, positions, velocities=None):
        x = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1, 3)
        v = None if velocities is None else np.ascontiguousarray(velocities, dtype=np.float32).reshape(-1, 3)
        check(self._lib.vfd_batch_set_fluid(self._h, self._scene, ptr(x), ptr(v), len(x)), "vfd_set_fluid")

    def SetRigidBodies(self, rigid_bodies):
        check(self._lib.vfd_clear_boundaries(self._h, self._scene), "vfd_clear_boundaries")
        for rb in rigid_bodies:
            if isinstance(rb, RigidBodyDescription) and rb.is_mesh:
                m = rb.mesh.to_abi()
                check(self._lib.vfd_batch_add_boundary_mesh(self._h, self._scene, _abi.ctypes.byref(rb.to_abi()),
                                                            _abi.ctypes.byref(m)), "vfd_add_boundary_mesh")
                continue
            b = rb.to_abi() if isinstance(rb, RigidBodyDescription) else rb
            check(self._lib.vfd_batch_add_boundary(self._h, self._scene, _abi.ctypes.byref(b)),
                  "vfd_add_boundary")

    # ---- stepping -----------------------------------------------------------------------------------------------
    def OnUpdate(self):
        check(self._lib.vfd_step(self._h), "vfd_step")

    def step_no_alloc(self):
        """OnUpdate without device allocation (no reference counterpart): the grid and curvature-table scratch come
        from the library's static scratch; needs frame capture off and the state already allocated."""
        check(self._lib.vfd_step_no_alloc(self._h), "vfd_step_no_alloc")

    def Simulate(self, asynchronous=False):
        """DFSPHImplementation::Simulate (DFSPHImplementation.cu:33-61): restore the baked particles and run until
        FrameCount frames are captured. asynchronous=True runs the frame loop on a worker thread of the library and
        returns at once, as the reference's DFSPHSimulation::Simulate does (DFSPHSimulator.cpp:156-166); poll with
        GetSimulationState / GetDebugInfo / GetParticleFrameBuffer and finish with Wait()."""
        if asynchronous:
            check(self._lib.vfd_simulate_async(self._h), "vfd_simulate_async")
        else:
            check(self._lib.vfd_simulate(self._h), "vfd_simulate")

    def Wait(self):
        """Join the worker of Simulate(asynchronous=True); raises VfdError with its status if the run failed."""
        check(self._lib.vfd_wait(self._h), "vfd_wait")

    def GetSimulationState(self):
        """0 None, 1 Simulating, 2 Ready (DFSPHImplementation::SimulationState, DFSPHImplementation.h:27-32)."""
        st = _abi.ctypes.c_int32()
        check(self._lib.vfd_get_simulation_state(self._h, _abi.ctypes.byref(st)), "vfd_get_simulation_state")
        return st.value

    def advance(self, steps):
        check(self._lib.vfd_advance(self._h, int(steps)), "vfd_advance")

    def step_until(self, phase):
        p = _abi.PHASES.index(phase) if isinstance(phase, str) else int(phase)
        check(self._lib.vfd_step_until(self._h, p), "vfd_step_until")

    def set_frame_capture(self, enabled):
        check(self._lib.vfd_set_frame_capture(self._h, 1 if enabled else 0), "vfd_set_frame_capture")

    def synchronize(self):
        pass  # every ABI call returns after its work is complete or enqueued behind a synchronising read

    # ---- getters ------------------------------------------------------------------------------------------------
    def GetParticleCount(self):
        n = _abi.ctypes.c_uint32()
        check(self._lib.vfd_get_particle_count(self._h, self._scene, _abi.ctypes.byref(n)), "vfd_get_particle_count")
        return n.value

    def GetParticleRadius(self):
        return self.GetInfo().ParticleRadius

    def GetDebugInfo(self):
        d = _abi.DebugInfo()
        check(self._lib.vfd_get_debug_info(self._h, self._scene, _abi.ctypes.byref(d)), "vfd_get_debug_info")
        return d

    def GetMaxVelocityMagnitude(self):
        return self.GetDebugInfo().MaxVelocityMagnitude

    def GetCurrentTimeStepSize(self):
        return self.GetDebugInfo().TimeStepSize

    def GetRigidBodyCount(self):
        return self.GetInfo().RigidBodyCount

    def GetKernel(self):
        w = np.zeros(10000, np.float32)
        g = np.zeros(10001, np.float32)
        s = np.zeros(6, np.float32)
        check(self._lib.vfd_get_kernel_table(self._h, ptr(w), ptr(g), ptr(s)), "vfd_get_kernel_table")
        return w, g, dict(zip(["Radius", "Radius2", "InvStepSize", "WZero", "K", "L"], s.tolist()))

    def field(self, name, body=0):
        fid, comps = _abi.FIELDS[name]
        n = self.GetParticleCount()
        out = np.zeros((n, comps), np.float32)
        check(self._lib.vfd_get_field(self._h, self._scene, fid, body, ptr(out)), "vfd_get_field(%s)" % name)
        return out if comps > 1 else out[:, 0]

    def set_field(self, name, values):
        fid, comps = _abi.FIELDS[name]
        a = np.ascontiguousarray(values, dtype=np.float32).reshape(-1)
        check(self._lib.vfd_set_field(self._h, self._scene, fid, ptr(a)), "vfd_set_field(%s)" % name)

    def particles(self):
        n = self.GetParticleCount()
        x = np.zeros((n, 3), np.float32)
        v = np.zeros((n, 3), np.float32)
        a = np.zeros((n, 3), np.float32)
        check(self._lib.vfd_batch_get_particles(self._h, self._scene, ptr(x), ptr(v), ptr(a)), "vfd_get_particles")
        return x, v, a

    def neighbors(self):
        n = self.GetParticleCount()
        counts = np.zeros(n, np.uint32)
        offsets = np.zeros(n, np.uint32)
        total = _abi.ctypes.c_uint64()
        check(self._lib.vfd_get_neighbors(self._h, self._scene, ptr(counts), ptr(offsets), None,
                                          _abi.ctypes.byref(total)), "vfd_get_neighbors")
        ids = np.zeros(max(total.value, 1), np.uint32)
        check(self._lib.vfd_get_neighbors(self._h, self._scene, ptr(counts), ptr(offsets), ptr(ids),
                                          _abi.ctypes.byref(total)), "vfd_get_neighbors")
        return counts, offsets, ids[:total.value]

    def boundary_map(self, body, field):
        cnt = _abi.ctypes.c_uint32()
        check(self._lib.vfd_get_boundary_map(self._h, self._scene, body, field, None, _abi.ctypes.byref(cnt)),
              "vfd_get_boundary_map")
        out = np.zeros(cnt.value, np.float32)
        check(self._lib.vfd_get_boundary_map(self._h, self._scene, body, field, ptr(out), _abi.ctypes.byref(cnt)),
              "vfd_get_boundary_map")
        return out

    def set_profiling(self, enabled):
        check(self._lib.vfd_set_profiling(self._h, 1 if enabled else 0), "vfd_set_profiling")

    def profile(self):
        """{kernel: (launches, total_ms)} accumulated since profiling was enabled."""
        n = _abi.ctypes.c_uint32()
        check(self._lib.vfd_get_profile(self._h, None, 0, _abi.ctypes.byref(n)), "vfd_get_profile")
        arr = (_abi.KernelStat * max(n.value, 1))()
        check(self._lib.vfd_get_profile(self._h, arr, n.value, _abi.ctypes.byref(n)), "vfd_get_profile")
        return {arr[i].name.decode(): (arr[i].launches, arr[i].total_ms) for i in range(n.value)}

    # ---- checkpoints (no reference counterpart: the reference only keeps per-frame snapshots for playback,
    #      DFSPHParticleBuffer.h:13-45, and Simulate restarts from the baked particles, DFSPHImplementation.cu:36) ---
    def checkpoint(self):
        """Everything a step reads from the previous one, in original particle order: positions, velocities, the
        viscosity warm start (velocity_difference), the surface-tension curvature blend and the time step."""
        return {"x": self.field("position"), "v": self.field("velocity"), "dv": self.field("velocity_difference"),
                "dk": self.field("delta_final_curvature"), "dt": self.GetDebugInfo().TimeStepSize}

    def restore(self, cp):
        """Put a checkpoint back. The next step from it repeats, bit for bit, the step that followed it when it was
        taken (same description and boundaries)."""
        d = self.GetDescription()
        d.TimeStepSize = cp["dt"]
        self.SetDescription(d)  # SetDescription resets the time step to desc.TimeStepSize (:352-356)
        self.set_field("position", cp["x"])
        self.set_field("velocity", cp["v"])
        self.set_field("velocity_difference", cp["dv"])
        self.set_field("delta_final_curvature", cp["dk"])

    # ---- adjoint (no reference counterpart) ------------------------------------------------------------------
    def tape(self, on):
        """Start (clearing earlier records) or stop recording steps for adjoint()."""
        check(self._lib.vfd_tape(self._h, 1 if on else 0), "vfd_tape")

    def taped_steps(self):
        n = _abi.ctypes.c_uint32()
        check(self._lib.vfd_tape_steps(self._h, _abi.ctypes.byref(n)), "vfd_tape_steps")
        return n.value

    def adjoint(self, lx_out, lv_out=None):
        """Cotangents of the last recorded step's positions/velocities (original order) -> those of the first
        recorded step's input, and d/dmu (mu = dynamic viscosity). For a scene of a batch, that scene's particles
        and viscosity (vfd_batch_adjoint)."""
        n = self.GetParticleCount()
        lx = np.ascontiguousarray(lx_out, np.float32).reshape(n, 3)
        lv = None if lv_out is None else np.ascontiguousarray(lv_out, np.float32).reshape(n, 3)
        ox, ov = np.zeros((n, 3), np.float32), np.zeros((n, 3), np.float32)
        mu = _abi.ctypes.c_double()
        args = (ptr(lx), None if lv is None else ptr(lv), ptr(ox), ptr(ov), _abi.ctypes.byref(mu))
        if self._scene == 0:
            check(self._lib.vfd_adjoint(self._h, *args), "vfd_adjoint")
        else:
            check(self._lib.vfd_batch_adjoint(self._h, self._scene, *args), "vfd_batch_adjoint")
        return ox, ov, mu.value

    def tuning(self):
        """{SolverCheckInterval, HaloLimit, UseGraphs, PredictIterations, AdjointGroupLoops}: launch and solver-control
        knobs (results do not depend on them)."""
        t = _abi.Tuning()
        check(self._lib.vfd_get_tuning(self._h, _abi.ctypes.byref(t)), "vfd_get_tuning")
        return {name: getattr(t, name) for name, _ in _abi.Tuning._fields_}

    def set_tuning(self, **kw):
        t = _abi.Tuning(**{**self.tuning(), **kw})
        check(self._lib.vfd_set_tuning(self._h, _abi.ctypes.byref(t)), "vfd_set_tuning")

    def GetParticleFrameBuffer(self):
        cnt = _abi.ctypes.c_uint32()
        check(self._lib.vfd_get_frame_count(self._h, self._scene, _abi.ctypes.byref(cnt)), "vfd_get_frame_count")
        n = self.GetParticleCount()
        frames = []
        for k in range(cnt.value):
            xva = np.zeros((n, 9), np.float32)
            mv = _abi.ctypes.c_float()
            dt = _abi.ctypes.c_float()
            check(self._lib.vfd_get_frame(self._h, self._scene, k, ptr(xva), _abi.ctypes.byref(mv),
                                          _abi.ctypes.byref(dt)), "vfd_get_frame")
            frames.append(DFSPHParticleFrame(xva, mv.value, dt.value))
        return frames


class DFSPHBatch:
    """Many independent scenes advanced by one set of launches (no reference counterpart)."""

    def __init__(self, descs):
        self._lib = _abi.load_library()
        arr = (_abi.Desc * len(descs))(*descs)
        h = _abi.ctypes.c_void_p()
        check(self._lib.vfd_batch_create(arr, len(descs), _abi.ctypes.byref(h)), "vfd_batch_create")
        self._h = h
        self.scenes = [DFSPHSimulation(_handle=h, _scene=i, _owner=self) for i in range(len(descs))]

    def step(self):
        check(self._lib.vfd_step(self._h), "vfd_step")

    def step_no_alloc(self):
        check(self._lib.vfd_step_no_alloc(self._h), "vfd_step_no_alloc")

    def advance(self, steps):
        check(self._lib.vfd_advance(self._h, int(steps)), "vfd_advance")

    def set_frame_capture(self, enabled):
        check(self._lib.vfd_set_frame_capture(self._h, 1 if enabled else 0), "vfd_set_frame_capture")

    def tape(self, on):
        """Start (clearing earlier records) or stop recording the batch's steps for adjoint()."""
        self.scenes[0].tape(on)

    def taped_steps(self):
        return self.scenes[0].taped_steps()

    def adjoint(self, scene, lx_out, lv_out=None):
        """Scene `scene`'s cotangents through the recorded steps and its d/dmu (DFSPHSimulation.adjoint)."""
        return self.scenes[scene].adjoint(lx_out, lv_out)

    def set_profiling(self, enabled):
        self.scenes[0].set_profiling(enabled)

    def profile(self):
        return self.scenes[0].profile()

    def tuning(self):
        return self.scenes[0].tuning()

    def set_tuning(self, **kw):
        self.scenes[0].set_tuning(**kw)

    def close(self):
        if getattr(self, "_h", None):
            self._lib.vfd_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
