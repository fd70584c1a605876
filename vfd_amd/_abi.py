"""ctypes mirror of include/vfd.h (structs, enums, prototypes) and the loader of the in-tree HIP library.

The library is loaded from vfd_amd/libvfd_hip.so (built by vfd_amd/build.py). There is no fallback: if the
library is missing or fails to load, load_library() raises.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# VFD_LIBRARY selects another build of the library (the autotuner points its trial processes at candidate builds)
LIB_PATH = os.environ.get("VFD_LIBRARY") or os.path.join(HERE, "libvfd_hip.so")

c_float3 = ctypes.c_float * 3
c_uint3 = ctypes.c_uint32 * 3

VFD_OK = 0
STATUS_NAMES = {0: "OK", 1: "INVALID_ARGUMENT", 2: "HIP", 3: "STATE", 4: "OUT_OF_MEMORY", 5: "NOT_SUPPORTED",
                6: "NONFINITE"}

SHAPE_BOX, SHAPE_SPHERE, SHAPE_PLANE, SHAPE_CYLINDER = 0, 1, 2, 3

# DFSPHImplementation::SimulationState (DFSPHImplementation.h:27-32)
STATE_NONE, STATE_SIMULATING, STATE_READY = 0, 1, 2
STATE_NAMES = {STATE_NONE: "None", STATE_SIMULATING: "Simulating", STATE_READY: "Ready"}

FIELDS = {
    "position": (0, 3), "velocity": (1, 3), "acceleration": (2, 3), "pressure_acceleration": (3, 3),
    "pressure_residuum": (4, 1), "density": (5, 1), "density_advection": (6, 1), "pressure_rho2": (7, 1),
    "pressure_rho2_v": (8, 1), "factor": (9, 1), "velocity_difference": (10, 3), "mc_surface_normal": (11, 3),
    "mc_surface_normal_smooth": (12, 3), "mc_surface_curvature": (13, 1), "mc_surface_curvature_smooth": (14, 1),
    "delta_final_curvature": (15, 1), "boundary_xj": (16, 3), "boundary_volume": (17, 1),
    "viscosity_solution": (18, 3), "viscosity_preconditioner": (19, 9),
}

PHASES = ["search", "boundary", "density", "factor", "divergence", "clear_acceleration", "surface_tension",
          "viscosity", "time_step", "velocity", "pressure", "position"]


class Desc(ctypes.Structure):
    """vfd_desc == DFSPHSimulationDescription (Structures/DFSPHSimulationDescription.h:9-50)."""
    _fields_ = [
        ("TimeStepSize", ctypes.c_float), ("MinTimeStepSize", ctypes.c_float), ("MaxTimeStepSize", ctypes.c_float),
        ("FrameLength", ctypes.c_float), ("FrameCount", ctypes.c_uint32),
        ("MinPressureSolverIterations", ctypes.c_uint32), ("MaxPressureSolverIterations", ctypes.c_uint32),
        ("MaxPressureSolverError", ctypes.c_float),
        ("EnableDivergenceSolverError", ctypes.c_int32), ("MinDivergenceSolverIterations", ctypes.c_uint32),
        ("MaxDivergenceSolverIterations", ctypes.c_uint32), ("MaxDivergenceSolverError", ctypes.c_float),
        ("EnableViscositySolver", ctypes.c_int32), ("MinViscositySolverIterations", ctypes.c_uint32),
        ("MaxViscositySolverIterations", ctypes.c_uint32), ("MaxViscositySolverError", ctypes.c_float),
        ("Viscosity", ctypes.c_float), ("BoundaryViscosity", ctypes.c_float),
        ("TangentialDistanceFactor", ctypes.c_float),
        ("EnableSurfaceTensionSolver", ctypes.c_int32), ("SurfaceTensionSmoothPassCount", ctypes.c_uint32),
        ("SurfaceTension", ctypes.c_float), ("TemporalSmoothing", ctypes.c_int32), ("CSDFix", ctypes.c_int32),
        ("CSD", ctypes.c_int32), ("ParticleRadius", ctypes.c_float), ("Gravity", c_float3),
    ]


# Reference defaults (DFSPHSimulationDescription.h:11-49)
DESC_DEFAULTS = dict(
    TimeStepSize=0.001, MinTimeStepSize=0.0001, MaxTimeStepSize=0.005, FrameLength=0.0016, FrameCount=200,
    MinPressureSolverIterations=0, MaxPressureSolverIterations=100, MaxPressureSolverError=10.0,
    EnableDivergenceSolverError=1, MinDivergenceSolverIterations=0, MaxDivergenceSolverIterations=100,
    MaxDivergenceSolverError=10.0, EnableViscositySolver=1, MinViscositySolverIterations=0,
    MaxViscositySolverIterations=100, MaxViscositySolverError=0.1, Viscosity=10.0, BoundaryViscosity=10.0,
    TangentialDistanceFactor=0.5, EnableSurfaceTensionSolver=1, SurfaceTensionSmoothPassCount=1,
    SurfaceTension=1.0, TemporalSmoothing=0, CSDFix=-1, CSD=10000, ParticleRadius=0.025,
    Gravity=(0.0, -9.81, 0.0),
)


def make_desc(**kw):
    vals = dict(DESC_DEFAULTS)
    for k, v in kw.items():
        if k not in vals:
            raise KeyError("unknown description field %r" % k)
        vals[k] = v
    d = Desc()
    for k, v in vals.items():
        if k == "Gravity":
            d.Gravity = c_float3(*v)
        elif isinstance(getattr(d, k), int):
            setattr(d, k, int(v))
        else:
            setattr(d, k, float(v))
    return d


def desc_to_dict(d):
    out = {}
    for name, _ in Desc._fields_:
        v = getattr(d, name)
        out[name] = tuple(v) if name == "Gravity" else v
    return out


class Info(ctypes.Structure):
    """vfd_info == DFSPHSimulationInfo (Structures/DFSPHSimulationInfo.h:5-44)."""
    _fields_ = [
        ("ParticleCount", ctypes.c_uint32), ("RigidBodyCount", ctypes.c_uint32),
        ("SupportRadius", ctypes.c_float), ("SupportRadius2", ctypes.c_float), ("ParticleRadius", ctypes.c_float),
        ("ParticleDiameter", ctypes.c_float), ("TimeStepSize", ctypes.c_float), ("TimeStepSize2", ctypes.c_float),
        ("TimeStepSizeInverse", ctypes.c_float), ("TimeStepSize2Inverse", ctypes.c_float),
        ("Volume", ctypes.c_float), ("Density0", ctypes.c_float), ("ParticleMass", ctypes.c_float),
        ("ParticleMassInverse", ctypes.c_float), ("Viscosity", ctypes.c_float),
        ("BoundaryViscosity", ctypes.c_float), ("DynamicViscosity", ctypes.c_float),
        ("DynamicBoundaryViscosity", ctypes.c_float), ("TangentialDistanceFactor", ctypes.c_float),
        ("TangentialDistance", ctypes.c_float), ("SurfaceTension", ctypes.c_float),
        ("SurfaceTensionSampleCount", ctypes.c_uint32), ("ClassifierSlope", ctypes.c_float),
        ("ClassifierConstant", ctypes.c_float), ("TemporalSmoothing", ctypes.c_int32),
        ("SmoothingFactor", ctypes.c_float), ("Factor", ctypes.c_float), ("NeighborParticleRadius", ctypes.c_float),
        ("MonteCarloFactor", ctypes.c_float), ("Gravity", c_float3),
    ]


class DebugInfo(ctypes.Structure):
    """vfd_debug_info == DFSPHDebugInfo (Structures/DFSPHDebugInfo.h:8-28) plus seven additions."""
    _fields_ = [
        ("IterationCount", ctypes.c_uint32), ("DivergenceSolverIterationCount", ctypes.c_uint32),
        ("PressureSolverIterationCount", ctypes.c_uint32), ("ViscositySolverIterationCount", ctypes.c_uint32),
        ("DivergenceSolverError", ctypes.c_float), ("PressureSolverError", ctypes.c_float),
        ("ViscositySolverError", ctypes.c_float), ("FrameTime", ctypes.c_float), ("FrameIndex", ctypes.c_uint32),
        ("NeighborhoodSearchTime", ctypes.c_float), ("BaseSolverTime", ctypes.c_float),
        ("DivergenceSolverTime", ctypes.c_float), ("SurfaceTensionSolverTime", ctypes.c_float),
        ("ViscositySolverTime", ctypes.c_float), ("PressureSolverTime", ctypes.c_float),
        ("TimeStepSize", ctypes.c_float), ("MaxVelocityMagnitude", ctypes.c_float),
        ("MaxNeighborCount", ctypes.c_uint32), ("TotalNeighborCount", ctypes.c_uint64),
        ("StagedGroupCount", ctypes.c_uint32), ("GatheredGroupCount", ctypes.c_uint32),
        ("MaxHaloSize", ctypes.c_uint32),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class BoundaryDesc(ctypes.Structure):
    """vfd_boundary_desc == RigidBodyDescription (RigidBody/RigidBody.cuh:12-21) with an analytic shape."""
    _fields_ = [
        ("Shape", ctypes.c_int32), ("Center", c_float3), ("HalfExtents", c_float3), ("Radius", ctypes.c_float),
        ("Normal", c_float3), ("Inverted", ctypes.c_int32), ("Padding", ctypes.c_float),
        ("CollisionMapResolution", c_uint3), ("HasDomain", ctypes.c_int32), ("DomainMin", c_float3),
        ("DomainMax", c_float3),
    ]


class Mesh(ctypes.Structure):
    """vfd_mesh: TriangleMesh::GetVertices / GetTriangles (Renderer/Mesh/TriangleMesh.cpp:39-128)."""
    _fields_ = [("Vertices", ctypes.c_void_p), ("VertexCount", ctypes.c_uint32), ("Triangles", ctypes.c_void_p),
                ("TriangleCount", ctypes.c_uint32)]


class MeshBoundaryDesc(ctypes.Structure):
    """vfd_mesh_boundary_desc == RigidBodyDescription (RigidBody/RigidBody.cuh:15-24) without its mesh pointer."""
    _fields_ = [("Transform", ctypes.c_float * 16), ("Inverted", ctypes.c_int32), ("Padding", ctypes.c_float),
                ("CollisionMapResolution", c_uint3)]


SAMPLE_MIN_DENSITY, SAMPLE_MEDIUM_DENSITY, SAMPLE_MAX_DENSITY = 0, 1, 2


def make_boundary(shape, center=(0, 0, 0), half_extents=(0, 0, 0), radius=0.0, normal=(0, 1, 0), inverted=False,
                  padding=0.0, resolution=(20, 20, 20), domain=None):
    b = BoundaryDesc()
    b.Shape = {"box": SHAPE_BOX, "sphere": SHAPE_SPHERE, "plane": SHAPE_PLANE, "cylinder": SHAPE_CYLINDER}.get(
        shape, shape)
    b.Center = c_float3(*center)
    b.HalfExtents = c_float3(*half_extents)
    b.Radius = float(radius)
    b.Normal = c_float3(*normal)
    b.Inverted = 1 if inverted else 0
    b.Padding = float(padding)
    b.CollisionMapResolution = c_uint3(*[int(r) for r in resolution])
    if domain is not None:
        b.HasDomain = 1
        b.DomainMin = c_float3(*domain[0])
        b.DomainMax = c_float3(*domain[1])
    return b


class KernelStat(ctypes.Structure):
    """vfd_kernel_stat: per-kernel launch count and summed HIP-event time."""
    _fields_ = [("name", ctypes.c_char * 48), ("launches", ctypes.c_uint32), ("pad", ctypes.c_uint32),
                ("total_ms", ctypes.c_double)]


class Tuning(ctypes.Structure):
    """vfd_tuning: launch and solver-control knobs (no reference counterpart)."""
    _fields_ = [("SolverCheckInterval", ctypes.c_uint32), ("HaloLimit", ctypes.c_uint32),
                ("UseGraphs", ctypes.c_uint32), ("PredictIterations", ctypes.c_uint32),
                ("AdjointGroupLoops", ctypes.c_uint32)]


_P = ctypes.c_void_p
_U32 = ctypes.c_uint32
_I32 = ctypes.c_int32
_PROTOS = {
    "vfd_abi_version": (ctypes.c_int, []),
    "vfd_last_error": (ctypes.c_char_p, []),
    "vfd_desc_default": (ctypes.c_int, [ctypes.POINTER(Desc)]),
    "vfd_read_descriptions": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(Desc), _U32,
                                             ctypes.POINTER(_U32)]),
    "vfd_write_description": (ctypes.c_int, [ctypes.POINTER(Desc), ctypes.c_char_p, ctypes.c_size_t,
                                             ctypes.POINTER(ctypes.c_size_t)]),
    "vfd_create": (ctypes.c_int, [ctypes.POINTER(Desc), ctypes.POINTER(_P)]),
    "vfd_batch_create": (ctypes.c_int, [ctypes.POINTER(Desc), _U32, ctypes.POINTER(_P)]),
    "vfd_destroy": (ctypes.c_int, [_P]),
    "vfd_set_stream": (ctypes.c_int, [_P, _P]),
    "vfd_scene_count": (ctypes.c_int, [_P, ctypes.POINTER(_U32)]),
    "vfd_set_description": (ctypes.c_int, [_P, ctypes.POINTER(Desc)]),
    "vfd_get_description": (ctypes.c_int, [_P, ctypes.POINTER(Desc)]),
    "vfd_get_info": (ctypes.c_int, [_P, ctypes.POINTER(Info)]),
    "vfd_batch_set_description": (ctypes.c_int, [_P, _U32, ctypes.POINTER(Desc)]),
    "vfd_batch_get_info": (ctypes.c_int, [_P, _U32, ctypes.POINTER(Info)]),
    "vfd_set_fluid": (ctypes.c_int, [_P, _P, _P, _U32]),
    "vfd_batch_set_fluid": (ctypes.c_int, [_P, _U32, _P, _P, _U32]),
    "vfd_add_boundary": (ctypes.c_int, [_P, ctypes.POINTER(BoundaryDesc)]),
    "vfd_batch_add_boundary": (ctypes.c_int, [_P, _U32, ctypes.POINTER(BoundaryDesc)]),
    "vfd_add_boundary_mesh": (ctypes.c_int, [_P, ctypes.POINTER(MeshBoundaryDesc), ctypes.POINTER(Mesh)]),
    "vfd_batch_add_boundary_mesh": (ctypes.c_int, [_P, _U32, ctypes.POINTER(MeshBoundaryDesc),
                                                   ctypes.POINTER(Mesh)]),
    "vfd_sample_mesh_volume": (ctypes.c_int, [ctypes.POINTER(Mesh), _P, ctypes.c_float, _P, ctypes.c_int32,
                                              ctypes.c_int32, _P, _U32, ctypes.POINTER(_U32)]),
    "vfd_clear_boundaries": (ctypes.c_int, [_P, _U32]),
    "vfd_get_boundary_map": (ctypes.c_int, [_P, _U32, _U32, _U32, _P, ctypes.POINTER(_U32)]),
    "vfd_step": (ctypes.c_int, [_P]),
    "vfd_step_no_alloc": (ctypes.c_int, [_P]),
    "vfd_advance": (ctypes.c_int, [_P, _U32]),
    "vfd_simulate": (ctypes.c_int, [_P]),
    "vfd_simulate_async": (ctypes.c_int, [_P]),
    "vfd_wait": (ctypes.c_int, [_P]),
    "vfd_get_simulation_state": (ctypes.c_int, [_P, ctypes.POINTER(_I32)]),
    "vfd_step_until": (ctypes.c_int, [_P, _I32]),
    "vfd_set_frame_capture": (ctypes.c_int, [_P, _I32]),
    "vfd_get_particles": (ctypes.c_int, [_P, _P, _P, _P]),
    "vfd_batch_get_particles": (ctypes.c_int, [_P, _U32, _P, _P, _P]),
    "vfd_get_field": (ctypes.c_int, [_P, _U32, _I32, _U32, _P]),
    "vfd_set_field": (ctypes.c_int, [_P, _U32, _I32, _P]),
    "vfd_get_particle_count": (ctypes.c_int, [_P, _U32, ctypes.POINTER(_U32)]),
    "vfd_get_neighbors": (ctypes.c_int, [_P, _U32, _P, _P, _P, ctypes.POINTER(ctypes.c_uint64)]),
    "vfd_get_debug_info": (ctypes.c_int, [_P, _U32, ctypes.POINTER(DebugInfo)]),
    "vfd_get_frame_count": (ctypes.c_int, [_P, _U32, ctypes.POINTER(_U32)]),
    "vfd_get_frame": (ctypes.c_int, [_P, _U32, _U32, _P, ctypes.POINTER(ctypes.c_float),
                                     ctypes.POINTER(ctypes.c_float)]),
    "vfd_get_kernel_table": (ctypes.c_int, [_P, _P, _P, _P]),
    "vfd_set_profiling": (ctypes.c_int, [_P, _I32]),
    "vfd_get_profile": (ctypes.c_int, [_P, ctypes.POINTER(KernelStat), _U32, ctypes.POINTER(_U32)]),
    "vfd_tape": (ctypes.c_int, [_P, _I32]),
    "vfd_tape_steps": (ctypes.c_int, [_P, ctypes.POINTER(_U32)]),
    "vfd_adjoint": (ctypes.c_int, [_P, _P, _P, _P, _P, ctypes.POINTER(ctypes.c_double)]),
    "vfd_batch_adjoint": (ctypes.c_int, [_P, _U32, _P, _P, _P, _P, ctypes.POINTER(ctypes.c_double)]),
    "vfd_get_tuning": (ctypes.c_int, [_P, ctypes.POINTER(Tuning)]),
    "vfd_set_tuning": (ctypes.c_int, [_P, ctypes.POINTER(Tuning)]),
}

EXPORTED = sorted(_PROTOS)

_lib = None
ABI_VERSION = 3  # VFD_ABI_VERSION of include/vfd.h, whose structs the ctypes classes above mirror


class VfdError(RuntimeError):
    def __init__(self, code, fn, msg):
        super().__init__("%s failed with %s: %s" % (fn, STATUS_NAMES.get(code, code), msg))
        self.code = code


def load_library(path=None):
    """Loads libvfd_hip.so and declares every prototype. Raises if the library is absent."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError("HIP extension not built: %s missing (run python -m vfd_amd.build)" % p)
    lib = ctypes.CDLL(p)
    for name, (res, args) in _PROTOS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.vfd_abi_version() != ABI_VERSION:
        raise RuntimeError("%s implements ABI version %d, this binding describes version %d (rebuild with python -m "
                           "vfd_amd.build)" % (p, lib.vfd_abi_version(), ABI_VERSION))
    if path is None:
        _lib = lib
    return lib


def check(code, fn="vfd call"):
    if code != VFD_OK:
        msg = load_library().vfd_last_error()
        raise VfdError(code, fn, msg.decode() if msg else "")
    return code


def ptr(a):
    """numpy array -> void* (None passes NULL)."""
    if a is None:
        return None
    return ctypes.c_void_p(a.ctypes.data)
