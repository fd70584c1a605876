"""Python binding of the CPU oracle (oracle/dfsph_oracle.cpp). TEST INFRASTRUCTURE ONLY.

Used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker; the product never imports
this module.

    from oracle.oracle import OracleSim, build_oracle
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "dfsph_oracle.cpp")
LIB = os.path.join(HERE, "liboracle.so")
CXXFLAGS = ["-std=c++17", "-O3", "-march=x86-64-v3", "-fopenmp", "-ffp-contract=off", "-fPIC", "-shared"]


def build_oracle(force=True):
    """Compiles the oracle with g++ (IEEE, no contraction, OpenMP)."""
    if force or not os.path.exists(LIB):
        tmp = LIB + ".tmp"
        subprocess.check_call(["g++"] + CXXFLAGS + ["-I", os.path.join(os.path.dirname(HERE), "include"), SRC,
                                                    "-o", tmp])
        os.replace(tmp, LIB)
    return LIB


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError("oracle not built: %s missing (python -c 'import oracle.oracle as o; o.build_oracle()')"
                               % LIB)
        L = ctypes.CDLL(LIB)
        P = ctypes.c_void_p
        L.orc_create.restype = P
        L.orc_create.argtypes = [P, ctypes.c_int]
        L.orc_destroy.argtypes = [P]
        for name, args in {
            "orc_set_description": [P, P], "orc_set_fluid": [P, P, P, ctypes.c_uint32],
            "orc_add_boundary": [P, P], "orc_step": [P], "orc_step_until": [P, ctypes.c_int],
            "orc_set_time_step": [P, ctypes.c_double], "orc_get_field": [P, ctypes.c_int, ctypes.c_uint32, P],
            "orc_set_field": [P, ctypes.c_int, P], "orc_particle_count": [P, P],
            "orc_get_neighbors": [P, P, P, P, P], "orc_get_debug_info": [P, P], "orc_get_info": [P, P],
            "orc_get_kernel_table": [P, P, P, P], "orc_kernel_eval": [P, ctypes.c_int, P, ctypes.c_uint32, P],
            "orc_get_boundary_map": [P, ctypes.c_uint32, ctypes.c_uint32, P, P, P],
            "orc_map_interpolate": [P, ctypes.c_uint32, ctypes.c_uint32, P, ctypes.c_uint32, P],
            "orc_brute_force_neighbors": [P, ctypes.c_uint32, ctypes.c_float, ctypes.c_uint32, P, P],
            "orc_search_grid": [P, ctypes.c_uint32, ctypes.c_float, P, P],
            "orc_viscosity_system": [P, P, P],
            "orc_set_smooth_kernel": [P, ctypes.c_int], "orc_tape": [P, ctypes.c_int],
            "orc_adjoint": [P, P, P, P, P, P, P],
            "orc_add_boundary_mesh": [P, P, P, ctypes.c_uint32, P, ctypes.c_uint32],
            "orc_sample_mesh_volume": [ctypes.c_int, P, ctypes.c_uint32, P, ctypes.c_uint32, P, ctypes.c_float, P,
                                       ctypes.c_int, ctypes.c_int, P, ctypes.c_uint32, P],
            "orc_mesh_signed_distance": [ctypes.c_int, P, ctypes.c_uint32, P, ctypes.c_uint32, P, P,
                                         ctypes.c_uint32, P],
        }.items():
            fn = getattr(L, name)
            fn.restype = ctypes.c_int
            fn.argtypes = args
        L.orc_halton.argtypes = [ctypes.c_int, P]
        L.orc_gauss_legendre.argtypes = [ctypes.c_int, P, P]
        L.orc_shape_function.argtypes = [P, P, P]
        L.orc_shape_function_hess.argtypes = [P, P]
        L.orc_morton3.restype = ctypes.c_uint32
        L.orc_morton3.argtypes = [ctypes.c_uint32] * 3
        L.orc_cell_nodes.argtypes = [P, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, P]
        L.orc_node_position.argtypes = [P, P, P, ctypes.c_uint32, P]
        L.orc_node_count.restype = ctypes.c_uint32
        L.orc_node_count.argtypes = [P]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


FIELD_IDS = {
    "position": (0, 3), "velocity": (1, 3), "acceleration": (2, 3), "pressure_acceleration": (3, 3),
    "pressure_residuum": (4, 1), "density": (5, 1), "density_advection": (6, 1), "pressure_rho2": (7, 1),
    "pressure_rho2_v": (8, 1), "factor": (9, 1), "velocity_difference": (10, 3), "mc_surface_normal": (11, 3),
    "mc_surface_normal_smooth": (12, 3), "mc_surface_curvature": (13, 1), "mc_surface_curvature_smooth": (14, 1),
    "delta_final_curvature": (15, 1), "boundary_xj": (16, 3), "boundary_volume": (17, 1),
    "viscosity_solution": (18, 3), "viscosity_preconditioner": (19, 9),
}
PHASES = ["search", "boundary", "density", "factor", "divergence", "clear_acceleration", "surface_tension",
          "viscosity", "time_step", "velocity", "pressure", "position"]
INFO_NAMES = ["ParticleCount", "RigidBodyCount", "SupportRadius", "SupportRadius2", "ParticleRadius",
              "ParticleDiameter", "TimeStepSize", "TimeStepSize2", "TimeStepSizeInverse", "TimeStepSize2Inverse",
              "Volume", "Density0", "ParticleMass", "ParticleMassInverse", "Viscosity", "BoundaryViscosity",
              "DynamicViscosity", "DynamicBoundaryViscosity", "TangentialDistanceFactor", "TangentialDistance",
              "SurfaceTension", "SurfaceTensionSampleCount", "ClassifierSlope", "ClassifierConstant",
              "TemporalSmoothing", "SmoothingFactor", "Factor", "NeighborParticleRadius", "MonteCarloFactor",
              "GravityX", "GravityY", "GravityZ"]


class OracleSim:
    """CPU DFSPH step (fp64=False: FP32 mirror; fp64=True: FP64 twin) with the same calls as the product."""

    def __init__(self, desc, fp64=False):
        self._L = lib()
        self._h = self._L.orc_create(ctypes.byref(desc), 1 if fp64 else 0)
        self.fp64 = fp64

    def close(self):
        if self._h:
            self._L.orc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_fluid(self, x, v=None):
        x = np.ascontiguousarray(x, np.float32).reshape(-1, 3)
        v = None if v is None else np.ascontiguousarray(v, np.float32).reshape(-1, 3)
        self._L.orc_set_fluid(self._h, _p(x), _p(v), len(x))
        self.n = len(x)

    def add_boundary(self, b):
        self._L.orc_add_boundary(self._h, ctypes.byref(b))

    def add_boundary_mesh(self, desc, vertices, triangles):
        """A mesh body: desc is a vfd_mesh_boundary_desc (ctypes), vertices (n, 3) float32, triangles (m, 3)."""
        v = np.ascontiguousarray(vertices, np.float32)
        t = np.ascontiguousarray(triangles, np.uint32)
        self._L.orc_add_boundary_mesh(self._h, ctypes.byref(desc), _p(v), len(v), _p(t), len(t))

    def step(self):
        self._L.orc_step(self._h)

    # ---- reference adjoint ----------------------------------------------------------------------------------
    def set_smooth_kernel(self, on):
        """Evaluate W and gradW from the spline instead of the table (a differentiable forward for adjoint pins)."""
        self._L.orc_set_smooth_kernel(self._h, 1 if on else 0)

    def tape(self, on):
        """Start (clearing earlier records) or stop recording steps for adjoint()."""
        self._L.orc_tape(self._h, 1 if on else 0)

    def adjoint(self, lx_out, lv_out=None):
        """Cotangents of the last recorded step's (x, v) -> those of the first recorded step's input and d/dmu
        (mu = dynamic viscosity, Viscosity * rho0)."""
        n = self.n
        lx = np.ascontiguousarray(lx_out, np.float64).reshape(n, 3)
        lv = np.zeros((n, 3)) if lv_out is None else np.ascontiguousarray(lv_out, np.float64).reshape(n, 3)
        ox, ov, mu = np.zeros((n, 3)), np.zeros((n, 3)), ctypes.c_double(0)
        err = ctypes.create_string_buffer(256)
        k = self._L.orc_adjoint(self._h, _p(lx), _p(lv), _p(ox), _p(ov), ctypes.byref(mu), err)
        if k < 0:
            raise RuntimeError("adjoint: " + err.value.decode())
        return ox, ov, mu.value

    def step_until(self, phase):
        self._L.orc_step_until(self._h, PHASES.index(phase) if isinstance(phase, str) else int(phase))

    def set_time_step(self, dt):
        self._L.orc_set_time_step(self._h, float(dt))

    def count(self):
        n = ctypes.c_uint32()
        self._L.orc_particle_count(self._h, ctypes.byref(n))
        return n.value

    def field(self, name, body=0):
        fid, comps = FIELD_IDS[name]
        out = np.zeros((self.count(), comps), np.float64)
        rc = self._L.orc_get_field(self._h, fid, body, _p(out))
        if rc:
            raise ValueError("oracle field %s: status %d" % (name, rc))
        return out if comps > 1 else out[:, 0]

    def set_field(self, name, values):
        fid, comps = FIELD_IDS[name]
        a = np.ascontiguousarray(values, np.float64).reshape(-1)
        self._L.orc_set_field(self._h, fid, _p(a))

    def neighbors(self):
        n = self.count()
        total = ctypes.c_uint64()
        counts = np.zeros(n, np.uint32)
        offsets = np.zeros(n, np.uint32)
        self._L.orc_get_neighbors(self._h, _p(counts), _p(offsets), None, ctypes.byref(total))
        ids = np.zeros(max(total.value, 1), np.uint32)
        self._L.orc_get_neighbors(self._h, _p(counts), _p(offsets), _p(ids), ctypes.byref(total))
        return counts, offsets, ids[:total.value]

    def debug_info(self):
        from vfd_amd._abi import DebugInfo
        d = DebugInfo()
        self._L.orc_get_debug_info(self._h, ctypes.byref(d))
        return d

    def info(self):
        out = np.zeros(32, np.float64)
        self._L.orc_get_info(self._h, _p(out))
        return dict(zip(INFO_NAMES, out.tolist()))

    def kernel_table(self):
        w = np.zeros(10000)
        g = np.zeros(10001)
        s = np.zeros(6)
        self._L.orc_get_kernel_table(self._h, _p(w), _p(g), _p(s))
        return w, g, dict(zip(["Radius", "Radius2", "InvStepSize", "WZero", "K", "L"], s.tolist()))

    def kernel_eval(self, which, r):
        r = np.ascontiguousarray(r, np.float64).reshape(-1, 3)
        out = np.zeros_like(r)
        self._L.orc_kernel_eval(self._h, {"W": 0, "Ws": 1, "gradW": 2}[which], _p(r), len(r), _p(out))
        return out[:, 0] if which != "gradW" else out

    def boundary_map(self, body, field):
        cnt = ctypes.c_uint32()
        geom = np.zeros(9)
        self._L.orc_get_boundary_map(self._h, body, field, None, ctypes.byref(cnt), _p(geom))
        out = np.zeros(cnt.value)
        self._L.orc_get_boundary_map(self._h, body, field, _p(out), ctypes.byref(cnt), _p(geom))
        return out, geom

    def viscosity_system(self):
        n = self.count()
        A = np.zeros((3 * n, 3 * n))
        b = np.zeros(3 * n)
        self._L.orc_viscosity_system(self._h, _p(A), _p(b))
        return A, b

    def map_interpolate(self, body, field, pts):
        pts = np.ascontiguousarray(pts, np.float64).reshape(-1, 3)
        out = np.zeros((len(pts), 4))
        self._L.orc_map_interpolate(self._h, body, field, _p(pts), len(pts), _p(out))
        return out


def halton(fp64=True):
    out = np.zeros(49152)
    lib().orc_halton(1 if fp64 else 0, _p(out))
    return out


def gauss_legendre(n=16):
    x = np.zeros(n)
    w = np.zeros(n)
    lib().orc_gauss_legendre(n, _p(x), _p(w))
    return x, w


def shape_function_hess(xi):
    """(32, 6) second derivatives (xx, yy, zz, xy, xz, yz) of the shape functions at local coordinates xi."""
    xi = np.ascontiguousarray(xi, np.float64)
    H = np.zeros(192)
    lib().orc_shape_function_hess(_p(xi), _p(H))
    return H.reshape(32, 6)


def shape_function(xi):
    xi = np.ascontiguousarray(xi, np.float64)
    N = np.zeros(32)
    dN = np.zeros(96)
    lib().orc_shape_function(_p(xi), _p(N), _p(dN))
    return N, dN.reshape(32, 3)


def morton3(x, y, z):
    return lib().orc_morton3(x, y, z)


def cell_nodes(res, i, j, k):
    r = np.ascontiguousarray(res, np.uint32)
    out = np.zeros(32, np.uint32)
    lib().orc_cell_nodes(_p(r), i, j, k, _p(out))
    return out


def node_position(dmin, dmax, res, i):
    a = np.ascontiguousarray(dmin, np.float64)
    b = np.ascontiguousarray(dmax, np.float64)
    r = np.ascontiguousarray(res, np.uint32)
    out = np.zeros(3)
    lib().orc_node_position(_p(a), _p(b), _p(r), i, _p(out))
    return out


def node_count(res):
    r = np.ascontiguousarray(res, np.uint32)
    return lib().orc_node_count(_p(r))


def brute_force_neighbors(x, R, cap=70):
    x = np.ascontiguousarray(x, np.float32).reshape(-1, 3)
    n = len(x)
    counts = np.zeros(n, np.uint32)
    ids = np.zeros(n * cap, np.uint32)
    lib().orc_brute_force_neighbors(_p(x), n, float(R), cap, _p(counts), _p(ids))
    return counts, ids.reshape(n, cap)


def search_grid(x, R):
    x = np.ascontiguousarray(x, np.float32).reshape(-1, 3)
    f = np.zeros(6, np.float32)
    u = np.zeros(7, np.uint32)
    lib().orc_search_grid(_p(x), len(x), float(R), _p(f), _p(u))
    return dict(gridMin=f[:3], gridDelta=f[3:], dims=u[:3], meta=u[3:6], cellCount=int(u[6]))


def _mesh_args(vertices, triangles, transform):
    v = np.ascontiguousarray(vertices, np.float32).reshape(-1, 3)
    t = np.ascontiguousarray(triangles, np.uint32).reshape(-1, 3)
    if len(t) == 0 or int(t.max()) >= len(v):
        raise ValueError("bad mesh")
    tr = None if transform is None else np.ascontiguousarray(np.asarray(transform, np.float32).reshape(4, 4).T).ravel()
    return v, t, tr


def sample_mesh_volume(vertices, triangles, radius, resolution, inverted=False, mode=0, transform=None, fp64=False):
    """ParticleSampler::SampleMeshVolume of the transformed mesh (transform: 4x4 acting on (x, y, z, 1))."""
    v, t, tr = _mesh_args(vertices, triangles, transform)
    res = np.ascontiguousarray(resolution, np.uint32)
    n = ctypes.c_uint32()
    lib().orc_sample_mesh_volume(1 if fp64 else 0, _p(v), len(v), _p(t), len(t), _p(tr), float(radius), _p(res),
                                 int(bool(inverted)), int(mode), None, 0, ctypes.byref(n))
    out = np.zeros((n.value, 3))
    lib().orc_sample_mesh_volume(1 if fp64 else 0, _p(v), len(v), _p(t), len(t), _p(tr), float(radius), _p(res),
                                 int(bool(inverted)), int(mode), _p(out), n.value, ctypes.byref(n))
    return out


def mesh_signed_distance(vertices, triangles, points, transform=None, fp64=True):
    """MeshDistance::SignedDistance at each point (negative inside)."""
    v, t, tr = _mesh_args(vertices, triangles, transform)
    p = np.ascontiguousarray(points, np.float64).reshape(-1, 3)
    out = np.zeros(len(p))
    lib().orc_mesh_signed_distance(1 if fp64 else 0, _p(v), len(v), _p(t), len(t), _p(tr), _p(p), len(p), _p(out))
    return out
