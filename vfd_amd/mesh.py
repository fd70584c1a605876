"""Triangle meshes: TriangleMesh (VFD/Source/Renderer/Mesh/TriangleMesh.cpp) as the DFSPH inputs use it - vertices
and triangles, loaded from Wavefront OBJ files or built as a box (or, for tests and benchmarks, an icosphere).

    mesh = TriangleMesh.load_obj("Cube.obj", scale=(1, 1, 1))
    body = RigidBodyDescription("mesh", mesh=mesh, transform=M, resolution=(20, 20, 20))
    fluid = FluidObject.sample(FluidObjectDescription(mesh, transform=M, sample_mode=1), particle_radius)
"""
import numpy as np

from . import _abi


def _f32(x):
    return np.float32(x)


class TriangleMesh:
    """Vertices (n, 3) float32 and triangles (m, 3) uint32 of zero-based vertex indices."""

    def __init__(self, vertices, triangles):
        self.vertices = np.ascontiguousarray(vertices, dtype=np.float32).reshape(-1, 3)
        self.triangles = np.ascontiguousarray(triangles, dtype=np.uint32).reshape(-1, 3)
        if len(self.triangles) and int(self.triangles.max()) >= len(self.vertices):
            raise ValueError("a triangle names a vertex beyond the vertex list")

    def GetVertices(self):
        return self.vertices

    def GetTriangles(self):
        return self.triangles

    @classmethod
    def load_obj(cls, path, scale=(1.0, 1.0, 1.0)):
        """TriangleMesh::LoadOBJ (TriangleMesh.cpp:39-128) over tinyobjloader's LoadObj with triangulation
        (ThirdParty/tinyobjloader/tiny_obj_loader.h): every `v` line gives a vertex (times `scale`), every `f` line a
        polygon of 1-based (or negative, relative) vertex indices, faces of every object in file order. A quad is
        split along its shorter diagonal as tinyobjloader does (|x2 - x0|^2 < |x3 - x1|^2 in float: (0,1,2), (0,2,3);
        else (0,1,3), (1,2,3)); a polygon of more vertices is split as a fan from its first vertex (tinyobjloader
        clips ears instead; both cover the same planar convex polygon). Faces with fewer than 3 vertices are
        skipped."""
        verts, tris = [], []
        with open(path, "r") as f:
            for line in f:
                tok = line.split()
                if not tok:
                    continue
                if tok[0] == "v":
                    verts.append((float(tok[1]), float(tok[2]), float(tok[3])))
                elif tok[0] == "f":
                    idx = []
                    for t in tok[1:]:
                        k = int(t.split("/")[0])
                        idx.append(k - 1 if k > 0 else len(verts) + k)
                    if len(idx) < 3:
                        continue
                    if len(idx) == 3:
                        tris.append(idx)
                    elif len(idx) == 4:
                        tris.extend(cls._split_quad(verts, idx))
                    else:
                        tris.extend([idx[0], idx[k], idx[k + 1]] for k in range(1, len(idx) - 1))
        v = np.asarray(verts, dtype=np.float32).reshape(-1, 3)
        s = np.asarray(scale, dtype=np.float32)
        return cls(v * s, np.asarray(tris, dtype=np.uint32).reshape(-1, 3))

    @staticmethod
    def _split_quad(verts, idx):
        p = [np.asarray(verts[i], dtype=np.float32) for i in idx]

        def sq(a, b):
            d = b - a
            return (d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]  # float32 scalars, no fused terms

        if sq(p[0], p[2]) < sq(p[1], p[3]):
            return [[idx[0], idx[1], idx[2]], [idx[0], idx[2], idx[3]]]
        return [[idx[0], idx[1], idx[3]], [idx[1], idx[2], idx[3]]]

    @classmethod
    def box(cls, position, width, height, depth):
        """TriangleMesh(const AABB&) (TriangleMesh.cpp:18-37): the 8 corners from `position` and the 12 triangles of
        the reference's table, wound outward."""
        x, y, z = (_f32(c) for c in position)
        w, h, d = _f32(width), _f32(height), _f32(depth)
        v = [(x, y, z), (x + w, y, z), (x + w, y, z + d), (x, y, z + d),
             (x, y + h, z), (x + w, y + h, z), (x + w, y + h, z + d), (x, y + h, z + d)]
        t = [(0, 1, 2), (0, 2, 3), (4, 7, 6), (4, 6, 5), (0, 3, 7), (0, 7, 4), (1, 5, 6), (1, 6, 2),
             (0, 4, 5), (0, 5, 1), (3, 2, 6), (3, 6, 7)]
        return cls(np.asarray(v, dtype=np.float32), np.asarray(t, dtype=np.uint32))

    @classmethod
    def icosphere(cls, level, radius=1.0, bumps=0.0, seed=0):
        """A closed, outward-wound sphere mesh: the icosahedron with every triangle split in four `level` times
        (20 * 4^level triangles), vertices on the sphere of `radius`, then moved radially by up to `bumps` * radius
        with a smooth random field (a test and benchmark mesh; no reference counterpart)."""
        t = (1.0 + 5.0 ** 0.5) / 2.0
        v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
             (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
        v = [np.asarray(p, np.float64) / np.linalg.norm(p) for p in v]
        f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
             (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
             (6, 2, 10), (8, 6, 7), (9, 8, 1)]
        for _ in range(level):
            mid = {}

            def m(a, b):
                k = (min(a, b), max(a, b))
                if k not in mid:
                    c = v[a] + v[b]
                    v.append(c / np.linalg.norm(c))
                    mid[k] = len(v) - 1
                return mid[k]
            f = [g for a, b, c in f for g in ((a, m(a, b), m(a, c)), (b, m(b, c), m(a, b)), (c, m(a, c), m(b, c)),
                                              (m(a, b), m(b, c), m(a, c)))]
        x = np.asarray(v)
        if bumps:
            rng = np.random.default_rng(seed)
            k = rng.normal(size=(6, 3))
            ph = rng.uniform(0, 2 * np.pi, 6)
            h = sum(np.sin(3.0 * x @ k[i] + ph[i]) for i in range(6)) / 6.0
            x = x * (1.0 + bumps * h)[:, None]
        return cls((x * radius).astype(np.float32), np.asarray(f, np.uint32))

    def to_abi(self):
        """The vfd_mesh view; keep `self` alive while the struct is in use."""
        m = _abi.Mesh()
        m.Vertices = self.vertices.ctypes.data
        m.VertexCount = len(self.vertices)
        m.Triangles = self.triangles.ctypes.data
        m.TriangleCount = len(self.triangles)
        return m


def glm_transform(matrix):
    """A 4x4 matrix M acting as M @ (x, y, z, 1) in the column-major float[16] glm::mat4 layout."""
    m = np.asarray(matrix, dtype=np.float32).reshape(4, 4)
    return np.ascontiguousarray(m.T, dtype=np.float32).ravel()


def transform_vertices(matrix, vertices):
    """Transform * vec4(v, 1) with glm's evaluation order (m0 x + m1 y) + (m2 z + m3 w), in float32."""
    m = np.asarray(matrix, dtype=np.float32).reshape(4, 4)
    v = np.asarray(vertices, dtype=np.float32).reshape(-1, 3)
    one = np.float32(1.0)
    return np.stack([(m[r, 0] * v[:, 0] + m[r, 1] * v[:, 1]) + (m[r, 2] * v[:, 2] + m[r, 3] * one)
                     for r in range(3)], axis=1).astype(np.float32)
