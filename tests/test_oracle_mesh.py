"""Pins the oracle's triangle-mesh inputs (CPU only): MeshDistance::SignedDistance (Utility/SDF/MeshDistance.cpp),
ParticleSampler::SampleMeshVolume (Utility/Sampler/ParticleSampler.cpp) and the OBJ loading of TriangleMesh.

Independent checks, not restatements of the oracle's formulas:
- the distance to a mesh of a box equals the closed-form box distance;
- on the reference's Sphere, Cone and Torus meshes the distance equals a brute-force closest point (the plane
  projection when it falls inside the triangle, else the nearest of the three clamped segment projections - a
  different construction from the oracle's region classification);
- the sign equals inside/outside by plane tests (convex meshes) or by ray-crossing parity (the torus);
- at the border of an open mesh the edge normal is the face's own normal;
- the sampler keeps exactly the lattice points of the reference's float loops that lie inside the analytic box.
Fixtures: tests/golden/models/*.obj, copied from VFD/Resources/Models by tests/golden/make_golden.py."""
import os

import numpy as np
import pytest

from oracle import oracle as orc
from vfd_amd.mesh import TriangleMesh, glm_transform, transform_vertices

HERE = os.path.dirname(os.path.abspath(__file__))
MODELS = os.path.join(HERE, "golden", "models")


def load(name, scale=(1.0, 1.0, 1.0)):
    return TriangleMesh.load_obj(os.path.join(MODELS, name), scale)


def brute_force_distance(points, verts, tris):
    """Unsigned distance to the closest triangle, by projection and clamping (float64)."""
    p = np.asarray(points, np.float64)
    best = np.full(len(p), np.inf)
    V = verts.astype(np.float64)
    for a, b, c in tris:
        A, B, C = V[a], V[b], V[c]
        n = np.cross(B - A, C - A)
        n = n / np.linalg.norm(n)
        h = (p - A) @ n
        q = p - h[:, None] * n  # projection onto the plane
        # barycentric coordinates of q
        v0, v1, v2 = B - A, C - A, q - A
        d00, d01, d11 = v0 @ v0, v0 @ v1, v1 @ v1
        d20, d21 = v2 @ v0, v2 @ v1
        den = d00 * d11 - d01 * d01
        bv = (d11 * d20 - d01 * d21) / den
        bw = (d00 * d21 - d01 * d20) / den
        inside = (bv >= 0) & (bw >= 0) & (bv + bw <= 1)
        d = np.where(inside, np.abs(h), np.inf)
        for X, Y in ((A, B), (B, C), (C, A)):
            e = Y - X
            t = np.clip(((p - X) @ e) / (e @ e), 0.0, 1.0)
            d = np.minimum(d, np.linalg.norm(p - (X + t[:, None] * e), axis=1))
        best = np.minimum(best, d)
    return best


def ray_parity_inside(points, verts, tris, direction=(0.5773, 0.5774, 0.5775)):
    """Inside by the parity of crossings of a ray (Moller-Trumbore), float64."""
    p = np.asarray(points, np.float64)
    dvec = np.asarray(direction, np.float64)
    count = np.zeros(len(p), int)
    V = verts.astype(np.float64)
    for a, b, c in tris:
        A, B, C = V[a], V[b], V[c]
        e1, e2 = B - A, C - A
        h = np.cross(dvec, e2)
        det = e1 @ h
        if abs(det) < 1e-14:
            continue
        s = p - A
        u = (s @ h) / det
        q = np.cross(s, e1)
        v = (q @ dvec) / det
        t = (q @ e2) / det
        count += ((u >= 0) & (v >= 0) & (u + v <= 1) & (t > 0)).astype(int)
    return count % 2 == 1


def test_obj_loader_splits_quads_along_the_shorter_diagonal():
    cube = load("Cube.obj")
    assert cube.vertices.shape == (8, 3) and cube.triangles.shape == (12, 3)
    text = open(os.path.join(MODELS, "Cube.obj")).read().split("\n")
    quads = [[int(t.split("/")[0]) - 1 for t in l.split()[1:]] for l in text if l.startswith("f ")]
    for q, (t0, t1) in zip(quads, cube.triangles.reshape(-1, 2, 3)):
        assert set(t0) | set(t1) == set(q)
        shared = sorted(set(t0) & set(t1))
        v = cube.vertices.astype(np.float64)
        diag = lambda i, j: np.sum((v[q[i]] - v[q[j]]) ** 2)
        # tinyobjloader keeps (0, 2) only when it is strictly shorter than (1, 3)
        assert shared == sorted([q[0], q[2]] if diag(0, 2) < diag(1, 3) else [q[1], q[3]])
    for name, nv, nt in (("Sphere.obj", 42, 80), ("Cone.obj", 33, 62), ("Torus.obj", 576, 1152)):
        m = load(name)
        assert m.vertices.shape == (nv, 3) and m.triangles.shape == (nt, 3), name


def test_box_mesh_distance_is_the_box_distance():
    cube = load("Cube.obj")
    rng = np.random.default_rng(0)
    p = rng.uniform(-2.0, 2.0, (4000, 3))
    p = np.concatenate([p, rng.uniform(-1.05, 1.05, (2000, 3))])
    q = np.abs(p) - 1.0
    exact = np.linalg.norm(np.maximum(q, 0.0), axis=1) + np.minimum(q.max(axis=1), 0.0)
    d64 = orc.mesh_signed_distance(cube.vertices, cube.triangles, p, fp64=True)
    assert np.abs(d64 - exact).max() < 1e-7  # sqrt of the double rounding of |x0 - p|^2 near the surface
    d32 = orc.mesh_signed_distance(cube.vertices, cube.triangles, p, fp64=False)
    # the reference's float quadratic form loses ~1e-7 of |x0 - p|^2, i.e. up to ~1e-3 in sqrt near the surface
    assert np.abs(d32 - exact).max() < 2e-3
    far = np.abs(exact) > 3e-3
    assert np.array_equal(np.sign(d32[far]), np.sign(exact[far]))


@pytest.mark.parametrize("name,convex", [("Sphere.obj", True), ("Cone.obj", True), ("Torus.obj", False)])
def test_mesh_distance_against_brute_force_and_sign(name, convex):
    m = load(name)
    lo, hi = m.vertices.min(0), m.vertices.max(0)
    rng = np.random.default_rng(1)
    n = 600 if name == "Torus.obj" else 3000
    p = rng.uniform(lo - 0.4, hi + 0.4, (n, 3))
    d = orc.mesh_signed_distance(m.vertices, m.triangles, p, fp64=True)
    ref = brute_force_distance(p, m.vertices, m.triangles)
    assert np.abs(np.abs(d) - ref).max() < 1e-7
    if convex:
        V = m.vertices.astype(np.float64)
        A, B, C = V[m.triangles[:, 0]], V[m.triangles[:, 1]], V[m.triangles[:, 2]]
        nrm = np.cross(B - A, C - A)
        nrm /= np.linalg.norm(nrm, axis=1)[:, None]
        inside = (np.einsum("fk,nfk->nf", nrm, p[:, None, :] - A[None]) < 0).all(axis=1)
    else:
        inside = ray_parity_inside(p, m.vertices, m.triangles)
    clear = ref > 1e-6
    assert clear.sum() > 0.9 * n
    assert np.array_equal(d[clear] < 0, inside[clear])
    assert inside.sum() > 0.05 * n  # both sides are exercised


def test_open_mesh_border_edges_use_the_face_normal():
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    t = np.array([[0, 1, 2]], np.uint32)
    rng = np.random.default_rng(2)
    p = rng.uniform(-1.0, 2.0, (2000, 3))
    p[:, 2] = rng.choice([-1.0, 1.0], 2000) * rng.uniform(0.01, 0.5, 2000)
    d = orc.mesh_signed_distance(v, t, p, fp64=True)
    # vertex normals of a lone triangle are (angle) * n, edge normals n, face normal n: all along +z
    assert np.array_equal(d < 0, p[:, 2] < 0)
    assert np.abs(np.abs(d) - brute_force_distance(p, v, t)).max() < 1e-7


def test_transform_is_applied_in_glm_order():
    cube = load("Cube.obj")
    M = np.array([[2.0, 0, 0, 0.25], [0, 0.2, 0, -0.5], [0, 0, 2.0, 0.125], [0, 0, 0, 1]], np.float32)
    rng = np.random.default_rng(3)
    p = rng.uniform(-3.0, 3.0, (500, 3))
    a = orc.mesh_signed_distance(cube.vertices, cube.triangles, p, transform=M, fp64=False)
    b = orc.mesh_signed_distance(transform_vertices(M, cube.vertices), cube.triangles, p, fp64=False)
    assert np.array_equal(a, b)
    assert glm_transform(M)[12:15].tolist() == [np.float32(0.25), np.float32(-0.5), np.float32(0.125)]


def reference_lattice(lo, hi, radius, mode):
    """The candidate points of SampleMeshVolume's float loops (ParticleSampler.cpp:19-77), in loop order."""
    f = np.float32
    r = f(radius)
    d = f(2.0) * r
    sx = sy = d
    if mode == 1:
        sy = np.sqrt(f(3.0)) * r
    elif mode == 2:
        sx = np.sqrt(f(3.0)) * r
        sy = np.sqrt(f(6.0)) * d / f(3.0)

    def loop(a, b, step):
        out, v = [], f(a)
        while v <= f(b):
            out.append(v)
            v = f(v + step)
        return out

    pts = []
    for z in loop(lo[2], hi[2], d):
        for j, y in enumerate(loop(lo[1], hi[1], sy)):
            for i, x in enumerate(loop(lo[0], hi[0], sx)):
                if mode == 0:
                    p = (x + r, y + r, z + r)
                elif mode == 1:
                    p = (x, y + r, z + r) if j % 2 == 0 else (x + r, y + r, z)
                else:
                    sh = [f(0), f(0), f(0)]
                    if i % 2:
                        sh[2] = f(sh[2] + d / (f(2.0) * f(-1.0 if j % 2 else 1.0)))
                    if j % 2:
                        sh[0] = f(sh[0] + sx / f(2.0))
                        sh[2] = f(sh[2] + d / f(2.0))
                    p = (f(x + sh[0]), f(f(y + r) + sh[1]), f(f(z + r) + sh[2]))
                pts.append(p)
    return np.asarray(pts, np.float32)


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_sampler_keeps_the_lattice_points_inside_a_box(mode):
    cube = load("Cube.obj")
    # a box [-0.6, 1.2] x [-0.3, 0.9] x [-0.9, 0.45]: holds the origin, so the sampler's bounds are the box
    M = np.array([[0.9, 0, 0, 0.3], [0, 0.6, 0, 0.3], [0, 0, 0.675, -0.225], [0, 0, 0, 1]], np.float32)
    v = transform_vertices(M, cube.vertices)
    lo, hi = v.min(0), v.max(0)
    radius = 0.0625
    got = orc.sample_mesh_volume(cube.vertices, cube.triangles, radius, (24, 24, 24), mode=mode, transform=M)
    cand = reference_lattice(lo, hi, radius, mode)
    c64 = cand.astype(np.float64)
    q = np.abs(c64 - (lo + hi) / 2) - (hi - lo) / 2
    sd = np.linalg.norm(np.maximum(q, 0.0), axis=1) + np.minimum(q.max(axis=1), 0.0)
    # every sample is a candidate, in loop order
    index = {tuple(c): k for k, c in enumerate(cand.tolist())}
    picked = np.array([index[tuple(p)] for p in got.astype(np.float32).tolist()])
    assert np.all(np.diff(picked) > 0)
    kept = np.zeros(len(cand), bool)
    kept[picked] = True
    # away from the surface the choice is the geometry's (the map's interpolation error is far below 5e-3 here)
    clear = np.abs(sd) > 5e-3
    assert np.array_equal(kept[clear], sd[clear] < 0)
    assert clear.sum() > 0.9 * len(cand) and len(got) > 100  # medium/max rows start on the faces
    # the FP64 twin differs only on candidates within 5e-3 of the surface
    got64 = orc.sample_mesh_volume(cube.vertices, cube.triangles, radius, (24, 24, 24), mode=mode, transform=M,
                                   fp64=True)
    assert abs(len(got64) - len(got)) <= (~clear).sum()


def test_box_mesh_body_matches_the_analytic_box_body():
    """RigidBody ctor (RigidBody.cu:10-73) for Cube.obj scaled to a box against the same box as an analytic shape
    (FP64 twin): the map domains agree (mesh bounding box grown by 8R + tol) and so do both fields, the signed
    distance to rounding and the boundary volume to the interpolation of that rounding."""
    from vfd_amd.simulation import DFSPHSimulationDescription, RigidBodyDescription
    r = 0.02
    desc = DFSPHSimulationDescription(ParticleRadius=r)
    x = np.zeros((1, 3), np.float32)
    cube = load("Cube.obj")
    h = (0.5, 0.3, 0.4)
    M = np.diag([h[0], h[1], h[2], 1.0]).astype(np.float32)
    mesh_body = RigidBodyDescription("mesh", mesh=cube, transform=M, padding=0.01, resolution=(12, 10, 12))
    box_body = RigidBodyDescription("box", half_extents=h, padding=0.01, resolution=(12, 10, 12))
    a = orc.OracleSim(desc, fp64=True)
    a.set_fluid(x)
    a.add_boundary_mesh(mesh_body.to_abi(), cube.vertices, cube.triangles)
    b = orc.OracleSim(desc, fp64=True)
    b.set_fluid(x)
    b.add_boundary(box_body.to_abi())
    (m0, ga), (a0, gb) = a.boundary_map(0, 0), b.boundary_map(0, 0)
    assert np.allclose(ga, gb, rtol=0, atol=1e-6)  # domain and cell geometry
    assert np.abs(m0 - a0).max() < 1e-6
    m1, _ = a.boundary_map(0, 1)
    a1, _ = b.boundary_map(0, 1)
    assert np.abs(m1 - a1).max() < 1e-6 * max(1.0, np.abs(a1).max()) + 1e-9 and a1.max() > 0
