"""Pins of the CPU oracle (no GPU): reference-held data tables, closed forms, invariants and brute force.

The reference has no tests or golden vectors for the DFSPH path; the only reference-held data are the Halton and
Gauss-Legendre tables (tests/golden/*.json, written by tests/golden/make_golden.py). Everything else is pinned here
by mathematics chosen so that a dropped term, a wrong sign or index, or a transposed operand fails a check.
"""
import json
import math
import os

import numpy as np
import pytest

from oracle import oracle as O
from vfd_amd import scenes
from vfd_amd.simulation import DFSPHSimulationDescription, RigidBodyDescription

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module", autouse=True)
def _built(oracle_lib):
    return oracle_lib


# ---- reference-held tables -------------------------------------------------------------------------------------
def test_halton_matches_reference_table():
    g = json.load(open(os.path.join(GOLDEN, "halton_samples.json")))
    h64 = O.halton(fp64=True)
    h32 = O.halton(fp64=False).astype(np.float32)
    for k, v in g["points"].items():
        i = int(k)
        assert np.abs(h64[3 * i: 3 * i + 3] - np.array(v)).max() < 1e-14
        assert np.array_equal(h32[3 * i: 3 * i + 3], np.array(v, dtype=np.float32))
    pts = h64.reshape(-1, 3)
    assert np.abs(np.linalg.norm(pts, axis=1) - 1).max() < 1e-14
    # the whole table, 49,152 values as the reference's source writes them (15 decimals)
    full = np.load(os.path.join(GOLDEN, "halton_table.npz"))["values"]
    assert full.shape == (49152,)
    assert np.abs(h64 - full).max() < 4e-15  # the rounding of the 15th decimal
    assert np.array_equal(h32, full.astype(np.float32))


def test_gauss_legendre_matches_reference_table():
    g = json.load(open(os.path.join(GOLDEN, "gauss_legendre_16.json")))
    x, w = O.gauss_legendre(16)
    assert np.abs(x - g["abscissae"]).max() < 1e-15
    assert np.abs(w - g["weights"]).max() < 1e-14
    assert np.array_equal(x.astype(np.float32), np.array(g["abscissae"], dtype=np.float32))
    assert np.array_equal(w.astype(np.float32), np.array(g["weights"], dtype=np.float32))
    # exact for polynomials up to degree 31
    t = np.linspace(-1, 1, 7)
    for deg in (0, 5, 30, 31):
        assert abs((w * x ** deg).sum() - (2.0 / (deg + 1) if deg % 2 == 0 else 0.0)) < 1e-13


# ---- cubic spline kernel table ---------------------------------------------------------------------------------
def _sim(fp64=True, r=0.025, **kw):
    d = DFSPHSimulationDescription(ParticleRadius=r, **kw)
    return O.OracleSim(d, fp64=fp64)


def _spline(r, h):
    q = r / h
    k = 8.0 / (math.pi * h ** 3)
    return np.where(q <= 0.5, k * (6 * q ** 3 - 6 * q ** 2 + 1), np.where(q <= 1, k * 2 * (1 - q) ** 3, 0.0))


@pytest.mark.parametrize("r", [0.025, 0.004])
def test_kernel_table_closed_form(r):
    o = _sim(fp64=True, r=r)
    w, g, s = o.kernel_table()
    h = 4 * float(np.float32(r))  # the description holds r in float
    assert s["Radius"] == pytest.approx(h, rel=1e-15)
    assert s["K"] == pytest.approx(8 / (math.pi * h ** 3), rel=1e-12)
    assert s["L"] == pytest.approx(48 / (math.pi * h ** 3), rel=1e-12)
    x = np.arange(10000) * (h / 9999)
    assert np.allclose(w, _spline(x, h), rtol=1e-12, atol=1e-12 * s["K"])
    # G[i] * x = dW/dr at x (central differences of the closed form), zero below |r| = 1e-5 (CalculateGradientW)
    dw = (_spline(x + 1e-9 * h, h) - _spline(x - 1e-9 * h, h)) / (2e-9 * h)
    m = x > 1e-5
    assert np.allclose(g[:10000][m] * x[m], dw[m], rtol=1e-5, atol=1e-6 * s["L"])
    assert np.all(g[:10000][~m] == 0) and g[10000] == 0
    # normalisation: 4 pi int_0^h W r^2 dr = 1 with the table lookup
    rr = np.linspace(0, h, 200001)
    wl = o.kernel_eval("Ws", np.stack([rr, 0 * rr, 0 * rr], 1))
    assert abs(4 * math.pi * np.trapezoid(wl * rr ** 2, rr) - 1) < 2e-3
    # cutoffs, W(0), gradW(0)
    assert o.kernel_eval("W", [[h * 1.0001, 0, 0]])[0] == 0.0
    assert o.kernel_eval("W", [[0, 0, 0]])[0] == pytest.approx(0.5 * (w[0] + w[1]))
    assert np.all(o.kernel_eval("gradW", [[0, 0, 0]]) == 0)
    gv = o.kernel_eval("gradW", [[0.3 * h, 0.2 * h, -0.1 * h]])[0]
    assert np.allclose(np.cross(gv, [0.3, 0.2, -0.1]), 0, atol=1e-12 * s["L"])  # parallel to r
    assert np.dot(gv, [0.3, 0.2, -0.1]) < 0  # W decreases outward


def test_fp32_kernel_table_rounds_the_same_formula():
    a = _sim(fp64=False).kernel_table()
    b = _sim(fp64=True).kernel_table()
    assert np.abs(a[0] - b[0]).max() <= 2e-6 * b[0].max()


# ---- neighbour search ------------------------------------------------------------------------------------------
def test_morton_known_answers():
    assert O.morton3(1, 0, 0) == 1 and O.morton3(0, 1, 0) == 2 and O.morton3(0, 0, 1) == 4
    assert O.morton3(1023, 1023, 1023) == 2 ** 30 - 1
    rng = np.random.default_rng(0)
    for _ in range(200):
        x, y, z = (int(v) for v in rng.integers(0, 1024, 3))
        m = 0
        for b in range(10):
            m |= ((x >> b) & 1) << (3 * b) | ((y >> b) & 1) << (3 * b + 1) | ((z >> b) & 1) << (3 * b + 2)
        assert O.morton3(x, y, z) == m


def _search(x, r):
    d = DFSPHSimulationDescription(ParticleRadius=r, EnableViscositySolver=0, EnableSurfaceTensionSolver=0)
    o = O.OracleSim(d)
    o.set_fluid(x, np.zeros_like(x))
    o.step_until("search")
    return o.neighbors()


@pytest.mark.parametrize("seed", [0, 1])
def test_neighbor_sets_match_brute_force(seed):
    rng = np.random.default_rng(seed)
    r = 0.01
    x = rng.uniform(-0.2, 0.25, size=(6000, 3)).astype(np.float32)
    counts, offsets, ids = _search(x, r)
    bc, bid = O.brute_force_neighbors(x, 4 * r)
    assert counts.max() < 70
    assert np.array_equal(counts, bc)
    for i in range(len(x)):
        got = np.sort(ids[offsets[i]: offsets[i] + counts[i]])
        assert np.array_equal(got, np.sort(bid[i, :bc[i]])), i


def test_neighbor_cutoff_is_strict_and_symmetric():
    r = 0.025
    R = np.float32(4 * r)
    # pairs exactly at R (excluded), just inside (included), coincident (excluded: d^2 > 0)
    x = np.array([[0, 0, 0], [R, 0, 0], [0, np.nextafter(R, np.float32(0)), 0], [0, 0, 0]], np.float32)
    counts, offsets, ids = _search(x, r)
    lists = [set(ids[offsets[i]: offsets[i] + counts[i]].tolist()) for i in range(4)]
    assert lists[0] == {2} and lists[3] == {2} and lists[1] == set()
    assert 0 in lists[2] and 3 in lists[2]


def test_neighbor_cap_keeps_subset_of_70():
    rng = np.random.default_rng(3)
    r = 0.01
    x = rng.uniform(0, 0.1, size=(5000, 3)).astype(np.float32)
    counts, offsets, ids = _search(x, r)
    bc, bid = O.brute_force_neighbors(x, 4 * r, cap=5000)
    assert np.all(counts == np.minimum(bc, 70))
    for i in range(0, len(x), 97):
        assert set(ids[offsets[i]: offsets[i] + counts[i]].tolist()) <= set(bid[i, :bc[i]].tolist())


# ---- volume maps -----------------------------------------------------------------------------------------------
def _corner_xi():
    xi = []
    for c in range(8):
        xi.append([(-1, 1)[c & 1], (-1, 1)[(c >> 1) & 1], (-1, 1)[(c >> 2) & 1]])
    for iy in range(2):  # x edges: 8 + 2*(iz + 2*iy) + side
        for iz in range(2):
            for side in range(2):
                xi.append([(-1 / 3, 1 / 3)[side], (-1, 1)[iy], (-1, 1)[iz]])
    for iz in range(2):  # y edges: 16 + 2*(ix + 2*iz) + side
        for ix in range(2):
            for side in range(2):
                xi.append([(-1, 1)[ix], (-1 / 3, 1 / 3)[side], (-1, 1)[iz]])
    for ix in range(2):  # z edges: 24 + 2*(iy + 2*ix) + side
        for iy in range(2):
            for side in range(2):
                xi.append([(-1, 1)[ix], (-1, 1)[iy], (-1 / 3, 1 / 3)[side]])
    return np.array(xi)


def test_shape_functions_interpolate_cubics():
    nodes = _corner_xi()
    # nodal property
    for k, xi in enumerate(nodes):
        N, _ = O.shape_function(xi)
        e = np.zeros(32)
        e[k] = 1
        assert np.allclose(N, e, atol=1e-14), k
    rng = np.random.default_rng(5)
    mon = [(a, b, c) for a in range(4) for b in range(4) for c in range(4) if a + b + c <= 3]
    for _ in range(20):
        coef = rng.normal(size=len(mon))
        f = lambda p: sum(cf * p[0] ** a * p[1] ** b * p[2] ** c for cf, (a, b, c) in zip(coef, mon))
        fn = np.array([f(p) for p in nodes])
        xi = rng.uniform(-1, 1, 3)
        N, dN = O.shape_function(xi)
        assert abs(N @ fn - f(xi)) < 1e-12
        assert abs(N.sum() - 1) < 1e-14 and np.abs(dN.sum(0)).max() < 1e-13
        fd = np.array([(f(xi + e) - f(xi - e)) / 2e-6 for e in np.eye(3) * 1e-6])
        assert np.allclose(dN.T @ fn, fd, atol=1e-7)


def test_shape_function_hessian_matches_differences():
    """Second derivatives used by the adjoint's boundary terms: a cubic's interpolant is the cubic, so its Hessian is
    the cubic's own (closed form per monomial); the Hessians sum to zero (partition of unity); and they are the
    central differences of the first derivatives."""
    nodes = _corner_xi()
    rng = np.random.default_rng(6)
    mon = [(a, b, c) for a in range(4) for b in range(4) for c in range(4) if a + b + c <= 3]
    pairs = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]

    def d2(p, cf_mon, i, j):
        s = 0.0
        for cf, e in cf_mon:
            e = list(e)
            c = cf
            for ax in (i, j):
                c *= e[ax]
                e[ax] -= 1
                if e[ax] < 0:
                    c = 0.0
                    break
            if c:
                s += c * p[0] ** e[0] * p[1] ** e[1] * p[2] ** e[2]
        return s

    for _ in range(20):
        coef = rng.normal(size=len(mon))
        f = lambda p: sum(cf * p[0] ** a * p[1] ** b * p[2] ** c for cf, (a, b, c) in zip(coef, mon))
        fn = np.array([f(p) for p in nodes])
        xi = rng.uniform(-1, 1, 3)
        H = O.shape_function_hess(xi)
        assert np.abs(H.sum(0)).max() < 1e-12
        for k, (i, j) in enumerate(pairs):
            assert abs(H[:, k] @ fn - d2(xi, list(zip(coef, mon)), i, j)) < 1e-11, (i, j)
        for k, (i, j) in enumerate(pairs):
            e = np.eye(3)[j] * 1e-6
            fd = (O.shape_function(xi + e)[1][:, i] - O.shape_function(xi - e)[1][:, i]) / 2e-6
            assert np.allclose(H[:, k], fd, atol=1e-8), (i, j)


def test_cell_node_layout_matches_local_coordinates():
    res = (3, 4, 5)
    dmin, dmax = np.array([-1.0, 0.5, 2.0]), np.array([2.0, 2.5, 4.5])
    cs = (dmax - dmin) / res
    xi = _corner_xi()
    assert O.node_count(res) == 4 * 5 * 6 + 2 * (3 * 5 * 6 + 4 * 4 * 6 + 4 * 5 * 5)
    for (i, j, k) in [(0, 0, 0), (2, 3, 4), (1, 2, 3)]:
        nodes = O.cell_nodes(res, i, j, k)
        assert len(set(nodes.tolist())) == 32
        origin = dmin + np.array([i, j, k]) * cs
        for q in range(32):
            p = O.node_position(dmin, dmax, res, int(nodes[q]))
            assert np.allclose(p, origin + (xi[q] + 1) / 2 * cs, atol=1e-12), (i, j, k, q)


def _box_scene(fp64=True, res=(20, 20, 20)):
    d = DFSPHSimulationDescription(ParticleRadius=0.025, EnableViscositySolver=0, EnableSurfaceTensionSolver=0)
    o = O.OracleSim(d, fp64=fp64)
    x = np.zeros((1, 3), np.float32)
    o.set_fluid(x, x)
    b = RigidBodyDescription("box", half_extents=(1, 1, 1), inverted=True, resolution=res)
    o.add_boundary(b.to_abi())
    return o


def test_volume_map_distance_field_is_exact():
    o = _box_scene()
    nodes, geom = o.boundary_map(0, 0)
    res = geom[6:9].astype(int)
    idx = np.linspace(0, len(nodes) - 1, 400).astype(int)
    pts = np.array([O.node_position(geom[:3], geom[3:6], res, int(i)) for i in idx])
    q = np.abs(pts) - 1
    d = np.linalg.norm(np.maximum(q, 0), axis=1) + np.minimum(q.max(1), 0)
    assert np.allclose(nodes[idx], -(d + 0.025), atol=1e-12)
    # inside a face region the distance is linear, which the cubic map reproduces
    p = np.array([[-0.97, 0.1, 0.2], [0.3, 0.99, -0.5], [0.2, -0.3, 0.95]])
    got = o.map_interpolate(0, 0, p)
    assert np.allclose(got[:, 0], [0.005, -0.015, 0.025], atol=1e-12)
    assert np.allclose(got[:, 1:], [[1, 0, 0], [0, -1, 0], [0, 0, -1]], atol=1e-9)


def test_volume_field_plane_closed_form():
    """Plane wall, phi(x + xi) = phi(x) + xi_y exactly. (a) Every node equals an independent numpy evaluation of
    0.8 * sum_ijk w_i w_j w_k R^3 gamma(phi + R t_j) [|R t| <= R] (16-point Gauss-Legendre, RigidBody.cu:44-72),
    gamma = 1 (phi <= 0), W(phi)/W(0) (0 < phi < R), 0 beyond, and 0 when phi(x) > 2R. (b) That rule stays within 5 %
    of the continuous integral 0.8 int_{-R}^{R} gamma(phi + t) pi (R^2 - t^2) dt."""
    r = 0.025
    R = float(4 * np.float32(r))
    d = DFSPHSimulationDescription(ParticleRadius=r, EnableViscositySolver=0, EnableSurfaceTensionSolver=0)
    o = O.OracleSim(d, fp64=True)
    x = np.zeros((1, 3), np.float32)
    o.set_fluid(x, x)
    b = RigidBodyDescription("plane", center=(0, 0, 0), normal=(0, 1, 0), resolution=(4, 24, 4),
                             domain=((-0.2, -0.3, -0.2), (0.2, 0.3, 0.2)))
    o.add_boundary(b.to_abi())
    w, g, s = o.kernel_table()

    def gamma(phi):
        p = np.minimum((np.abs(phi) * s["InvStepSize"]).astype(int), 9998)
        wv = np.where(np.abs(phi) <= R, 0.5 * (w[p] + w[p + 1]), 0.0) / s["WZero"]
        return np.where(phi <= 0, 1.0, np.where(phi < R, wv, 0.0))

    gx, gw = O.gauss_legendre(16)
    T = np.stack(np.meshgrid(gx, gx, gx, indexing="ij"), -1) * R
    Wt = gw[:, None, None] * gw[None, :, None] * gw[None, None, :]
    inside = (T ** 2).sum(-1) <= R * R
    t = np.linspace(-R, R, 200001)
    nodes, geom = o.boundary_map(0, 1)
    phi_nodes, _ = o.boundary_map(0, 0)
    res = geom[6:9].astype(int)
    checked = 0
    for i in range(len(nodes)):
        phi = phi_nodes[i]
        pos = O.node_position(geom[:3], geom[3:6], res, i)
        if np.any(pos - R < geom[:3]) or np.any(pos + R > geom[3:6]):
            continue  # support ball leaves the map domain (outside points contribute 0)
        rule = 0.0 if phi > 2 * R else 0.8 * R ** 3 * (Wt * inside * gamma(phi + T[..., 1])).sum()
        assert nodes[i] == pytest.approx(rule, rel=1e-10, abs=1e-16), (i, phi)
        if phi < 1.5 * R:
            exact = 0.8 * np.trapezoid(gamma(phi + t) * np.pi * (R * R - t * t), t)
            assert rule == pytest.approx(exact, rel=0.05, abs=2e-2 * 0.8 * 4 / 3 * np.pi * R ** 3), (i, phi)
        checked += 1
    assert checked > 50


# ---- densities and factors -------------------------------------------------------------------------------------
def _lattice_state(fp64, spacing_factor=2.1, n=9, r=0.025):
    d = DFSPHSimulationDescription(ParticleRadius=r, EnableViscositySolver=0, EnableSurfaceTensionSolver=0,
                                   Gravity=(0, 0, 0))
    o = O.OracleSim(d, fp64=fp64)
    h = spacing_factor * r
    idx = np.arange(n) - n // 2
    X, Y, Z = np.meshgrid(idx, idx, idx, indexing="ij")
    x = (np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1) * h).astype(np.float32)
    o.set_fluid(x, np.zeros_like(x))
    return o, x


@pytest.mark.parametrize("fp64", [True, False])
def test_lattice_density_and_factor_closed_form(fp64):
    r = 0.025
    o, x = _lattice_state(fp64)
    o.step_until("factor")
    w, g, s = o.kernel_table()
    c = len(x) // 2  # centre particle
    d = x.astype(np.float64) - x[c].astype(np.float64)
    dist = np.linalg.norm(d, axis=1)
    nb = (dist > 0) & (dist < 4 * r)
    p = np.minimum((dist * s["InvStepSize"]).astype(int), 9998)
    wv = 0.5 * (w[p] + w[p + 1])
    gv = 0.5 * (g[p] + g[p + 1])
    r32 = float(np.float32(r))  # the description holds r in float
    V = 0.8 * (2 * r32) ** 3
    rho = 1000 * (V * s["WZero"] + V * wv[nb].sum())
    gj = -V * gv[nb, None] * (-d[nb])  # gradW(x_c - x_j) = G * (x_c - x_j)
    alpha = 1.0 / ((gj ** 2).sum() + ((-gj.sum(0)) ** 2).sum())
    tol = 1e-12 if fp64 else 2e-6
    assert o.field("density")[c] == pytest.approx(rho, rel=tol)
    assert o.field("factor")[c] == pytest.approx(alpha, rel=tol)
    assert nb.sum() == o.neighbors()[0][c]


def test_pressure_forces_conserve_momentum():
    """Compressed lattice, no gravity or walls: the pair terms (kappa_i + kappa_j) V gradW_ij are antisymmetric, so
    the divergence and density solves leave the total momentum at zero."""
    o, x = _lattice_state(True, spacing_factor=1.7, n=7)
    o.step()
    v = o.field("velocity")
    assert np.abs(v).max() > 1e-3  # the solve did act
    assert np.abs(v.sum(0)).max() < 1e-12 * len(v) * np.abs(v).max()


# ---- viscosity -------------------------------------------------------------------------------------------------
def _visc_scene(n=4, tol=1e-6, max_it=1000, boundary=False, vscale=0.1):
    r = 0.02
    rng = np.random.default_rng(9)
    x = scenes.lattice_block((-0.05, -0.05, -0.05), (n, n, n), 2 * r, 0.05 * r, rng)
    v = rng.normal(size=x.shape).astype(np.float32) * vscale
    d = DFSPHSimulationDescription(ParticleRadius=r, Viscosity=0.5, BoundaryViscosity=0.3, EnableSurfaceTensionSolver=0,
                                   MaxViscositySolverError=tol, MaxViscositySolverIterations=max_it,
                                   Gravity=(0, 0, 0))
    o = O.OracleSim(d, fp64=True)
    o.set_fluid(x, v)
    if boundary:
        o.add_boundary(RigidBodyDescription("box", half_extents=(0.11, 0.11, 0.11), inverted=True,
                                            resolution=(8, 8, 8)).to_abi())
    return o, x, v


def test_viscosity_matvec_closed_form():
    """A x = x - dt/rho_i sum_j 10 mu (m/rho_j) ((x_i - x_j).r_ij) / (|r_ij|^2 + 0.01 h^2) gradW(r_ij)."""
    o, x, v = _visc_scene(n=3)
    o.step_until("viscosity")
    A, b = o.viscosity_system()
    info = o.info()
    rho = o.field("density")
    pos = o.field("position")
    n = len(pos)
    rng = np.random.default_rng(1)
    u = rng.normal(size=(n, 3))
    cnt, off, ids = o.neighbors()
    h2 = info["SupportRadius2"]
    y = u.copy()
    for i in range(n):
        acc = np.zeros(3)
        for j in ids[off[i]: off[i] + cnt[i]]:
            rij = pos[i] - pos[j]
            gw = o.kernel_eval("gradW", rij[None])[0]
            acc += 10 * info["DynamicViscosity"] * (info["ParticleMass"] / rho[j]) * np.dot(u[i] - u[j], rij) / (
                rij @ rij + 0.01 * h2) * gw
        y[i] = u[i] - info["TimeStepSize"] / rho[i] * acc
    assert np.allclose(A @ u.ravel(), y.ravel(), rtol=1e-12, atol=1e-12)
    assert np.allclose(b, o.field("velocity").ravel())


def test_viscosity_pcg_solves_the_system():
    o, x, v = _visc_scene(n=4)
    o.step_until("viscosity")
    A, b = o.viscosity_system()
    xs = o.field("viscosity_solution").ravel()
    dbg = o.debug_info()
    assert dbg.ViscositySolverIterationCount > 2
    assert np.linalg.norm(A @ xs - b) <= 1.01e-6 * np.linalg.norm(b)
    assert np.allclose(xs, np.linalg.solve(A, b), rtol=0, atol=1e-5 * np.abs(b).max())
    # preconditioner: inverse of the diagonal 3x3 blocks (no boundary)
    P = o.field("viscosity_preconditioner").reshape(-1, 3, 3)
    for i in range(0, len(P), 7):
        assert np.allclose(P[i], np.linalg.inv(A[3 * i: 3 * i + 3, 3 * i: 3 * i + 3]), rtol=1e-10, atol=1e-12)


def test_viscosity_quirks():
    """Min iterations > 0 disables the CG loop (DFSPHImplementation.cu:693); a zero rhs zeroes the solution
    (:654-659)."""
    r = 0.02
    rng = np.random.default_rng(2)
    x = scenes.lattice_block((0, 0, 0), (4, 4, 4), 2 * r, 0.05 * r, rng)
    v = rng.normal(size=x.shape).astype(np.float32)
    d = DFSPHSimulationDescription(ParticleRadius=r, Viscosity=0.5, EnableSurfaceTensionSolver=0,
                                   MinViscositySolverIterations=1, Gravity=(0, 0, 0))
    o = O.OracleSim(d, fp64=True)
    o.set_fluid(x, v)
    o.step_until("viscosity")
    assert o.debug_info().ViscositySolverIterationCount == 0
    d2 = DFSPHSimulationDescription(ParticleRadius=r, Viscosity=0.5, EnableSurfaceTensionSolver=0, Gravity=(0, 0, 0))
    o2 = O.OracleSim(d2, fp64=True)
    o2.set_fluid(x, np.zeros_like(x))
    o2.step_until("viscosity")
    assert np.all(o2.field("viscosity_solution") == 0) and o2.debug_info().ViscositySolverError == 0


# ---- time step, Jacobi loop semantics ---------------------------------------------------------------------------
def test_time_step_closed_form():
    o, x, v = _visc_scene(n=4, tol=0.1, vscale=5.0)
    o.step_until("viscosity")
    vv, aa = o.field("velocity"), o.field("acceleration")
    dt0 = o.info()["TimeStepSize"]
    o2, _, _ = _visc_scene(n=4, tol=0.1, vscale=5.0)
    o2.step_until("time_step")
    m = max(0.1, ((vv + aa * dt0) ** 2).sum(1).max())
    r = float(np.float32(0.02))
    desc_dt = min(max(0.4 * 2 * r / math.sqrt(m), float(np.float32(0.0001))), float(np.float32(0.005)))
    assert desc_dt < 0.004  # the CFL branch, not the clamp
    assert o2.info()["TimeStepSize"] == pytest.approx(desc_dt, rel=1e-12)
    assert o2.debug_info().MaxVelocityMagnitude == pytest.approx(m, rel=1e-6)


def test_jacobi_minimum_iterations_quirk():
    """With Min iterations 0 the Jacobi loops run zero iterations (DFSPHImplementation.cu:458-461, 529-532)."""
    o, x = _lattice_state(True, spacing_factor=1.8, n=5)
    o.step()
    assert o.debug_info().PressureSolverIterationCount == 0
    assert o.debug_info().DivergenceSolverIterationCount == 0


# ---- surface tension -------------------------------------------------------------------------------------------
def _ball(radius_cells, r=0.02):
    h = 2 * r
    m = radius_cells
    idx = np.arange(-m, m + 1)
    X, Y, Z = np.meshgrid(idx, idx, idx, indexing="ij")
    p = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1) * h
    p = p[np.linalg.norm(p, axis=1) <= m * h]
    return p.astype(np.float32)


def test_surface_tension_sphere():
    r = 0.02
    out = {}
    for m in (4, 7):
        x = _ball(m, r)
        d = DFSPHSimulationDescription(ParticleRadius=r, EnableViscositySolver=0, EnableSurfaceTensionSolver=1,
                                       SurfaceTension=0.05, CSDFix=64, Gravity=(0, 0, 0))
        o = O.OracleSim(d, fp64=True)
        o.set_fluid(x, np.zeros_like(x))
        o.step_until("surface_tension")
        n = o.field("mc_surface_normal")
        surf = np.abs(n).sum(1) > 0
        rad = np.linalg.norm(x, axis=1)
        assert not surf[rad < (m - 2.5) * 2 * r].any()  # deep interior is never surface
        assert surf[rad > (m - 0.5) * 2 * r].mean() > 0.9  # outer shell is
        outward = (n[surf] * x[surf]).sum(1) > 0
        assert outward.mean() > 0.95
        k = o.field("mc_surface_curvature_smooth")[surf & (rad > (m - 0.5) * 2 * r)]
        out[m] = k.mean()
        a = o.field("acceleration")[surf & (rad > (m - 0.5) * 2 * r)]
        assert ((a * x[surf & (rad > (m - 0.5) * 2 * r)]).sum(1) < 0).mean() > 0.9  # pulls inward
    assert out[4] > out[7] > 0  # smaller drop, larger curvature
