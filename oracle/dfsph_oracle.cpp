// dfsph_oracle.cpp - CPU oracle for the DFSPH step. TEST INFRASTRUCTURE ONLY.
//
// This file restates, on the CPU, the algorithm of VFD's GPU DFSPH step (Goubermouche/VFD; the reference has no CPU
// path of its own). It exists to check the HIP product: only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load it, and only as the result they compare against. The product (vfd_amd/) never links or calls it.
//
// One template, two instantiations:
//   T = float   "FP32 mirror": the reference's operation order in IEEE single precision, compiled with
//               -ffp-contract=off. Every fused multiply-add is an explicit std::fma at the places named below.
//   T = double  "FP64 twin": the same code in double precision, used to state each field's FP32 error.
//
// Arithmetic contract shared with the HIP kernels (the reference's nvcc --use_fast_math build leaves these
// unpinned; this implementation fixes them):
//   dot3(a,b)  = fma(a.z,b.z, fma(a.y,b.y, a.x*b.x))      (every dot, length2 and neighbour-search distance)
//   length(a)  = sqrt(dot3(a,a)), IEEE sqrt and division (no fast math)
//   reductions: Jacobi residual sums and CG dot products accumulate per-particle float terms in double,
//               sequentially here (the reference uses thrust float reductions; `minus` makes its residual
//               order-dependent, DFSPHImplementation.cu:483-489)
//   neighbour order: 27 cells in z,y,x order (ParticleSearchKernels.cu:301-303), within a cell ascending particle
//               index (the reference's order there follows atomicAdd, ParticleSearchKernels.cu:77)
//
// Parity status of each function is stated in the function comments ("pinned by ..." names the test that pins it).
// The reference adjoint (Sim::reverse_step) is pinned by central finite differences of the smooth-kernel forward
// (tests/test_adjoint_oracle.py).
//
// Parity unpinned (pinnable only by running the reference's CUDA build, which cannot be built here): the order of
// thrust's `minus` reduction for the solver errors, fast-math sqrt/division rounding, and the atomicAdd order inside a
// cell that decides which neighbours fall past the cap of 70. This oracle fixes all three deterministically.

#include <algorithm>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <map>
#include <string>
#include <vector>
#include <type_traits>

#include "../include/vfd.h"

namespace orc {

// Core/Math/Math.h:6 - EPS is a double literal, compared against float values promoted to double.
static const double kEps = 1.0e-5;
static const unsigned kMaxNeighbors = 70u;      // ParticleSearchKernels.cuh:7
static const unsigned kMetaGroup = 8u;          // ParticleSearchKernels.cuh:8
static const unsigned kMetaBlock = 512u;        // ParticleSearchKernels.cuh:9
static const unsigned kTableResolution = 10000u;  // Kernel/DFSPHKernels.h:143
static const unsigned kHaltonCount = 16384u;    // HaltonVec323.cuh: 49152 floats = 16384 points
static const unsigned kHaltonMod = 49152u;      // DFSPHKernels.cu:923

// ---------------------------------------------------------------------------------------------------------------
// Small vector / matrix types (glm semantics where the reference relies on them)
// ---------------------------------------------------------------------------------------------------------------
template <class T> struct V3 {
    T x, y, z;
};
template <class T> inline V3<T> v3(T x, T y, T z) { return V3<T>{x, y, z}; }
template <class T> inline V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> inline V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> inline V3<T> operator-(V3<T> a) { return {-a.x, -a.y, -a.z}; }
template <class T> inline V3<T> operator*(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> inline V3<T> operator*(T s, V3<T> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class T> inline V3<T> operator/(V3<T> a, T s) { return {a.x / s, a.y / s, a.z / s}; }
template <class T> inline V3<T> cmul(V3<T> a, V3<T> b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
template <class T> inline bool is_zero(V3<T> a) { return a.x == T(0) && a.y == T(0) && a.z == T(0); }
template <class T> inline T dot3(V3<T> a, V3<T> b) { return std::fma(a.z, b.z, std::fma(a.y, b.y, a.x * b.x)); }
template <class T> inline T length2(V3<T> a) { return dot3(a, a); }
template <class T> inline T length(V3<T> a) { return std::sqrt(dot3(a, a)); }
// glm::cross
template <class T> inline V3<T> cross(V3<T> a, V3<T> b) {
    return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
// glm::normalize(v) = v * inversesqrt(dot(v,v)), inversesqrt(x) = 1/sqrt(x)
template <class T> inline V3<T> normalize(V3<T> a) { return a * (T(1) / std::sqrt(dot3(a, a))); }

template <class T> struct M3 {  // row-major
    T m[9];
};
template <class T> inline M3<T> mzero() { M3<T> r; for (auto& e : r.m) e = T(0); return r; }

// 3x3 inverse by the adjugate: inv = adj(A) / det(A). Follows glm::inverse's use at DFSPHKernels.cu:683.
template <class T> M3<T> inverse3(const M3<T>& A) {
    const T a = A.m[0], b = A.m[1], c = A.m[2];
    const T d = A.m[3], e = A.m[4], f = A.m[5];
    const T g = A.m[6], h = A.m[7], i = A.m[8];
    const T c00 = e * i - f * h, c01 = f * g - d * i, c02 = d * h - e * g;
    const T det = a * c00 + b * c01 + c * c02;
    const T s = T(1) / det;
    M3<T> r;
    r.m[0] = c00 * s;
    r.m[1] = (c * h - b * i) * s;
    r.m[2] = (b * f - c * e) * s;
    r.m[3] = c01 * s;
    r.m[4] = (a * i - c * g) * s;
    r.m[5] = (c * d - a * f) * s;
    r.m[6] = c02 * s;
    r.m[7] = (b * g - a * h) * s;
    r.m[8] = (a * e - b * d) * s;
    return r;
}
template <class T> inline V3<T> mulmv(const M3<T>& A, V3<T> v) {  // glm mat3 * vec3 (row i = sum_j A_ij v_j)
    return {A.m[0] * v.x + A.m[1] * v.y + A.m[2] * v.z, A.m[3] * v.x + A.m[4] * v.y + A.m[5] * v.z,
            A.m[6] * v.x + A.m[7] * v.y + A.m[8] * v.z};
}
template <class T> inline M3<T> mmul(const M3<T>& A, const M3<T>& B) {
    M3<T> C;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            C.m[3 * r + c] = A.m[3 * r] * B.m[c] + A.m[3 * r + 1] * B.m[3 + c] + A.m[3 * r + 2] * B.m[6 + c];
    return C;
}
template <class T> inline M3<T> skew(V3<T> a) {  // skew(a) b = cross(a, b)
    M3<T> S;
    S.m[0] = T(0), S.m[1] = -a.z, S.m[2] = a.y;
    S.m[3] = a.z, S.m[4] = T(0), S.m[5] = -a.x;
    S.m[6] = -a.y, S.m[7] = a.x, S.m[8] = T(0);
    return S;
}
template <class T> inline M3<T> mlin(T a, const M3<T>& A, T b, const M3<T>& B) {  // a A + b B
    M3<T> C;
    for (int e = 0; e < 9; ++e) C.m[e] = a * A.m[e] + b * B.m[e];
    return C;
}
template <class T> inline M3<T> normalize_jacobian(V3<T> u, T len) {  // d(a / |a|)/da = (I - u u^T) / |a|, u = a / |a|
    const T uv[3] = {u.x, u.y, u.z};
    M3<T> J;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) J.m[3 * r + c] = ((r == c ? T(1) : T(0)) - uv[r] * uv[c]) / len;
    return J;
}
template <class T> inline V3<T> mulmtv(const M3<T>& A, V3<T> v) {  // A^T v
    return {A.m[0] * v.x + A.m[3] * v.y + A.m[6] * v.z, A.m[1] * v.x + A.m[4] * v.y + A.m[7] * v.z,
            A.m[2] * v.x + A.m[5] * v.y + A.m[8] * v.z};
}

// Math.h:19-33 GetOrthogonalVectors
template <class T> inline void orthogonal_vectors(V3<T> vec, V3<T>& x, V3<T>& y) {
    V3<T> v = v3<T>(1, 0, 0);
    if ((double)std::fabs(dot3(v, vec)) > 0.999) v = v3<T>(0, 1, 0);
    x = cross(vec, v);
    y = cross(vec, x);
    x = normalize(x);
    y = normalize(y);
}

// ---------------------------------------------------------------------------------------------------------------
// Tabulated cubic spline kernel. Kernel/DFSPHKernels.h:7-143 (SetRadius :13-41, GetW :43-65, GetGradientW :67-80,
// CalculateW :86-104, CalculateGradientW :110-131). Pinned by tests/test_oracle_pins.py
// (test_kernel_table_closed_form: closed-form spline and gradient at both radii;
// test_fp32_kernel_table_rounds_the_same_formula) and by GPU bit-equality of the tables.
// ---------------------------------------------------------------------------------------------------------------
template <class T> struct CubicKernel {
    std::vector<T> W, G;  // W[10000], G[10001]
    T R = 0, R2 = 0, invStep = 0, W0 = 0, K = 0, L = 0;

    T calcW(T r) const {
        T res = 0;
        const T q = r / R;
        if (q <= T(1)) {
            if (q <= T(0.5)) {
                const T q2 = q * q;
                const T q3 = q2 * q;
                res = K * (T(6) * q3 - T(6) * q2 + T(1));
            } else {
                const T q1 = T(1) - q;
                const T q3 = q1 * q1 * q1;
                res = K * (T(2) * q3);
            }
        }
        return res;
    }
    V3<T> calcGradW(V3<T> r) const {
        const T rl = std::sqrt(dot3(r, r));
        const T q = rl / R;
        if (rl > T(1.0e-5) && q <= T(1)) {
            const V3<T> gq = r * (T(1) / (rl * R));
            if (q <= T(0.5)) {
                const T s = L * q * (T(3) * q - T(2));
                return s * gq;
            }
            const T f = T(1) - q;
            const T s = L * (-f * f);
            return s * gq;
        }
        return v3<T>(0, 0, 0);
    }
    void setRadius(T radius) {
        const T pi = T(3.14159265358979323846);
        R = radius;
        R2 = R * R;
        const T r3 = R * R * R;
        K = T(8) / (pi * r3);
        L = T(48) / (pi * r3);
        W0 = calcW(std::sqrt(dot3(v3<T>(0, 0, 0), v3<T>(0, 0, 0))));
        const T step = R / T(kTableResolution - 1);
        invStep = T(1) / step;
        W.assign(kTableResolution, T(0));
        G.assign(kTableResolution + 1, T(0));
        for (unsigned i = 0; i < kTableResolution; ++i) {
            const T posX = step * T(i);
            W[i] = calcW(posX);
            G[i] = posX > T(1.0e-9) ? calcGradW(v3<T>(posX, 0, 0)).x / posX : T(0);
        }
        G[kTableResolution] = T(0);
    }
    unsigned index(T rl) const {
        const unsigned p = (unsigned)(rl * invStep);
        return p < kTableResolution - 2u ? p : kTableResolution - 2u;
    }
    T getW(T r) const {  // scalar overload, cutoff r <= R
        if (r <= R) {
            const unsigned p = index(r);
            return T(0.5) * (W[p] + W[p + 1]);
        }
        return T(0);
    }
    T getW(V3<T> r) const {  // vector overload, cutoff |r|^2 <= R^2
        if (smooth) return calcW(std::sqrt(dot3(r, r)));
        const T r2 = dot3(r, r);
        if (r2 <= R2) {
            const unsigned p = index(std::sqrt(r2));
            return T(0.5) * (W[p] + W[p + 1]);
        }
        return T(0);
    }
    V3<T> gradW(V3<T> r) const {  // cutoff |r| <= R
        if (smooth) return calcGradW(r);
        const T rl = std::sqrt(dot3(r, r));
        if (rl <= R) {
            const unsigned p = index(rl);
            const T s = T(0.5) * (G[p] + G[p + 1]);
            return s * r;
        }
        return v3<T>(0, 0, 0);
    }

    // ---- adjoint support -------------------------------------------------------------------------------------
    // smooth = true evaluates W and gradW from the spline itself instead of the table (a differentiable forward, used
    // only to pin the adjoint against finite differences).
    bool smooth = false;
    // gradW(r) / r: the table value the forward uses (or the spline's value when smooth)
    T gval(T rl) const {
        if (smooth) return rl > T(1.0e-5) ? calcGradW(v3<T>(rl, 0, 0)).x / rl : T(0);
        if (!(rl <= R)) return T(0);
        const unsigned p = index(rl);
        return T(0.5) * (G[p] + G[p + 1]);
    }
    // derivatives of the spline (the table's smooth counterpart): dW/drho and d(gradW/r)/drho
    T dW(T rl) const {
        const T q = rl / R;
        if (!(q <= T(1))) return T(0);
        if (q <= T(0.5)) return K * (T(18) * q * q - T(12) * q) / R;
        return -T(6) * K * (T(1) - q) * (T(1) - q) / R;
    }
    T dG(T rl) const {
        const T q = rl / R;
        if (!(q <= T(1)) || !(rl > T(1.0e-5))) return T(0);
        if (q <= T(0.5)) return T(3) * L / (R * R * R);
        const T f = T(1) - q;
        return (L / R) * (T(2) * f / (R * rl) + f * f / (rl * rl));
    }
};

// ---------------------------------------------------------------------------------------------------------------
// Halton sphere points. The reference stores 16,384 precomputed points (HaltonVec323.cuh); they are the closed form
// H[i] = (sqrt(1-z^2) cos 2*pi*phi2(i), sqrt(1-z^2) sin 2*pi*phi2(i), z = 1 - 2*phi3(i)), phi_b = radical inverse,
// rounded to 15 decimals as printed in the reference table.
// Pinned by tests/golden/halton_samples.json (values read from the reference table).
// ---------------------------------------------------------------------------------------------------------------
static double radical_inverse(unsigned base, unsigned i) {
    double inv = 1.0 / base, f = inv, r = 0.0;
    while (i > 0) {
        r += f * (double)(i % base);
        i /= base;
        f *= inv;
    }
    return r;
}
template <class T> std::vector<T> halton_table() {
    std::vector<T> h(kHaltonMod);
    const double twopi = 6.283185307179586476925286766559;
    for (unsigned i = 0; i < kHaltonCount; ++i) {
        const double z = 1.0 - 2.0 * radical_inverse(3, i);
        const double rxy = std::sqrt(std::max(0.0, 1.0 - z * z));
        const double a = twopi * radical_inverse(2, i);
        // the reference table holds 15 decimals; rounding there reproduces its values (e.g. 0 for cos(pi/2))
        h[3 * i + 0] = (T)(std::nearbyint(rxy * std::cos(a) * 1e15) / 1e15);
        h[3 * i + 1] = (T)(std::nearbyint(rxy * std::sin(a) * 1e15) / 1e15);
        h[3 * i + 2] = (T)(std::nearbyint(z * 1e15) / 1e15);
    }
    return h;
}

// ---------------------------------------------------------------------------------------------------------------
// Gauss-Legendre rule (GaussQuadrature.cpp tables, p = 30 -> 16 points). Computed by Newton iteration on P_n.
// Pinned by tests/golden/gauss_legendre_16.json (values read from the reference table).
// ---------------------------------------------------------------------------------------------------------------
static void gauss_legendre(int n, std::vector<double>& x, std::vector<double>& w) {
    x.assign(n, 0.0);
    w.assign(n, 0.0);
    const double pi = 3.14159265358979323846;
    for (int i = 0; i < n; ++i) {
        double z = std::cos(pi * (i + 0.75) / (n + 0.5));
        double dp = 0.0;
        for (int it = 0; it < 100; ++it) {
            double p0 = 1.0, p1 = z;
            for (int k = 2; k <= n; ++k) {
                const double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            dp = n * (z * p1 - p0) / (z * z - 1.0);
            const double dz = p1 / dp;
            z -= dz;
            if (std::fabs(dz) < 1e-16) break;
        }
        {
            double p0 = 1.0, p1 = z;
            for (int k = 2; k <= n; ++k) {
                const double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            dp = n * (z * p1 - p0) / (z * z - 1.0);
        }
        x[n - 1 - i] = z;  // ascending order, like the reference table
        w[n - 1 - i] = 2.0 / ((1.0 - z * z) * dp * dp);
    }
    for (int i = 0; i < n / 2; ++i) {  // exact symmetry
        const double xs = 0.5 * (x[n - 1 - i] - x[i]);
        x[i] = -xs;
        x[n - 1 - i] = xs;
        const double ws = 0.5 * (w[i] + w[n - 1 - i]);
        w[i] = ws;
        w[n - 1 - i] = ws;
    }
}

// ---------------------------------------------------------------------------------------------------------------
// 32-node cubic serendipity shape functions on [-1,1]^3 (Discregrid layout): SDF.cu:403-...,
// SDFDeviceData.cuh:36-370. Node order: 8 corners (bit0 = x, bit1 = y, bit2 = z), then x-edge, y-edge and z-edge
// nodes at -1/3 and +1/3. Written from the polynomials (corner: (1/64)(1+-x)(1+-y)(1+-z)(9(x^2+y^2+z^2)-19);
// edge: (9/64)(1-s^2)(1+-3s)(1+-a)(1+-b)). Pinned by tests/test_oracle_pins.py
// (test_shape_functions_interpolate_cubics: exact reproduction of cubics and their gradients;
// test_cell_node_layout_matches_local_coordinates).
// ---------------------------------------------------------------------------------------------------------------
template <class T> void shape_function(V3<T> xi, T N[32], V3<T> dN[32]) {
    const T x = xi.x, y = xi.y, z = xi.z;
    const T x2 = x * x, y2 = y * y, z2 = z * z;
    const T mx = T(1) - x, px = T(1) + x, my = T(1) - y, py = T(1) + y, mz = T(1) - z, pz = T(1) + z;
    const T lx[2] = {mx, px}, ly[2] = {my, py}, lz[2] = {mz, pz};
    const T sg[2] = {T(-1), T(1)};
    // corners
    const T q = T(9) * (x2 + y2 + z2) - T(19);
    for (int c = 0; c < 8; ++c) {
        const int ix = c & 1, iy = (c >> 1) & 1, iz = (c >> 2) & 1;
        const T a = lx[ix], b = ly[iy], d = lz[iz];
        N[c] = T(1) / T(64) * q * a * b * d;
        // d/dx [a b d q] = sx*b*d*q + a*b*d*18x
        dN[c].x = (sg[ix] * q + T(18) * x * a) * b * d / T(64);
        dN[c].y = (sg[iy] * q + T(18) * y * b) * a * d / T(64);
        dN[c].z = (sg[iz] * q + T(18) * z * d) * a * b / T(64);
    }
    // 1D cubic factors for edge nodes at s = -1/3 (m) and s = +1/3 (p): e(s) = (1-s^2)(1 -+ 3s)
    auto edge = [](T s, int side, T& f, T& df) {
        const T t = side == 0 ? T(1) - T(3) * s : T(1) + T(3) * s;
        const T ds = side == 0 ? T(-3) : T(3);
        f = (T(1) - s * s) * t;
        df = T(-2) * s * t + (T(1) - s * s) * ds;
    };
    const T k = T(9) / T(64);
    // x edges: index 8 + 2*(iz + 2*iy) + side   (iy, iz select -1/+1 in y and z)
    for (int iy = 0; iy < 2; ++iy)
        for (int iz = 0; iz < 2; ++iz)
            for (int side = 0; side < 2; ++side) {
                T f, df;
                edge(x, side, f, df);
                const int n = 8 + 2 * (iz + 2 * iy) + side;
                N[n] = k * f * ly[iy] * lz[iz];
                dN[n] = v3<T>(k * df * ly[iy] * lz[iz], k * f * sg[iy] * lz[iz], k * f * ly[iy] * sg[iz]);
            }
    // y edges: index 16 + 2*(ix + 2*iz) + side
    for (int iz = 0; iz < 2; ++iz)
        for (int ix = 0; ix < 2; ++ix)
            for (int side = 0; side < 2; ++side) {
                T f, df;
                edge(y, side, f, df);
                const int n = 16 + 2 * (ix + 2 * iz) + side;
                N[n] = k * f * lx[ix] * lz[iz];
                dN[n] = v3<T>(k * f * sg[ix] * lz[iz], k * df * lx[ix] * lz[iz], k * f * lx[ix] * sg[iz]);
            }
    // z edges: index 24 + 2*(iy + 2*ix) + side
    for (int ix = 0; ix < 2; ++ix)
        for (int iy = 0; iy < 2; ++iy)
            for (int side = 0; side < 2; ++side) {
                T f, df;
                edge(z, side, f, df);
                const int n = 24 + 2 * (iy + 2 * ix) + side;
                N[n] = k * f * lx[ix] * ly[iy];
                dN[n] = v3<T>(k * f * sg[ix] * ly[iy], k * f * lx[ix] * sg[iy], k * df * lx[ix] * ly[iy]);
            }
}

// Second derivatives of the 32 shape functions in the local coordinates, per function (xx, yy, zz, xy, xz, yz):
// what the adjoint of the boundary pass needs for the derivative of the normal of the interpolated distance (no
// reference counterpart: the reference never differentiates its maps twice). Pinned against central differences
// of the first derivatives (test_shape_function_hessian_matches_differences).
template <class T> void shape_function_hess(V3<T> xi, T H[32][6]) {
    const T x = xi.x, y = xi.y, z = xi.z;
    const T lx[2] = {T(1) - x, T(1) + x}, ly[2] = {T(1) - y, T(1) + y}, lz[2] = {T(1) - z, T(1) + z};
    const T sg[2] = {T(-1), T(1)};
    const T q = T(9) * (x * x + y * y + z * z) - T(19);
    const T c = T(1) / T(64);
    // corner: N = c q a b d with a = 1 + sx x etc.; d2N/dx2 = c (18 a + 36 sx x) b d,
    // d2N/dxdy = c (18 x a sy + 18 y sx b + q sx sy) d
    for (int n = 0; n < 8; ++n) {
        const int ix = n & 1, iy = (n >> 1) & 1, iz = (n >> 2) & 1;
        const T a = lx[ix], b = ly[iy], d = lz[iz], sx = sg[ix], sy = sg[iy], sz = sg[iz];
        H[n][0] = c * (T(18) * a + T(36) * sx * x) * b * d;
        H[n][1] = c * (T(18) * b + T(36) * sy * y) * a * d;
        H[n][2] = c * (T(18) * d + T(36) * sz * z) * a * b;
        H[n][3] = c * (T(18) * x * a * sy + T(18) * y * sx * b + q * sx * sy) * d;
        H[n][4] = c * (T(18) * x * a * sz + T(18) * z * sx * d + q * sx * sz) * b;
        H[n][5] = c * (T(18) * y * b * sz + T(18) * z * sy * d + q * sy * sz) * a;
    }
    // edge: N = k e(s) u v with e = (1 - s^2)(1 + t 3s), t = -1 (side 0) or +1; e' = -2 s (1 + 3 t s) + 3 t (1 - s^2),
    // e'' = -2 (1 + 3 t s) - 12 t s
    const T k = T(9) / T(64);
    auto edge = [](T s, int side, T& f, T& df, T& ddf) {
        const T t = side == 0 ? T(-1) : T(1);
        const T u = T(1) + T(3) * t * s;
        f = (T(1) - s * s) * u;
        df = T(-2) * s * u + T(3) * t * (T(1) - s * s);
        ddf = T(-2) * u - T(12) * t * s;
    };
    for (int iy = 0; iy < 2; ++iy)  // x edges: 8 + 2 (iz + 2 iy) + side
        for (int iz = 0; iz < 2; ++iz)
            for (int side = 0; side < 2; ++side) {
                T f, df, ddf;
                edge(x, side, f, df, ddf);
                T* h = H[8 + 2 * (iz + 2 * iy) + side];
                h[0] = k * ddf * ly[iy] * lz[iz];
                h[1] = T(0);
                h[2] = T(0);
                h[3] = k * df * sg[iy] * lz[iz];
                h[4] = k * df * ly[iy] * sg[iz];
                h[5] = k * f * sg[iy] * sg[iz];
            }
    for (int iz = 0; iz < 2; ++iz)  // y edges: 16 + 2 (ix + 2 iz) + side
        for (int ix = 0; ix < 2; ++ix)
            for (int side = 0; side < 2; ++side) {
                T f, df, ddf;
                edge(y, side, f, df, ddf);
                T* h = H[16 + 2 * (ix + 2 * iz) + side];
                h[0] = T(0);
                h[1] = k * ddf * lx[ix] * lz[iz];
                h[2] = T(0);
                h[3] = k * df * sg[ix] * lz[iz];
                h[4] = k * f * sg[ix] * sg[iz];
                h[5] = k * df * lx[ix] * sg[iz];
            }
    for (int ix = 0; ix < 2; ++ix)  // z edges: 24 + 2 (iy + 2 ix) + side
        for (int iy = 0; iy < 2; ++iy)
            for (int side = 0; side < 2; ++side) {
                T f, df, ddf;
                edge(z, side, f, df, ddf);
                T* h = H[24 + 2 * (iy + 2 * ix) + side];
                h[0] = T(0);
                h[1] = T(0);
                h[2] = k * ddf * lx[ix] * ly[iy];
                h[3] = k * f * sg[ix] * sg[iy];
                h[4] = k * df * sg[ix] * ly[iy];
                h[5] = k * df * lx[ix] * sg[iy];
            }
}

// ---------------------------------------------------------------------------------------------------------------
// Analytic signed distances (stand-in for MeshDistance, Utility/SDF/MeshDistance.cpp). d < 0 inside the shape.
// ---------------------------------------------------------------------------------------------------------------
template <class T> T shape_distance(const vfd_boundary_desc& b, V3<T> p) {
    const V3<T> c = v3<T>((T)b.Center[0], (T)b.Center[1], (T)b.Center[2]);
    switch (b.Shape) {
        case VFD_SHAPE_BOX: {
            const V3<T> d = p - c;
            const V3<T> q = v3<T>(std::fabs(d.x) - (T)b.HalfExtents[0], std::fabs(d.y) - (T)b.HalfExtents[1],
                                  std::fabs(d.z) - (T)b.HalfExtents[2]);
            const V3<T> qp = v3<T>(std::max(q.x, T(0)), std::max(q.y, T(0)), std::max(q.z, T(0)));
            return std::sqrt(dot3(qp, qp)) + std::min(std::max(q.x, std::max(q.y, q.z)), T(0));
        }
        case VFD_SHAPE_SPHERE: {
            const V3<T> d = p - c;
            return std::sqrt(dot3(d, d)) - (T)b.Radius;
        }
        case VFD_SHAPE_PLANE: {
            const V3<T> n = v3<T>((T)b.Normal[0], (T)b.Normal[1], (T)b.Normal[2]);
            return dot3(p - c, n);
        }
        case VFD_SHAPE_CYLINDER: {
            const V3<T> d = p - c;
            const T rxz = std::sqrt(std::fma(d.z, d.z, d.x * d.x)) - (T)b.Radius;
            const T ry = std::fabs(d.y) - (T)b.HalfExtents[1];
            const T ox = std::max(rxz, T(0)), oy = std::max(ry, T(0));
            return std::sqrt(std::fma(oy, oy, ox * ox)) + std::min(std::max(rxz, ry), T(0));
        }
    }
    return T(0);
}

template <class T> struct Box {
    V3<T> min, max;
    bool contains(V3<T> v) const {  // BoundingBox.h:42-44
        return min.x <= v.x && min.y <= v.y && min.z <= v.z && max.x >= v.x && max.y >= v.y && max.z >= v.z;
    }
};

// ---------------------------------------------------------------------------------------------------------------
// Volume map: SDF(domain,res) SDF.cu:8-14, AddFunction :45-139, host Interpolate :141-215, IndexToNodePosition
// :313-373, device DetermineShapeFunction / Interpolate SDFDeviceData.cuh:372-474. Field 0 = signed distance
// sign*(d - (Padding - r)) (RigidBody.cu:39-42), field 1 = boundary volume 0.8 * GL16^3 integral (RigidBody.cu:44-72,
// GaussQuadrature.cpp:5621-5655). Pinned by tests/test_oracle_pins.py (test_volume_map_distance_field_is_exact:
// node values = analytic distance; test_volume_field_plane_closed_form: closed-form plane volume).
// ---------------------------------------------------------------------------------------------------------------
template <class T> struct VolumeMap {
    Box<T> domain;
    unsigned res[3] = {0, 0, 0};
    V3<T> cellSize, cellSizeInv;
    unsigned cellCount = 0, nodeCount = 0;
    std::vector<T> nodes[2];

    void init(Box<T> d, const unsigned r[3]) {
        domain = d;
        res[0] = r[0];
        res[1] = r[1];
        res[2] = r[2];
        const V3<T> diag = domain.max - domain.min;
        cellSize = v3<T>(diag.x / T(res[0]), diag.y / T(res[1]), diag.z / T(res[2]));
        cellSizeInv = v3<T>(T(1) / cellSize.x, T(1) / cellSize.y, T(1) / cellSize.z);
        cellCount = res[0] * res[1] * res[2];
        const unsigned nv = (res[0] + 1) * (res[1] + 1) * (res[2] + 1);
        const unsigned nex = res[0] * (res[1] + 1) * (res[2] + 1);
        const unsigned ney = (res[0] + 1) * res[1] * (res[2] + 1);
        const unsigned nez = (res[0] + 1) * (res[1] + 1) * res[2];
        nodeCount = nv + 2 * (nex + ney + nez);
    }
    V3<T> node_position(unsigned i) const {
        const unsigned nx = res[0], ny = res[1], nz = res[2];
        const unsigned nv = (nx + 1) * (ny + 1) * (nz + 1);
        const unsigned nex = nx * (ny + 1) * (nz + 1), ney = (nx + 1) * ny * (nz + 1);
        V3<T> idx;
        V3<T> r;
        if (i < nv) {
            idx.z = T(i / ((ny + 1) * (nx + 1)));
            const unsigned t = i % ((ny + 1) * (nx + 1));
            idx.y = T(t / (nx + 1));
            idx.x = T(t % (nx + 1));
            r = domain.min + cmul(cellSize, idx);
        } else if (i < nv + 2 * nex) {
            i -= nv;
            const unsigned e = i / 2;
            idx.z = T(e / ((ny + 1) * nx));
            const unsigned t = e % ((ny + 1) * nx);
            idx.y = T(t / nx);
            idx.x = T(t % nx);
            r = domain.min + cmul(cellSize, idx);
            r.x += (T(1) + T(i % 2)) / T(3) * cellSize.x;
        } else if (i < nv + 2 * (nex + ney)) {
            i -= nv + 2 * nex;
            const unsigned e = i / 2;
            idx.x = T(e / ((nz + 1) * ny));
            const unsigned t = e % ((nz + 1) * ny);
            idx.z = T(t / ny);
            idx.y = T(t % ny);
            r = domain.min + cmul(cellSize, idx);
            r.y += (T(1) + T(i % 2)) / T(3) * cellSize.y;
        } else {
            i -= nv + 2 * (nex + ney);
            const unsigned e = i / 2;
            idx.y = T(e / ((nx + 1) * nz));
            const unsigned t = e % ((nx + 1) * nz);
            idx.x = T(t / nz);
            idx.z = T(t % nz);
            r = domain.min + cmul(cellSize, idx);
            r.z += (T(1) + T(i % 2)) / T(3) * cellSize.z;
        }
        return r;
    }
    // SDF.cu:81-131 - the 32 node indices of cell (i,j,k)
    void cell_nodes(unsigned i, unsigned j, unsigned k, unsigned out[32]) const {
        const unsigned nx = res[0], ny = res[1], nz = res[2];
        const unsigned nv = (nx + 1) * (ny + 1) * (nz + 1);
        const unsigned nex = nx * (ny + 1) * (nz + 1), ney = (nx + 1) * ny * (nz + 1);
        for (unsigned c = 0; c < 8; ++c) {
            const unsigned di = c & 1u, dj = (c >> 1) & 1u, dk = (c >> 2) & 1u;
            out[c] = (nx + 1) * (ny + 1) * (k + dk) + (nx + 1) * (j + dj) + (i + di);
        }
        for (unsigned e = 0; e < 4; ++e) {  // x edges at (j+dj, k+dk), e = dk + 2 dj
            const unsigned dk = e & 1u, dj = e >> 1;
            const unsigned b = nv + 2 * (nx * (ny + 1) * (k + dk) + nx * (j + dj) + i);
            out[8 + 2 * e] = b;
            out[9 + 2 * e] = b + 1;
        }
        for (unsigned e = 0; e < 4; ++e) {  // y edges at (i+di, k+dk), e = di + 2 dk
            const unsigned di = e & 1u, dk = e >> 1;
            const unsigned b = nv + 2 * nex + 2 * (ny * (nz + 1) * (i + di) + ny * (k + dk) + j);
            out[16 + 2 * e] = b;
            out[17 + 2 * e] = b + 1;
        }
        for (unsigned e = 0; e < 4; ++e) {  // z edges at (i+di, j+dj), e = dj + 2 di
            const unsigned dj = e & 1u, di = e >> 1;
            const unsigned b = nv + 2 * nex + 2 * ney + 2 * (nz * (nx + 1) * (j + dj) + nz * (i + di) + k);
            out[24 + 2 * e] = b;
            out[25 + 2 * e] = b + 1;
        }
    }
    // Locate the cell, build xi and the shape functions. Returns false outside the domain.
    bool locate(V3<T> x, unsigned cell[32], V3<T>& c0, T N[32], V3<T> dN[32], V3<T>* xiOut = nullptr) const {
        if (!domain.contains(x)) return false;
        const V3<T> f = cmul(cellSizeInv, x - domain.min);
        unsigned mi[3] = {(unsigned)f.x, (unsigned)f.y, (unsigned)f.z};
        for (int a = 0; a < 3; ++a)
            if (mi[a] >= res[a]) mi[a] = res[a] - 1;
        const V3<T> origin = domain.min + cmul(v3<T>(T(mi[0]), T(mi[1]), T(mi[2])), cellSize);
        const V3<T> top = origin + cellSize;
        const V3<T> den = top - origin;
        c0 = v3<T>(T(2) / den.x, T(2) / den.y, T(2) / den.z);
        const V3<T> sum = top + origin;
        const V3<T> c1 = v3<T>(sum.x / den.x, sum.y / den.y, sum.z / den.z);
        const V3<T> xi = cmul(c0, x) - c1;
        cell_nodes(mi[0], mi[1], mi[2], cell);
        shape_function(xi, N, dN);
        if (xiOut) *xiOut = xi;
        return true;
    }
    // Value, gradient and Hessian (xx, yy, zz, xy, xz, yz) of a field's interpolant at x in world coordinates;
    // false outside the domain. The interpolant is a cubic polynomial per cell, so these are exact within a cell.
    bool derivatives(unsigned field, V3<T> x, T& phi, V3<T>& grad, T hess[6]) const {
        unsigned cell[32];
        V3<T> c0, xi;
        T N[32];
        V3<T> dN[32];
        if (!locate(x, cell, c0, N, dN, &xi)) return false;
        phi = interp_grad(field, cell, c0, N, dN, grad);
        T H[32][6];
        shape_function_hess(xi, H);
        for (int e = 0; e < 6; ++e) hess[e] = T(0);
        for (int k = 0; k < 32; ++k) {
            const T c = nodes[field][cell[k]];
            for (int e = 0; e < 6; ++e) hess[e] += c * H[k][e];
        }
        const T s[6] = {c0.x * c0.x, c0.y * c0.y, c0.z * c0.z, c0.x * c0.y, c0.x * c0.z, c0.y * c0.z};
        for (int e = 0; e < 6; ++e) hess[e] *= s[e];
        return true;
    }
    T interp(unsigned field, const unsigned cell[32], const T N[32]) const {
        T phi = 0;
        for (int k = 0; k < 32; ++k) phi += nodes[field][cell[k]] * N[k];
        return phi;
    }
    T interp_grad(unsigned field, const unsigned cell[32], V3<T> c0, const T N[32], const V3<T> dN[32],
                  V3<T>& grad) const {
        T phi = 0;
        grad = v3<T>(0, 0, 0);
        for (int k = 0; k < 32; ++k) {
            const T c = nodes[field][cell[k]];
            phi += c * N[k];
            grad = grad + c * dN[k];
        }
        grad = cmul(grad, c0);
        return phi;
    }
    // host Interpolate(field, x) without gradient; FLT_MAX outside the domain
    T interp_at(unsigned field, V3<T> x) const {
        unsigned cell[32];
        V3<T> c0;
        T N[32];
        V3<T> dN[32];
        if (!locate(x, cell, c0, N, dN)) return T(FLT_MAX);
        return interp(field, cell, N);
    }
};

// ---------------------------------------------------------------------------------------------------------------
// Triangle-mesh signed distance: MeshDistance (Utility/SDF/MeshDistance.cpp:13-608) with precomputed normals, over
// the half-edge mesh of EdgeMesh(vertices, faces) (Renderer/Mesh/EdgeMesh.cpp:132-166), vertices transformed as
// RigidBody.cu:18-22 / FluidObject.cpp:9-14 do. The reference finds the closest triangle through a bounding-sphere
// hierarchy; the result is the least squared distance over all triangles, which this restatement takes by a scan in
// triangle order (first minimum). Pinned by tests/test_oracle_mesh.py (cube and sphere meshes against the analytic
// box / polyhedron distance, brute-force closest points, sign = inside/outside by ray parity).
// ---------------------------------------------------------------------------------------------------------------
template <class T> inline T gdot(V3<T> a, V3<T> b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }  // glm::dot
template <class T> inline V3<T> gnormalize(V3<T> a) { return a * (T(1) / std::sqrt(gdot(a, a))); }

template <class T> struct TriMesh {
    std::vector<V3<T>> x;                 // transformed vertices
    std::vector<unsigned> f;              // 3 per triangle
    std::vector<V3<T>> faceN, vertexN;    // MeshDistance.cpp:28-56
    std::vector<long> opp;                // opposite half-edge (3 face + edge) or -1 at a border
    V3<T> lo{0, 0, 0}, hi{0, 0, 0};       // BoundingBox(vertices) from the zero box (BoundingBox.h:16-31)

    // glm::mat4 * vec4(v, 1) = (c0 x + c1 y) + (c2 z + c3 w), columns c of the column-major matrix
    static V3<T> transform(const float* m, V3<T> v) {
        T o[3];
        for (int r = 0; r < 3; ++r)
            o[r] = ((T)m[r] * v.x + (T)m[4 + r] * v.y) + ((T)m[8 + r] * v.z + (T)m[12 + r] * T(1));
        return v3<T>(o[0], o[1], o[2]);
    }
    void build(const float* verts, unsigned nv, const unsigned* tris, unsigned nt, const float* m) {
        x.resize(nv);
        for (unsigned i = 0; i < nv; ++i) {
            const V3<T> v = v3<T>((T)verts[3 * i], (T)verts[3 * i + 1], (T)verts[3 * i + 2]);
            x[i] = m ? transform(m, v) : v;
        }
        f.assign(tris, tris + 3 * (size_t)nt);
        lo = hi = v3<T>(0, 0, 0);
        for (const V3<T>& v : x) {  // glm::min(a, b) = b < a ? b : a; glm::max(a, b) = a < b ? b : a
            lo = v3<T>(v.x < lo.x ? v.x : lo.x, v.y < lo.y ? v.y : lo.y, v.z < lo.z ? v.z : lo.z);
            hi = v3<T>(hi.x < v.x ? v.x : hi.x, hi.y < v.y ? v.y : hi.y, hi.z < v.z ? v.z : hi.z);
        }
        faceN.resize(nt);
        vertexN.assign(nv, v3<T>(0, 0, 0));
        for (unsigned t = 0; t < nt; ++t) {
            const V3<T>& x0 = x[f[3 * t]];
            const V3<T>& x1 = x[f[3 * t + 1]];
            const V3<T>& x2 = x[f[3 * t + 2]];
            const V3<T> n = gnormalize(cross(x1 - x0, x2 - x0));
            const V3<T> e1 = gnormalize(x1 - x0), e2 = gnormalize(x2 - x1), e3 = gnormalize(x0 - x2);
            const T ax = std::acos(gdot(e1, -e3)), ay = std::acos(gdot(e2, -e1)), az = std::acos(gdot(e3, -e2));
            vertexN[f[3 * t]] = vertexN[f[3 * t]] + ax * n;
            vertexN[f[3 * t + 1]] = vertexN[f[3 * t + 1]] + ay * n;
            vertexN[f[3 * t + 2]] = vertexN[f[3 * t + 2]] + az * n;
            faceN[t] = n;
        }
        // pairing of opposite half-edges: a half-edge (a -> b) meets a waiting (b -> a); earliest waiting first
        opp.assign(3 * (size_t)nt, -1);
        std::map<std::pair<unsigned, unsigned>, std::vector<long>> waiting;
        for (unsigned t = 0; t < nt; ++t)
            for (unsigned e = 0; e < 3; ++e) {
                const unsigned a = f[3 * t + e], b = f[3 * t + (e + 1) % 3];
                const long he = 3 * (long)t + e;
                auto it = waiting.find({b, a});
                if (it != waiting.end() && !it->second.empty()) {
                    opp[he] = it->second.front();
                    opp[it->second.front()] = he;
                    it->second.erase(it->second.begin());
                } else {
                    waiting[{a, b}].push_back(he);
                }
            }
    }
    // PointTriangleDistanceSquared (MeshDistance.cpp:303-607): Eberly's region classification of the minimiser
    // (s, t) of |x0 + s e0 + t e1 - p|^2, entity 0..2 = vertex A, B, C; 3..5 = edge D (x0x1), E (x1x2), F (x2x0);
    // 6 = face G.
    T point_triangle_d2(V3<T> point, unsigned t3, V3<T>* closest, int* entity) const {
        const V3<T>& x0 = x[f[t3]];
        const V3<T> diff = x0 - point;
        const V3<T> edge0 = x[f[t3 + 1]] - x0;
        const V3<T> edge1 = x[f[t3 + 2]] - x0;
        const T a00 = gdot(edge0, edge0), a01 = gdot(edge0, edge1), a11 = gdot(edge1, edge1);
        const T b0 = gdot(diff, edge0), b1 = gdot(diff, edge1), c = gdot(diff, diff);
        const T det = std::fabs(a00 * a11 - a01 * a01);
        T s = a01 * b1 - a11 * b0;
        T t = a01 * b0 - a00 * b1;
        T d2 = T(-1);
        int ent = -1;
        const T two = T(2);
        auto quad = [&] { return s * (a00 * s + a01 * t + two * b0) + t * (a01 * s + a11 * t + two * b1) + c; };
        auto vA = [&] { ent = 0; s = 0; t = 0; d2 = c; };
        auto vB = [&] { ent = 1; s = 1; t = 0; d2 = a00 + two * b0 + c; };
        auto vC = [&] { ent = 2; s = 0; t = 1; d2 = a11 + two * b1 + c; };
        auto eD = [&] { ent = 3; t = 0; s = -b0 / a00; d2 = b0 * s + c; };
        auto eF = [&] { ent = 5; s = 0; t = -b1 / a11; d2 = b1 * t + c; };
        if (s + t <= det) {
            if (s < 0) {
                if (t < 0 && b0 < 0) {
                    if (-b0 >= a00) vB(); else eD();
                } else {  // region 4 with b0 >= 0, or region 3
                    if (b1 >= 0) vA(); else if (-b1 >= a11) vC(); else eF();
                }
            } else if (t < 0) {  // region 5
                if (b0 >= 0) vA(); else if (-b0 >= a00) vB(); else eD();
            } else {  // region 0
                ent = 6;
                const T invDet = T(1) / det;
                s *= invDet;
                t *= invDet;
                d2 = quad();
            }
        } else {
            const T denom = a00 - two * a01 + a11;
            if (s < 0) {  // region 2
                const T tmp0 = a01 + b0, tmp1 = a11 + b1;
                if (tmp1 > tmp0) {
                    const T numer = tmp1 - tmp0;
                    if (numer >= denom) vB();
                    else { ent = 4; s = numer / denom; t = 1 - s; d2 = quad(); }
                } else {
                    if (tmp1 <= 0) vC(); else if (b1 >= 0) vA(); else eF();
                }
            } else if (t < 0) {  // region 6
                const T tmp0 = a01 + b1, tmp1 = a00 + b0;
                if (tmp1 > tmp0) {
                    const T numer = tmp1 - tmp0;
                    if (numer >= denom) vC();
                    else { ent = 4; t = numer / denom; s = 1 - t; d2 = quad(); }
                } else {
                    if (tmp1 <= 0) vB(); else if (b0 >= 0) vA(); else eD();
                }
            } else {  // region 1
                const T numer = a11 + b1 - a01 - b0;
                if (numer <= 0) vC();
                else if (numer >= denom) vB();
                else { ent = 4; s = numer / denom; t = 1 - s; d2 = quad(); }
            }
        }
        if (d2 < 0) d2 = 0;
        if (closest) *closest = x0 + s * edge0 + t * edge1;
        if (entity) *entity = ent;
        return d2;
    }
    // SignedDistance (MeshDistance.cpp:186-225): distance to the closest triangle, negative when (p - closest point)
    // points against the closest entity's normal (vertex: angle-weighted; edge: sum of the two face normals, or the
    // face's alone at a border; face: its normal).
    T signed_distance(V3<T> p) const {
        const unsigned nt = (unsigned)(f.size() / 3);
        T best = std::numeric_limits<T>::infinity();
        unsigned face = 0;
        for (unsigned t = 0; t < nt; ++t) {
            const T d2 = point_triangle_d2(p, 3 * t, nullptr, nullptr);
            if (d2 < best) {
                best = d2;
                face = t;
            }
        }
        V3<T> cp;
        int ent;
        T dist = std::sqrt(point_triangle_d2(p, 3 * face, &cp, &ent));
        V3<T> n;
        if (ent <= 2)
            n = vertexN[f[3 * face + ent]];
        else if (ent <= 5) {
            const long o = opp[3 * (size_t)face + (ent - 3)];
            n = o < 0 ? faceN[face] : faceN[face] + faceN[(size_t)o / 3];
        } else
            n = faceN[face];
        if (gdot(p - cp, n) < T(0)) dist *= T(-1);
        return dist;
    }
};

// Builds both fields of a body's volume map (RigidBody.cu:10-73).
// A mesh body (mesh != nullptr) takes its domain from the transformed vertices' bounding box and its distance from
// TriMesh::signed_distance; b then supplies Inverted, Padding and CollisionMapResolution only.
template <class T>
void build_volume_map(VolumeMap<T>& map, const vfd_boundary_desc& b, const CubicKernel<T>& kernel, T particleRadius,
                      const TriMesh<T>* mesh = nullptr) {
    const T R = kernel.R;
    const T tol = (T)b.Padding - particleRadius;
    const T sign = b.Inverted ? T(-1) : T(1);
    Box<T> dom;
    if (mesh) {
        const T grow = T(8) * R + tol;  // RigidBody.cu:34-36
        dom.min = mesh->lo - v3<T>(grow, grow, grow);
        dom.max = mesh->hi + v3<T>(grow, grow, grow);
    } else if (b.HasDomain) {
        dom.min = v3<T>((T)b.DomainMin[0], (T)b.DomainMin[1], (T)b.DomainMin[2]);
        dom.max = v3<T>((T)b.DomainMax[0], (T)b.DomainMax[1], (T)b.DomainMax[2]);
    } else {
        V3<T> lo, hi;
        const V3<T> c = v3<T>((T)b.Center[0], (T)b.Center[1], (T)b.Center[2]);
        V3<T> e;
        if (b.Shape == VFD_SHAPE_SPHERE)
            e = v3<T>((T)b.Radius, (T)b.Radius, (T)b.Radius);
        else if (b.Shape == VFD_SHAPE_CYLINDER)
            e = v3<T>((T)b.Radius, (T)b.HalfExtents[1], (T)b.Radius);
        else
            e = v3<T>((T)b.HalfExtents[0], (T)b.HalfExtents[1], (T)b.HalfExtents[2]);
        lo = c - e;
        hi = c + e;
        const T grow = T(8) * R + tol;
        dom.min = lo - v3<T>(grow, grow, grow);
        dom.max = hi + v3<T>(grow, grow, grow);
    }
    map.init(dom, b.CollisionMapResolution);
    map.nodes[0].assign(map.nodeCount, T(0));
    map.nodes[1].assign(map.nodeCount, T(0));
#pragma omp parallel for schedule(static)
    for (long l = 0; l < (long)map.nodeCount; ++l) {
        const V3<T> x = map.node_position((unsigned)l);
        map.nodes[0][l] = sign * ((mesh ? mesh->signed_distance(x) : shape_distance<T>(b, x)) - tol);
    }
    // Gauss-Legendre 16 points on [-R, R]^3; nodes/weights stored in T like the reference's float table.
    std::vector<double> gx, gw;
    gauss_legendre(16, gx, gw);
    T xs[16], ws[16];
    for (int i = 0; i < 16; ++i) {
        xs[i] = (T)gx[i];
        ws[i] = (T)gw[i];
    }
    const double c0 = (double)R;  // 0.5 * Diagonal() of [-R, R]^3; c1 = 0
    const T W0 = kernel.W0;
#pragma omp parallel for schedule(dynamic, 64)
    for (long l = 0; l < (long)map.nodeCount; ++l) {
        const V3<T> x = map.node_position((unsigned)l);
        const T dX = map.interp_at(0, x);
        if (dX > T(2) * R) {
            map.nodes[1][l] = T(0);
            continue;
        }
        T res = 0;
        for (int i = 0; i < 16; ++i) {
            const T wi = ws[i];
            for (int j = 0; j < 16; ++j) {
                const T wij = wi * ws[j];
                for (int k = 0; k < 16; ++k) {
                    const T wijk = wij * ws[k];
                    const V3<T> xi = v3<T>((T)(c0 * (double)xs[i]), (T)(c0 * (double)xs[j]), (T)(c0 * (double)xs[k]));
                    T f = 0;
                    if (!(dot3(xi, xi) > R * R)) {
                        const T dist = map.interp_at(0, x + xi);
                        if (dist <= T(0))
                            f = T(1);
                        else if (dist < R)
                            f = kernel.getW(dist) / W0;
                    }
                    res += wijk * f;
                }
            }
        }
        res = (T)((double)res * (c0 * c0 * c0));
        map.nodes[1][l] = T(0.8) * res;
    }
}

// ---------------------------------------------------------------------------------------------------------------
// Neighbour search: ParticleSearch.cu:29-195, ParticleSearchKernels.cu:4-193.
// ---------------------------------------------------------------------------------------------------------------
inline unsigned part1by2(unsigned x) {  // spread the low 10 bits of x to every third bit
    unsigned r = 0;
    for (unsigned b = 0; b < 10; ++b) r |= ((x >> b) & 1u) << (3 * b);
    return r;
}
inline unsigned morton3(unsigned x, unsigned y, unsigned z) { return (part1by2(z) << 2) + (part1by2(y) << 1) + part1by2(x); }

template <class T> struct SearchGrid {
    V3<T> gridMin, gridDelta;
    unsigned dims[3], meta[3];
    unsigned cellCount;
    T R2;
    unsigned cell_index(int cx, int cy, int cz) const {  // ToCellIndexMortonMetaGrid :24-38
        const unsigned mx = (unsigned)cx / kMetaGroup, my = (unsigned)cy / kMetaGroup, mz = (unsigned)cz / kMetaGroup;
        const unsigned lin = mz * meta[1] * meta[0] + my * meta[0] + mx;
        return lin * kMetaBlock + morton3((unsigned)cx % kMetaGroup, (unsigned)cy % kMetaGroup, (unsigned)cz % kMetaGroup);
    }
};

template <class T> SearchGrid<T> make_grid(const std::vector<V3<T>>& x, T R) {
    int mn[3] = {INT_MAX, INT_MAX, INT_MAX}, mx[3] = {INT_MIN, INT_MIN, INT_MIN};
    for (const V3<T>& p : x) {  // ComputeMinMaxKernel :40-62
        const int c[3] = {(int)std::floor(p.x / R), (int)std::floor(p.y / R), (int)std::floor(p.z / R)};
        for (int a = 0; a < 3; ++a) {
            mn[a] = std::min(mn[a], c[a]);
            mx[a] = std::max(mx[a], c[a]);
        }
    }
    for (int a = 0; a < 3; ++a) mx[a] += 1;
    const V3<T> bmin = v3<T>(T(mn[0]) * R, T(mn[1]) * R, T(mn[2]) * R);
    const V3<T> bmax = v3<T>(T(mx[0]) * R, T(mx[1]) * R, T(mx[2]) * R);
    SearchGrid<T> g;  // ComputeCellInformation ParticleSearch.cu:57-93
    g.R2 = R * R;
    V3<T> size = bmax - bmin;
    g.dims[0] = (unsigned)std::ceil(size.x / R) + 4;
    g.dims[1] = (unsigned)std::ceil(size.y / R) + 4;
    g.dims[2] = (unsigned)std::ceil(size.z / R) + 4;
    g.gridMin = bmin - v3<T>(R, R, R) * T(2);
    for (int a = 0; a < 3; ++a) g.meta[a] = (unsigned)std::ceil(T(g.dims[a]) / T(kMetaGroup));
    size = v3<T>(T(g.dims[0]) * R, T(g.dims[1]) * R, T(g.dims[2]) * R);
    g.gridDelta = v3<T>(T(g.dims[0]) / size.x, T(g.dims[1]) / size.y, T(g.dims[2]) / size.z);
    g.cellCount = g.meta[0] * g.meta[1] * g.meta[2] * kMetaBlock;
    return g;
}

struct NeighborLists {
    std::vector<uint32_t> counts, offsets, ids;
};

template <class T> void find_neighbors(const std::vector<V3<T>>& x, T R, NeighborLists& out) {
    const unsigned n = (unsigned)x.size();
    out.counts.assign(n, 0);
    out.offsets.assign(n, 0);
    out.ids.clear();
    if (n == 0) return;
    const SearchGrid<T> g = make_grid(x, R);
    std::vector<uint32_t> cellOf(n), cellCount(g.cellCount, 0), cellStart(g.cellCount, 0), sorted(n);
    for (unsigned i = 0; i < n; ++i) {  // InsertParticlesMortonKernel :64-78 (truncation of a positive value)
        const V3<T> f = cmul(x[i] - g.gridMin, g.gridDelta);
        cellOf[i] = g.cell_index((int)f.x, (int)f.y, (int)f.z);
        cellCount[cellOf[i]]++;
    }
    uint32_t run = 0;
    for (unsigned c = 0; c < g.cellCount; ++c) {
        cellStart[c] = run;
        run += cellCount[c];
    }
    {
        std::vector<uint32_t> fill(cellStart);
        for (unsigned i = 0; i < n; ++i) sorted[fill[cellOf[i]]++] = i;  // ascending index within each cell
    }
    std::vector<uint32_t> tmp((size_t)n * kMaxNeighbors);
#pragma omp parallel for schedule(static)
    for (long il = 0; il < (long)n; ++il) {  // ComputeCountsKernel :93-142 / NeighborhoodQueryWithCountsKernel :144-193
        const unsigned i = (unsigned)il;
        const V3<T> p = x[i];
        const V3<T> f = cmul(p - g.gridMin, g.gridDelta);
        const int gc[3] = {(int)std::floor(f.x), (int)std::floor(f.y), (int)std::floor(f.z)};
        unsigned cnt = 0;
        uint32_t* dst = &tmp[(size_t)i * kMaxNeighbors];
        for (int dz = -1; dz < 2 && cnt < kMaxNeighbors; ++dz)
            for (int dy = -1; dy < 2 && cnt < kMaxNeighbors; ++dy)
                for (int dx = -1; dx < 2 && cnt < kMaxNeighbors; ++dx) {
                    const int cx = gc[0] + dx, cy = gc[1] + dy, cz = gc[2] + dz;
                    if (cx < 0 || cy < 0 || cz < 0 || cx >= (int)g.dims[0] || cy >= (int)g.dims[1] ||
                        cz >= (int)g.dims[2])
                        continue;
                    const unsigned c = g.cell_index(cx, cy, cz);
                    for (uint32_t s = cellStart[c]; s < cellStart[c] + cellCount[c]; ++s) {
                        const unsigned j = sorted[s];
                        const V3<T> d = x[j] - p;
                        const T d2 = std::fma(d.z, d.z, std::fma(d.y, d.y, d.x * d.x));
                        if (d2 < g.R2 && d2 > T(0)) dst[cnt++] = j;
                        if (cnt == kMaxNeighbors) break;
                    }
                }
        out.counts[i] = cnt;
    }
    uint64_t total = 0;
    for (unsigned i = 0; i < n; ++i) {
        out.offsets[i] = (uint32_t)total;
        total += out.counts[i];
    }
    out.ids.resize(total);
    for (unsigned i = 0; i < n; ++i)
        std::memcpy(&out.ids[out.offsets[i]], &tmp[(size_t)i * kMaxNeighbors], out.counts[i] * sizeof(uint32_t));
}

// ---------------------------------------------------------------------------------------------------------------
// The simulation. DFSPHImplementation.cu, DFSPHKernels.cu.
// ---------------------------------------------------------------------------------------------------------------
template <class T> struct Particle {  // Structures/DFSPHParticle.h:8-32
    V3<T> x, v, a, ap;
    T residuum, density, densityAdv, kappa, kappaV, factor;
    V3<T> dv;
    V3<T> nMC, nS;
    T kMC, kS, dK;
};

template <class T> struct Info {  // Structures/DFSPHSimulationInfo.h:5-44
    unsigned particleCount = 0, bodyCount = 0;
    T R, R2, r, diam, dt, dt2, dtInv, dt2Inv, V, rho0, m, mInv, nu, nuB, mu, muB, tdf, td;
    T sigma;
    unsigned S = 0;
    T slope, constant;
    bool temporal;
    T smoothing, factor, NR, mcFactor;
    V3<T> g;
};

// ParticleSampler::SampleMeshVolume (Utility/Sampler/ParticleSampler.cpp:7-91): a one-field map of the mesh's signed
// distance over its bounding box grown by 0.001 of the diagonal (SDF.cu:16-35; the upper corner first, then the lower
// one with the grown box's diagonal), and the lattice of the sample mode kept where the interpolated distance is
// negative, in the reference's z, y, x loop order with its float loop variables. Pinned by tests/test_oracle_mesh.py
// (sample counts of a box mesh against the closed-form lattice count).
template <class T>
void sample_mesh_volume(const TriMesh<T>& mesh, T radius, const unsigned res[3], bool inverted, int mode,
                        std::vector<T>& out) {
    const Box<T> bounds{mesh.lo, mesh.hi};
    Box<T> dom = bounds;
    const T g = T(0.001f);
    V3<T> diag = dom.max - dom.min;
    const T up = g * std::sqrt(gdot(diag, diag));
    dom.max = dom.max + v3<T>(up, up, up);
    diag = dom.max - dom.min;
    const T down = g * std::sqrt(gdot(diag, diag));
    dom.min = dom.min - v3<T>(down, down, down);
    VolumeMap<T> sdf;
    sdf.init(dom, res);
    sdf.nodes[0].assign(sdf.nodeCount, T(0));
    const T factor = inverted ? T(-1) : T(1);
#pragma omp parallel for schedule(dynamic, 256)
    for (long l = 0; l < (long)sdf.nodeCount; ++l) sdf.nodes[0][l] = factor * mesh.signed_distance(sdf.node_position((unsigned)l));
    const T diameter = T(2) * radius;
    T shiftX = diameter, shiftY = diameter;
    if (mode == VFD_SAMPLE_MEDIUM_DENSITY) {
        shiftY = std::sqrt(T(3)) * radius;
    } else if (mode == VFD_SAMPLE_MAX_DENSITY) {
        shiftX = std::sqrt(T(3)) * radius;
        shiftY = std::sqrt(T(6)) * diameter / T(3);
    }
    out.clear();
    unsigned counterX = 0, counterY = 0;
    for (T z = bounds.min.z; z <= bounds.max.z; z += diameter) {
        for (T y = bounds.min.y; y <= bounds.max.y; y += shiftY) {
            for (T x = bounds.min.x; x <= bounds.max.x; x += shiftX) {
                V3<T> p;
                if (mode == VFD_SAMPLE_MIN_DENSITY) {
                    p = v3<T>(x + radius, y + radius, z + radius);
                } else if (mode == VFD_SAMPLE_MEDIUM_DENSITY) {
                    p = counterY % 2 == 0 ? v3<T>(x, y + radius, z + radius) : v3<T>(x + radius, y + radius, z);
                } else {
                    p = v3<T>(x, y + radius, z + radius);
                    V3<T> shift = v3<T>(0, 0, 0);
                    if (counterX % 2) shift.z += diameter / (T(2) * T(counterY % 2 ? -1 : 1));
                    if (counterY % 2) {
                        shift.x += shiftX / T(2);
                        shift.z += diameter / T(2);
                    }
                    p = p + shift;
                }
                if (sdf.interp_at(0, p) - T(0) < T(0)) {  // SDF::GetDistance(p, 0) (SDF.cu:217-225)
                    out.push_back(p.x);
                    out.push_back(p.y);
                    out.push_back(p.z);
                }
                ++counterX;
            }
            counterX = 0;
            ++counterY;
        }
        counterY = 0;
    }
}

template <class T> struct Body {
    vfd_boundary_desc desc;
    VolumeMap<T> map;
    std::vector<V3<T>> xb;
    std::vector<T> vb;
};

template <class T> class Sim {
public:
    using Real = T;
    vfd_desc desc{};
    Info<T> info{};
    CubicKernel<T> kernel;
    std::vector<T> halton;
    std::vector<Particle<T>> p;
    std::vector<Body<T>> bodies;
    NeighborLists nb;
    std::vector<M3<T>> minv;
    std::vector<V3<T>> vb_b, vb_g, vb_r, vb_p, vb_z, vb_tmp, vb_ap;
    vfd_debug_info dbg{};
    T maxVel = 0;
    bool kernelSet = false;

    // Per-step record of everything the adjoint needs (frozen lists, clamp masks, solver iterates, time steps).
    struct StepTape {
        NeighborLists nb;
        std::vector<V3<T>> x, v0, v1, y, v2, v3;  // positions; velocity at entry, after divergence, viscosity solution,
                                                  // before and after the pressure solve
        std::vector<T> alpha, D, rhoadv;
        std::vector<uint8_t> m0, mp0;
        std::vector<std::vector<T>> kv, kp;            // kappa iterates (K + 1 each)
        std::vector<std::vector<uint8_t>> mv, mp;      // residual masks (K each)
        T h = 0, H = 0;
        bool visc = false;
        // Boundary contact sample of a particle against one body: r = x - xb = s n with s = max(phi + r/2, 2r) and
        // n = grad phi / |grad phi| at the particle, the sample volume vb (> 0; 0 = no sample) and the derivatives
        // of both in the particle's position (Jr = dr/dx, row-major).
        struct Contact {
            T vb = T(0);
            V3<T> dvb{T(0), T(0), T(0)}, r{T(0), T(0), T(0)};
            M3<T> Jr = mzero<T>();
            M3<T> Jn = mzero<T>();  // d n / dx (the boundary viscosity samples turn with the normal)
        };
        std::vector<std::vector<Contact>> contact;  // [body][particle]
        // Push-out of a particle found inside a body: x <- x + delta n, v <- 0; Jpush = dx_after / dx_before
        std::vector<uint8_t> pushed;
        std::vector<M3<T>> Jpush;
        // surface tension: the Monte-Carlo normals and curvatures of the step (their sample coverage and the
        // classification are discrete), the blend pass count and, with temporal smoothing, the blended curvature dK
        // each particle carried into the step
        bool st = false, temporal = false;
        unsigned stPasses = 0;
        std::vector<V3<T>> nMC;
        std::vector<T> kMC, dK0;
    };
    std::vector<StepTape> tapes;
    std::vector<V3<T>> tapePre;  // positions the boundary pass of a taped step started from
    bool taping = false;

    explicit Sim(const vfd_desc& d) {
        halton = halton_table<T>();
        set_description(d);
    }

    // SetDescription DFSPHImplementation.cu:318-368 (+ the sample count / Monte-Carlo factor initialised from dt0,
    // which the reference leaves uninitialised until the first ComputeTimeStepSize)
    void set_description(const vfd_desc& d) {
        desc = d;
        const T r = (T)d.ParticleRadius;
        if (!kernelSet || r != info.r) {
            kernel.setRadius(T(4) * r);
            kernelSet = true;
        }
        info.r = r;
        info.diam = T(2) * r;
        info.R = T(4) * r;
        info.R2 = info.R * info.R;
        info.V = T(0.8) * info.diam * info.diam * info.diam;
        info.rho0 = T(1000);
        info.m = info.V * info.rho0;
        info.mInv = T(1) / info.m;
        info.nu = (T)d.Viscosity;
        info.nuB = (T)d.BoundaryViscosity;
        info.mu = info.nu * info.rho0;
        info.muB = info.nuB * info.rho0;
        info.tdf = (T)d.TangentialDistanceFactor;
        info.td = info.tdf * info.R;
        set_dt((T)d.TimeStepSize);
        info.sigma = (T)d.SurfaceTension;
        info.slope = T(74.688796680497925);
        info.constant = T(12);
        info.temporal = d.TemporalSmoothing != 0;
        info.smoothing = T(0.5);
        info.factor = T(0.8);
        info.NR = info.r * info.factor;
        info.g = v3<T>((T)d.Gravity[0], (T)d.Gravity[1], (T)d.Gravity[2]);
        update_sample_count();
        info.mcFactor = std::asin(info.NR / info.r);
    }
    void set_dt(T dt) {
        info.dt = dt;
        info.dt2 = dt * dt;
        info.dtInv = T(1) / info.dt;
        info.dt2Inv = T(1) / info.dt2;
    }
    void update_sample_count() {
        if (desc.CSDFix > 0)
            info.S = (unsigned)desc.CSDFix;
        else
            info.S = (unsigned)(int)((T)desc.CSD * info.dt);
    }

    // SetFluidObjects DFSPHImplementation.cu:172-253
    void set_fluid(const float* x, const float* v, unsigned n) {
        p.assign(n, Particle<T>{});
        for (unsigned i = 0; i < n; ++i) {
            Particle<T>& q = p[i];
            std::memset(&q, 0, sizeof(q));
            q.x = v3<T>((T)x[3 * i], (T)x[3 * i + 1], (T)x[3 * i + 2]);
            q.v = v ? v3<T>((T)v[3 * i], (T)v[3 * i + 1], (T)v[3 * i + 2]) : v3<T>(0, 0, 0);
        }
        info.particleCount = n;
        for (auto& b : bodies) {
            b.xb.assign(n, v3<T>(0, 0, 0));
            b.vb.assign(n, T(0));
        }
        const size_t m = n;
        minv.assign(m, mzero<T>());
        vb_b.assign(m, v3<T>(0, 0, 0));
        vb_g = vb_r = vb_p = vb_z = vb_tmp = vb_ap = vb_b;
    }
    void add_boundary(const vfd_boundary_desc& d, const TriMesh<T>* mesh = nullptr) {
        bodies.emplace_back();
        Body<T>& b = bodies.back();
        b.desc = d;
        build_volume_map(b.map, d, kernel, info.r, mesh);
        b.xb.assign(info.particleCount, v3<T>(0, 0, 0));
        b.vb.assign(info.particleCount, T(0));
        info.bodyCount = (unsigned)bodies.size();
    }

    // ----- kernels --------------------------------------------------------------------------------------------
    // ComputeVolumeAndBoundaryKernel DFSPHKernels.cu:69-150
    void boundary() {
        const long n = (long)p.size();
#pragma omp parallel for schedule(static)
        for (long i = 0; i < n; ++i) {
            for (auto& b : bodies) {
                V3<T>& xb = b.xb[i];
                T& vb = b.vb[i];
                xb = v3<T>(0, 0, 0);
                vb = T(0);
                V3<T> normal = v3<T>(0, 0, 0);
                T dist = T(FLT_MAX);
                unsigned cell[32];
                V3<T> c0;
                T N[32];
                V3<T> dN[32];
                if (b.map.locate(p[i].x, cell, c0, N, dN)) dist = b.map.interp_grad(0, cell, c0, N, dN, normal);
                if (dist > T(0) && dist < info.R) {
                    const T vol = b.map.interp(1, cell, N);
                    if (vol > T(0) && vol != T(FLT_MAX)) {
                        vb = vol;
                        const T nl = length(normal);
                        if (nl > T(1.0e-9)) {
                            normal = normal / nl;
                            const T d = std::max(dist + T(0.5) * info.r, info.diam);
                            xb = p[i].x - d * normal;
                        } else {
                            vb = T(0);
                        }
                    } else {
                        vb = T(0);
                    }
                } else if (dist <= T(0)) {
                    if (dist != T(FLT_MAX)) {
                        const T nl = length(normal);
                        if (nl > T(1.0e-5)) {
                            normal = normal / nl;
                            T delta = info.diam - dist;
                            delta = std::min(delta, T(0.1) * info.r);
                            p[i].x = p[i].x + delta * normal;
                            p[i].v = v3<T>(0, 0, 0);
                        }
                    }
                    vb = T(0);
                } else {
                    vb = T(0);
                }
            }
        }
    }
    unsigned cnt(long i) const { return nb.counts[i]; }
    unsigned nbr(long i, unsigned j) const { return nb.ids[nb.offsets[i] + j]; }

    // ComputeDensityKernel DFSPHKernels.cu:152-187
    void density() {
        const long n = (long)p.size();
#pragma omp parallel for schedule(static)
        for (long i = 0; i < n; ++i) {
            const V3<T> xi = p[i].x;
            T d = info.V * kernel.W0;
            for (unsigned j = 0; j < cnt(i); ++j) d += info.V * kernel.getW(xi - p[nbr(i, j)].x);
            for (auto& b : bodies) {
                const T vb = b.vb[i];
                if (vb > T(0)) d += vb * kernel.getW(xi - b.xb[i]);
            }
            p[i].density = d * info.rho0;
        }
    }
    // ComputeDFSPHFactorKernel :189-231
    void factor() {
        const long n = (long)p.size();
#pragma omp parallel for schedule(static)
        for (long i = 0; i < n; ++i) {
            const V3<T> xi = p[i].x;
            V3<T> gi = v3<T>(0, 0, 0);
            T sum = 0;
            for (unsigned j = 0; j < cnt(i); ++j) {
                const V3<T> gj = (-info.V) * kernel.gradW(xi - p[nbr(i, j)].x);
                sum += dot3(gj, gj);
                gi = gi - gj;
            }
            for (auto& b : bodies) {
                const T vb = b.vb[i];
                if (vb > T(0)) {
                    const V3<T> gj = (-vb) * kernel.gradW(xi - b.xb[i]);
                    gi = gi - gj;
                }
            }
            sum += dot3(gi, gi);
            p[i].factor = (double)sum > kEps ? T(1) / sum : T(0);
        }
    }
    // v_i . sum_j gradW (velocity divergence term shared by :233-278 and :483-529)
    T velocity_divergence(long i) const {
        const V3<T> xi = p[i].x, vi = p[i].v;
        T d = 0;
        for (unsigned j = 0; j < cnt(i); ++j) {
            const unsigned k = nbr(i, j);
            d += dot3(vi - p[k].v, kernel.gradW(xi - p[k].x));
        }
        d *= info.V;
        for (auto& b : bodies) {
            const T vb = b.vb[i];
            if (vb > T(0)) d += vb * dot3(vi, kernel.gradW(xi - b.xb[i]));
        }
        return d;
    }
    // pressure acceleration from kappa (:342-386, :388-432, :434-481, :555-602); divergence selects kappaV
    V3<T> pressure_accel(long i, bool divergence) const {
        const V3<T> xi = p[i].x;
        const T ki = divergence ? p[i].kappaV : p[i].kappa;
        V3<T> a = v3<T>(0, 0, 0);
        for (unsigned j = 0; j < cnt(i); ++j) {
            const unsigned k = nbr(i, j);
            const T s = ki + (divergence ? p[k].kappaV : p[k].kappa);
            if ((double)std::fabs(s) > kEps) {
                const V3<T> gj = (-info.V) * kernel.gradW(xi - p[k].x);
                a = a + s * gj;
            }
        }
        if ((double)std::fabs(ki) > kEps) {
            for (auto& b : bodies) {
                const T vb = b.vb[i];
                if (vb > T(0)) {
                    const V3<T> gj = (-vb) * kernel.gradW(xi - b.xb[i]);
                    a = a + ki * gj;
                }
            }
        }
        return a;
    }
    // ComputeDensityPressureForce :280-315
    T density_pressure_force(long i) const {
        const V3<T> xi = p[i].x, ai = p[i].ap;
        T f = 0;
        for (unsigned j = 0; j < cnt(i); ++j) {
            const unsigned k = nbr(i, j);
            f += dot3(ai - p[k].ap, kernel.gradW(xi - p[k].x));
        }
        f *= info.V;
        for (auto& b : bodies) {
            const T vb = b.vb[i];
            if (vb > T(0)) f += vb * dot3(ai, kernel.gradW(xi - b.xb[i]));
        }
        return f;
    }
    // sequential residual error -rho0 * sum(r) / N in the build's FP64 accumulation (DFSPHImplementation.cu:483-491)
    T residual_error() const {
        double s = 0;
        for (const auto& q : p) s += (double)(info.rho0 * q.residuum);
        return (T)(-s) / (T)p.size();
    }

    // ComputeDivergence DFSPHImplementation.cu:506-575
    void divergence_solve() {
        if (!desc.EnableDivergenceSolverError) {
            dbg.DivergenceSolverError = 0;
            dbg.DivergenceSolverIterationCount = 0;
            return;
        }
        const long n = (long)p.size();
#pragma omp parallel for schedule(static)
        for (long i = 0; i < n; ++i) {  // ComputeDensityChangeKernel :483-529
            T d = velocity_divergence(i);
            d = cnt(i) < 20u ? T(0) : std::max(d, T(0));
            p[i].densityAdv = d;
            p[i].factor *= info.dtInv;
            p[i].kappaV = d * p[i].factor;
        }
        if (taping) {
            StepTape& t = tapes.back();
            t.h = info.dt;
            t.D.resize(n);
            t.m0.resize(n);
            t.kv.assign(1, std::vector<T>(n));
            for (long i = 0; i < n; ++i) {
                t.D[i] = p[i].densityAdv;
                t.m0[i] = cnt(i) >= 20u && velocity_divergence(i) > T(0);
                t.kv[0][i] = p[i].kappaV;
            }
        }
        const T eta = info.dtInv * (T)desc.MaxDivergenceSolverError * T(0.0001) * info.rho0;
        unsigned it = 0;
        T err = 0;
        while ((err > eta || it < desc.MinDivergenceSolverIterations) && it < desc.MaxDivergenceSolverIterations) {
#pragma omp parallel for schedule(static)
            for (long i = 0; i < n; ++i) p[i].ap = pressure_accel(i, true);
#pragma omp parallel for schedule(static)
            for (long i = 0; i < n; ++i) {  // DivergenceSolveIterationKernel :531-553
                const T f = density_pressure_force(i) * info.dt;
                p[i].residuum = cnt(i) < 20u ? T(0) : std::min(-p[i].densityAdv - f, T(0));
                p[i].kappaV -= p[i].residuum * p[i].factor;
            }
            if (taping) {
                StepTape& t = tapes.back();
                t.mv.emplace_back(n);
                t.kv.emplace_back(n);
                for (long i = 0; i < n; ++i) {
                    t.mv.back()[i] = p[i].residuum < T(0);
                    t.kv.back()[i] = p[i].kappaV;
                }
            }
            err = residual_error();
            ++it;
        }
        dbg.DivergenceSolverIterationCount = it;
        dbg.DivergenceSolverError = (float)err;
#pragma omp parallel for schedule(static)
        for (long i = 0; i < n; ++i) {  // ComputePressureAccelerationAndFactorKernel :555-602
            p[i].ap = pressure_accel(i, true);
            p[i].v = p[i].v + info.dt * p[i].ap;
            p[i].factor *= info.dt;
        }
    }
    // ComputePressure DFSPHImplementation.cu:443-504
    void pressure_solve() {
        const long n = (long)p.size();
#pragma omp parallel for schedule(static)
        for (long i = 0; i < n; ++i) {  // ComputeDensityAdvectionKernel :233-278
            const T d = velocity_divergence(i);
            p[i].densityAdv = p[i].density / info.rho0 + info.dt * d;
            p[i].factor *= info.dt2Inv;
            const T si = T(1) - p[i].densityAdv;
            const T r = std::min(si, T(0));
            p[i].kappa = -r * p[i].factor;
        }
        if (taping) {
            StepTape& t = tapes.back();
            t.rhoadv.resize(n);
            t.mp0.resize(n);
            t.kp.assign(1, std::vector<T>(n));
            for (long i = 0; i < n; ++i) {
                t.rhoadv[i] = p[i].densityAdv;
                t.mp0[i] = T(1) - p[i].densityAdv < T(0);
                t.kp[0][i] = p[i].kappa;
            }
        }
        const T eta = (T)desc.MaxPressureSolverError * T(0.0001) * info.rho0;
        unsigned it = 0;
        T err = 0;
        while ((err > eta || it < desc.MinPressureSolverIterations) && it < desc.MaxPressureSolverIterations) {
#pragma omp parallel for schedule(static)
            for (long i = 0; i < n; ++i) p[i].ap = pressure_accel(i, false);
#pragma omp parallel for schedule(static)
            for (long i = 0; i < n; ++i) {  // PressureSolveIterationKernel :317-340
                const T f = density_pressure_force(i) * info.dt2;
                p[i].residuum = std::min(T(1) - p[i].densityAdv - f, T(0));
                p[i].kappa -= p[i].residuum * p[i].factor;
            }
            if (taping) {
                StepTape& t = tapes.back();
                t.mp.emplace_back(n);
                t.kp.emplace_back(n);
                for (long i = 0; i < n; ++i) {
                    t.mp.back()[i] = p[i].residuum < T(0);
                    t.kp.back()[i] = p[i].kappa;
                }
            }
            err = residual_error();
            ++it;
        }
        dbg.PressureSolverIterationCount = it;
        dbg.PressureSolverError = (float)err;
#pragma omp parallel for schedule(static)
        for (long i = 0; i < n; ++i) {  // ComputePressureAccelerationAndVelocityKernel :434-481
            p[i].ap = pressure_accel(i, false);
            p[i].v = p[i].v + info.dt * p[i].ap;
        }
    }

    // boundary tangential samples shared by :636-681, :706-752, :793-839
    struct Tangential {
        bool ok;
        V3<T> pd[4];
        V3<T> gw[4];
        T vol;
    };
    Tangential tangential(long i, const Body<T>& b) const {
        Tangential t;
        t.ok = false;
        const V3<T> xb = b.xb[i];
        const V3<T> xi = p[i].x;
        V3<T> normal = -(xi - xb);
        const T nl = length(normal);
        if (!(nl > T(0.0001))) return t;
        normal = normal / nl;
        V3<T> t1, t2;
        orthogonal_vectors(normal, t1, t2);
        const V3<T> pos[4] = {xb - t1 * info.td, xb + t1 * info.td, xb - t2 * info.td, xb + t2 * info.td};
        for (int k = 0; k < 4; ++k) {
            t.pd[k] = xi - pos[k];
            t.gw[k] = kernel.gradW(t.pd[k]);
        }
        t.vol = T(0.25) * b.vb[i];
        t.ok = true;
        return t;
    }

    // ComputeViscosity DFSPHImplementation.cu:577-615
    void viscosity_solve() {
        if (!desc.EnableViscositySolver) {
            dbg.ViscositySolverError = 0;
            dbg.ViscositySolverIterationCount = 0;
            return;
        }
        const long n = (long)p.size();
        const T d001 = T(0.01) * info.R2;
#pragma omp parallel for schedule(static)
        for (long i = 0; i < n; ++i) {  // ComputeViscosityPreconditionerKernel :604-684
            M3<T> M = mzero<T>();
            const V3<T> xi = p[i].x;
            for (unsigned j = 0; j < cnt(i); ++j) {
                const unsigned k = nbr(i, j);
                const V3<T> xij = xi - p[k].x;
                const V3<T> gw = kernel.gradW(xij);
                const T s = T(10) * info.mu * (info.m / p[k].density) / (length2(xij) + d001);
                const T a[3] = {xij.x, xij.y, xij.z}, g[3] = {gw.x, gw.y, gw.z};
                for (int r = 0; r < 3; ++r)
                    for (int c = 0; c < 3; ++c) M.m[3 * r + c] += s * (a[r] * g[c]);
            }
            if (info.muB != T(0)) {
                for (auto& b : bodies) {
                    if (!(b.vb[i] > T(0))) continue;
                    const Tangential t = tangential(i, b);
                    if (!t.ok) continue;
                    const T a[3] = {t.pd[0].x, t.pd[0].y, t.pd[0].z};  // the reference uses direction 1 for all four
                    for (int k = 0; k < 4; ++k) {
                        const T s = T(10) * info.muB * t.vol / (length2(t.pd[k]) + d001);
                        const T g[3] = {t.gw[k].x, t.gw[k].y, t.gw[k].z};
                        for (int r = 0; r < 3; ++r)
                            for (int c = 0; c < 3; ++c) M.m[3 * r + c] += s * (a[r] * g[c]);
                    }
                }
            }
            const T f = info.dt / p[i].density;
            M3<T> A;
            for (int e = 0; e < 9; ++e) A.m[e] = (e % 4 == 0 ? T(1) : T(0)) - f * M.m[e];
            minv[i] = inverse3(A);
            // ComputeViscosityGradientKernel :686-756 (the boundary rhs term is identically zero: dot(0, .) = 0)
            vb_b[i] = p[i].v;
            vb_g[i] = p[i].v + p[i].dv;
        }
        cg_solve();
#pragma omp parallel for schedule(static)
        for (long i = 0; i < n; ++i) {  // ApplyViscosityForceKernel :844-862
            const V3<T> d = vb_g[i] - p[i].v;
            p[i].a = p[i].a + (T(1) / info.dt) * d;
            p[i].dv = d;
        }
    }
    // ComputeMatrixVecProdFunctionKernel :758-842
    void matvec(const std::vector<V3<T>>& in, std::vector<V3<T>>& out) {
        const long n = (long)p.size();
        const T d001 = T(0.01) * info.R2;
#pragma omp parallel for schedule(static)
        for (long i = 0; i < n; ++i) {
            const V3<T> xi = p[i].x, vi = in[i];
            V3<T> acc = v3<T>(0, 0, 0);
            for (unsigned j = 0; j < cnt(i); ++j) {
                const unsigned k = nbr(i, j);
                const V3<T> xij = xi - p[k].x;
                const V3<T> gw = kernel.gradW(xij);
                const T s = T(10) * info.mu * (info.m / p[k].density) * dot3(vi - in[k], xij) / (length2(xij) + d001);
                acc = acc + s * gw;
            }
            if (info.muB != T(0)) {
                for (auto& b : bodies) {
                    if (!(b.vb[i] > T(0))) continue;
                    const Tangential t = tangential(i, b);
                    if (!t.ok) continue;
                    V3<T> s4 = v3<T>(0, 0, 0);
                    for (int k = 0; k < 4; ++k) {
                        const T s = T(10) * info.muB * t.vol * dot3(vi, t.pd[k]) / (length2(t.pd[k]) + d001);
                        s4 = k == 0 ? s * t.gw[k] : s4 + s * t.gw[k];
                    }
                    acc = acc + s4;
                }
            }
            out[i] = vi - (info.dt / p[i].density) * acc;
        }
    }
    static double dotsum(const std::vector<V3<T>>& a, const std::vector<V3<T>>& b) {
        double s = 0;
        for (size_t i = 0; i < a.size(); ++i) s += (double)dot3(a[i], b[i]);
        return s;
    }
    // SolveViscosity DFSPHImplementation.cu:617-808 (preconditioned CG)
    void cg_solve() {
        const long n = (long)p.size();
        dbg.ViscositySolverIterationCount = 0;
        matvec(vb_g, vb_tmp);
        for (long i = 0; i < n; ++i) vb_r[i] = vb_b[i] - vb_tmp[i];
        const T rhs2 = (T)dotsum(vb_b, vb_b);
        if (rhs2 == T(0)) {
            for (long i = 0; i < n; ++i) vb_g[i] = v3<T>(0, 0, 0);
            dbg.ViscositySolverError = 0;
            return;
        }
        const T tol = (T)desc.MaxViscositySolverError;
        const T thr = std::max(tol * tol * T(0.0001) * rhs2, std::numeric_limits<T>::min());
        T res2 = (T)dotsum(vb_r, vb_r);
        if (res2 < thr) {
            dbg.ViscositySolverError = (float)std::sqrt(res2 / rhs2);
            return;
        }
        for (long i = 0; i < n; ++i) vb_p[i] = mulmv(minv[i], vb_r[i]);
        T absNew = std::fabs((T)dotsum(vb_r, vb_p));
        unsigned it = 0;
        while (it >= desc.MinViscositySolverIterations && it < desc.MaxViscositySolverIterations) {
            matvec(vb_p, vb_ap);
            const T alpha = absNew / (T)dotsum(vb_p, vb_ap);
            for (long i = 0; i < n; ++i) {
                vb_g[i] = vb_g[i] + vb_p[i] * alpha;
                vb_r[i] = vb_r[i] - vb_ap[i] * alpha;
            }
            res2 = (T)dotsum(vb_r, vb_r);
            if (res2 < thr) break;
            for (long i = 0; i < n; ++i) vb_z[i] = mulmv(minv[i], vb_r[i]);
            const T absOld = absNew;
            absNew = std::fabs((T)dotsum(vb_r, vb_z));
            const T beta = absNew / absOld;
            for (long i = 0; i < n; ++i) vb_p[i] = vb_p[i] * beta + vb_z[i];
            ++it;
        }
        dbg.ViscositySolverIterationCount = it;
        dbg.ViscositySolverError = (float)std::sqrt(res2 / rhs2);
    }

    // SolveSurfaceTension DFSPHImplementation.cu:810-839; kernels DFSPHKernels.cu:864-1046
    void surface_tension() {
        if (!desc.EnableSurfaceTensionSolver) return;
        const long n = (long)p.size();
        const T R = info.R;
#pragma omp parallel for schedule(dynamic, 256)
        for (long i = 0; i < n; ++i) {  // ComputeSurfaceTensionClassificationKernel :881-958
            Particle<T>& q = p[i];
            q.nMC = v3<T>(0, 0, 0);
            const V3<T> xi = q.x;
            const unsigned c = cnt(i);
            if (c == 0) {
                q.kMC = T(1) / R;
                continue;
            }
            V3<T> com = v3<T>(0, 0, 0);
            for (unsigned j = 0; j < c; ++j) com = com + (p[nbr(i, j)].x - xi);
            com = com / R;
            const T ci = length(com) / (T)c;
            if (!((T)c <= info.slope * ci + info.constant + T(0))) continue;
            unsigned count = 0;
            const unsigned S = info.S;
            const unsigned s = (unsigned)i * S / 3u * 3u;
            const T rr = info.NR / info.r;
            const T lim = rr * rr * info.R2;
            V3<T> nsum = v3<T>(0, 0, 0);
            for (unsigned k = 0; k < S; ++k) {
                const unsigned i3 = s + 3u * k;
                const V3<T> pt = R * v3<T>(halton[i3 % kHaltonMod], halton[(i3 + 1) % kHaltonMod],
                                           halton[(i3 + 2) % kHaltonMod]);
                bool covered = false;
                for (unsigned j = 0; j < c; ++j) {
                    const V3<T> pd = p[nbr(i, j)].x - xi;
                    if (length2(pt - pd) <= lim) {
                        covered = true;
                        break;
                    }
                }
                if (covered) continue;
                nsum = nsum + pt;
                ++count;
            }
            if (count > 0) {
                if (length2(nsum) > T(0)) nsum = normalize(nsum);
                q.nMC = nsum;
                q.kMC = T(1) / R * T(-2) * std::sqrt(T(1) - info.NR * info.NR / (info.r * info.r)) *
                        std::cos(std::acos(T(1) - T(2) * ((T)count / (T)S)) + info.mcFactor);
            } else {
                q.nMC = v3<T>(0, 0, 0);
                q.kMC = T(0);
            }
        }
        const T sf = info.smoothing;
#pragma omp parallel for schedule(static)
        for (long i = 0; i < n; ++i) {  // ComputeSurfaceTensionNormalsAndCurvatureKernel :960-1006
            Particle<T>& q = p[i];
            if (is_zero(q.nMC)) continue;
            const V3<T> xi = q.x;
            V3<T> nc = v3<T>(0, 0, 0);
            T cc = 0, cf = 0;
            for (unsigned j = 0; j < cnt(i); ++j) {
                const Particle<T>& o = p[nbr(i, j)];
                if (is_zero(o.nMC)) continue;
                const T dist = length(o.x - xi);
                const T num = T(1) - dist / R;
                nc = nc + o.nMC * num;
                cc += o.kMC * num;
                cf += num;
            }
            if (length2(nc) > T(0)) nc = normalize(nc);
            V3<T> ns = (T(1) - sf) * q.nMC + sf * nc;
            if (length2(ns) > T(0)) ns = normalize(ns);
            q.nS = ns;
            q.kS = ((T(1) - sf) * q.kMC + sf * cc) / (T(1) - sf + sf * cf);
        }
        for (unsigned pass = 0; pass < desc.SurfaceTensionSmoothPassCount; ++pass) {
#pragma omp parallel for schedule(static)
            for (long i = 0; i < n; ++i) {  // ComputeSurfaceTensionBlendingKernel :1008-1046
                Particle<T>& q = p[i];
                if (!is_zero(q.nMC)) {
                    T fk = q.kS;
                    if (info.temporal) fk = T(0.05) * fk + T(0.95) * q.dK;
                    const V3<T> force = q.nS * info.sigma * fk;
                    q.a = q.a - info.mInv * force;
                    q.dK = fk;
                } else {
                    T fk = 0;
                    if (info.temporal) fk = T(0.95) * q.dK;
                    q.dK = fk;
                }
            }
        }
    }

    // ComputeTimeStepSize DFSPHImplementation.cu:395-441
    void time_step() {
        T mx = T(0.1);
        for (const auto& q : p) mx = std::max(mx, length2(q.v + q.a * info.dt));
        maxVel = mx;
        if (maxVel < T(1.0e-9)) maxVel = T(1.0e-9);
        T dt = T(0.4) * (info.diam / std::sqrt(maxVel));
        dt = std::min(dt, (T)desc.MaxTimeStepSize);
        dt = std::max(dt, (T)desc.MinTimeStepSize);
        set_dt(dt);
        update_sample_count();
        info.mcFactor = std::asin(info.NR / info.r);
        dbg.FrameTime += (float)info.dt;
    }

    // OnUpdate DFSPHImplementation.cu:63-170, stopping after `last` (VFD_PHASE_*)
    void step(int last = VFD_PHASE_POSITION) {
        if (p.empty()) return;
        const long n = (long)p.size();
        std::vector<V3<T>> xs(n);
        for (long i = 0; i < n; ++i) xs[i] = p[i].x;
        find_neighbors(xs, kernel.R, nb);
        if (last == VFD_PHASE_SEARCH) return;
        if (taping) tapePre = xs;
        boundary();
        if (last == VFD_PHASE_BOUNDARY) return;
        if (taping) tape_begin();
        density();
        if (last == VFD_PHASE_DENSITY) return;
        factor();
        if (taping) {
            tapes.back().alpha.resize(n);
            for (long i = 0; i < n; ++i) tapes.back().alpha[i] = p[i].factor;
        }
        if (last == VFD_PHASE_FACTOR) return;
        divergence_solve();
        if (taping) tape_velocity(tapes.back().v1);
        if (last == VFD_PHASE_DIVERGENCE) return;
        for (auto& q : p) q.a = info.g;  // ClearAccelerationKernel :24-37
        if (last == VFD_PHASE_CLEAR_ACCELERATION) return;
        if (taping && desc.EnableSurfaceTensionSolver && info.temporal) {
            StepTape& t = tapes.back();
            t.dK0.resize(n);
            for (long i = 0; i < n; ++i) t.dK0[i] = p[i].dK;
        }
        surface_tension();
        if (taping && desc.EnableSurfaceTensionSolver) {
            StepTape& t = tapes.back();
            t.st = true;
            t.temporal = info.temporal;
            t.stPasses = desc.SurfaceTensionSmoothPassCount;
            t.nMC.resize(n);
            t.kMC.resize(n);
            for (long i = 0; i < n; ++i) {
                t.nMC[i] = p[i].nMC;
                t.kMC[i] = p[i].kMC;
            }
        }
        if (last == VFD_PHASE_SURFACE_TENSION) return;
        viscosity_solve();
        if (taping) {
            tapes.back().visc = desc.EnableViscositySolver != 0;
            tapes.back().y = vb_g;
        }
        if (last == VFD_PHASE_VISCOSITY) return;
        time_step();
        if (last == VFD_PHASE_TIME_STEP) return;
        for (auto& q : p) q.v = q.v + info.dt * q.a;  // ComputeVelocityKernel :39-52
        if (taping) {
            tapes.back().H = info.dt;
            tape_velocity(tapes.back().v2);
        }
        if (last == VFD_PHASE_VELOCITY) return;
        pressure_solve();
        if (taping) tape_velocity(tapes.back().v3);
        if (last == VFD_PHASE_PRESSURE) return;
        for (auto& q : p) q.x = q.x + info.dt * q.v;  // ComputePositionKernel :54-67
        if (dbg.FrameIndex < desc.FrameCount && dbg.FrameTime >= desc.FrameLength) {
            dbg.FrameTime = 0;
            dbg.FrameIndex++;
        }
        dbg.IterationCount++;
    }

    // ==========================================================================================================
    // Adjoint of the step (reference gradient for the HIP adjoint; TEST INFRASTRUCTURE like the rest of this file).
    //
    // Differentiated map: one step (x, v; mu) -> (x', v') with the neighbour lists, the time steps, the solver
    // iteration counts, the Jacobi clamp masks and the boundary pass's branches (which particles sample which body,
    // which are pushed out, the max/min of the sample distance and the push depth) of the forward held fixed. Kernel
    // values are the forward's (table or spline); their derivatives are the spline's (dW/drho, d(gradW/r)/drho), so
    // with `smooth` kernels this is the exact derivative of the computed step and with the table it is the
    // derivative of the step the table approximates. The Jacobi solves are reversed iteration by iteration (their
    // transposed iteration), the viscosity solve by one transposed linear solve. Boundary contact enters through the
    // sample terms of density, factor, divergence and pressure acceleration (vb W(x - xb), vb gradW(x - xb)), whose
    // position derivatives follow the volume map's interpolant to second order (the normal's derivative is
    // (I - n n^T) Hess(phi) / |grad phi|), through the push-out's Jacobian and, with a boundary viscosity, through the
    // tangential samples' block of the viscosity matrix (tan_samples). Surface tension enters through its smoothing
    // pass (st_vjp): the Monte-Carlo normals and curvatures are piecewise constant in the positions (they change only
    // where a sample's coverage or the classification flips) and are held at their recorded values; with temporal
    // smoothing the blended curvature dK each particle carries from step to step has a cotangent of its own (ldK),
    // passed from step to step through the reversal (zero after the last step: the loss reads x and v only; the dK
    // carried into the first step is a constant). lmu is d/dmu of the fluid viscosity; the boundary viscosity is held.
    // Scope: no particle pushed out twice or pushed after a contact sample of an earlier body, symmetric lists (no
    // particle at the 70-neighbour cap); tape_begin reports anything outside it.
    // ==========================================================================================================
    std::string tape_error;
    void tape_begin() {
        tapes.emplace_back();
        StepTape& t = tapes.back();
        const long n = (long)p.size();
        t.nb = nb;
        t.x.resize(n);
        for (long i = 0; i < n; ++i) t.x[i] = p[i].x;
        tape_velocity(t.v0);
        for (long i = 0; i < n; ++i)
            if (cnt(i) >= kMaxNeighbors) tape_error = "a particle reached the neighbour cap (lists not symmetric)";
        tape_boundary(t);
    }
    // (I - n n^T) Hs / |g|: the derivative of n = g / |g| for g = grad phi, Hs = Hess(phi) as (xx, yy, zz, xy, xz, yz)
    static M3<T> normal_jacobian(V3<T> n, T gl, const T hs[6]) {
        const T Hm[9] = {hs[0], hs[3], hs[4], hs[3], hs[1], hs[5], hs[4], hs[5], hs[2]};
        const T nv[3] = {n.x, n.y, n.z};
        M3<T> J;
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) {
                T s = Hm[3 * r + c];
                for (int k = 0; k < 3; ++k) s -= nv[r] * nv[k] * Hm[3 * k + c];
                J.m[3 * r + c] = s / gl;
            }
        return J;
    }
    // Replays the boundary pass (boundary()) from the positions it started from, with the same branches, recording
    // each contact sample and each push-out with their derivatives.
    void tape_boundary(StepTape& t) {
        const long n = (long)p.size();
        t.contact.assign(bodies.size(), std::vector<typename StepTape::Contact>(n));
        t.pushed.assign(n, 0);
        t.Jpush.assign(n, mzero<T>());
        for (long i = 0; i < n; ++i) {
            V3<T> x = tapePre[i];
            bool sampled = false;
            for (size_t bi = 0; bi < bodies.size(); ++bi) {
                const Body<T>& b = bodies[bi];
                T phi, hs[6];
                V3<T> g;
                if (!b.map.derivatives(0, x, phi, g, hs)) continue;
                const T gl = length(g);
                if (b.vb[i] > T(0)) {
                    const V3<T> nn = g / gl;
                    const M3<T> Jn = normal_jacobian(nn, gl, hs);
                    const T s0 = phi + T(0.5) * info.r;
                    const T s = std::max(s0, info.diam);
                    const V3<T> ds = s0 < info.diam ? v3<T>(0, 0, 0) : g;
                    typename StepTape::Contact& c = t.contact[bi][i];
                    c.vb = b.vb[i];
                    T vol, hv[6];
                    b.map.derivatives(1, x, vol, c.dvb, hv);
                    c.r = x - b.xb[i];
                    const T nv[3] = {nn.x, nn.y, nn.z}, dv[3] = {ds.x, ds.y, ds.z};
                    for (int r = 0; r < 3; ++r)
                        for (int k = 0; k < 3; ++k) c.Jr.m[3 * r + k] = nv[r] * dv[k] + s * Jn.m[3 * r + k];
                    c.Jn = Jn;
                    sampled = true;
                } else if (phi <= T(0) && gl > T(1.0e-5)) {
                    if (sampled || t.pushed[i])
                        tape_error = "a particle pushed out twice or after a contact sample is outside the adjoint's scope";
                    const V3<T> nn = g / gl;
                    const M3<T> Jn = normal_jacobian(nn, gl, hs);
                    const T d0 = info.diam - phi;
                    const T delta = std::min(d0, T(0.1) * info.r);
                    const V3<T> dd = T(0.1) * info.r < d0 ? v3<T>(0, 0, 0) : -g;
                    const T nv[3] = {nn.x, nn.y, nn.z}, dv[3] = {dd.x, dd.y, dd.z};
                    for (int r = 0; r < 3; ++r)
                        for (int k = 0; k < 3; ++k)
                            t.Jpush[i].m[3 * r + k] = (r == k ? T(1) : T(0)) + nv[r] * dv[k] + delta * Jn.m[3 * r + k];
                    t.pushed[i] = 1;
                    x = x + delta * nn;
                }
            }
            if (x.x != p[i].x.x || x.y != p[i].x.y || x.z != p[i].x.z)
                tape_error = "the replay of the boundary pass differs from the forward";
        }
    }
    // boundary part of particle i's kernel-gradient sum: sum over bodies of vb gradW(x_i - xb)
    V3<T> contact_sum(const StepTape& t, long i) const {
        V3<T> s = v3<T>(0, 0, 0);
        for (const auto& cb : t.contact) {
            const auto& c = cb[i];
            if (c.vb > T(0)) s = s + (c.vb * kernel.gval(length(c.r))) * c.r;
        }
        return s;
    }
    // d/dx_i of q . contact_sum(i) for a fixed vector q
    V3<T> contact_vjp(const StepTape& t, long i, V3<T> q) const {
        V3<T> o = v3<T>(0, 0, 0);
        for (const auto& cb : t.contact) {
            const auto& c = cb[i];
            if (!(c.vb > T(0))) continue;
            const T g = kernel.gval(length(c.r));
            o = o + (g * dot3(c.r, q)) * c.dvb + mulmtv(c.Jr, vjp_grad(c.r, c.vb * q));
        }
        return o;
    }
    // sum over bodies of vb W(x_i - xb), and (grad != nullptr) its derivative in x_i
    T contact_density(const StepTape& t, long i, V3<T>* grad) const {
        T d = T(0);
        for (const auto& cb : t.contact) {
            const auto& c = cb[i];
            if (!(c.vb > T(0))) continue;
            const T w = kernel.getW(c.r);
            d += c.vb * w;
            if (grad) {
                const T rl = length(c.r);
                const T dw = rl > T(0) ? kernel.dW(rl) / rl : T(0);
                *grad = *grad + w * c.dvb + mulmtv(c.Jr, (c.vb * dw) * c.r);
            }
        }
        return d;
    }
    void tape_velocity(std::vector<V3<T>>& out) const {
        out.resize(p.size());
        for (size_t i = 0; i < p.size(); ++i) out[i] = p[i].v;
    }

    // gradient-of-gradW contraction: d(c . G(|r|) r)/dr = G c + G'(|r|) (c . r)/|r| r
    V3<T> vjp_grad(V3<T> r, V3<T> c) const {
        const T rl = std::sqrt(dot3(r, r));
        const T g = kernel.gval(rl);
        V3<T> o = g * c;
        if (rl > T(0)) o = o + (kernel.dG(rl) * dot3(c, r) / rl) * r;
        return o;
    }
    // A(k)_i = -V sum_j (k_i + k_j) gradW_ij   (pressure acceleration of a kappa field)
    void op_A(const StepTape& t, const std::vector<T>& k, std::vector<V3<T>>& out) const {
        const long n = (long)t.x.size();
        out.assign(n, v3<T>(0, 0, 0));
        for (long i = 0; i < n; ++i) {
            V3<T> a = v3<T>(0, 0, 0);
            for (unsigned jj = 0; jj < t.nb.counts[i]; ++jj) {
                const unsigned j = t.nb.ids[t.nb.offsets[i] + jj];
                const V3<T> r = t.x[i] - t.x[j];
                a = a + (-info.V * (k[i] + k[j]) * kernel.gval(std::sqrt(dot3(r, r)))) * r;
            }
            out[i] = a - k[i] * contact_sum(t, i);
        }
    }
    // D(u)_i = V sum_j (u_i - u_j) . gradW_ij   (divergence of a vector field)
    void op_D(const StepTape& t, const std::vector<V3<T>>& u, std::vector<T>& out) const {
        const long n = (long)t.x.size();
        out.assign(n, T(0));
        for (long i = 0; i < n; ++i) {
            T d = 0;
            for (unsigned jj = 0; jj < t.nb.counts[i]; ++jj) {
                const unsigned j = t.nb.ids[t.nb.offsets[i] + jj];
                const V3<T> r = t.x[i] - t.x[j];
                d += dot3(u[i] - u[j], kernel.gval(std::sqrt(dot3(r, r))) * r);
            }
            out[i] = info.V * d + dot3(u[i], contact_sum(t, i));
        }
    }
    // position cotangents of A(k) against la, and of D(u) against lf
    void vjp_x_A(const StepTape& t, const std::vector<T>& k, const std::vector<V3<T>>& la,
                 std::vector<V3<T>>& lx) const {
        const long n = (long)t.x.size();
        for (long i = 0; i < n; ++i) {
            for (unsigned jj = 0; jj < t.nb.counts[i]; ++jj) {
                const unsigned j = t.nb.ids[t.nb.offsets[i] + jj];
                const V3<T> g = vjp_grad(t.x[i] - t.x[j], (-info.V * (k[i] + k[j])) * la[i]);
                lx[i] = lx[i] + g;
                lx[j] = lx[j] - g;
            }
            lx[i] = lx[i] + contact_vjp(t, i, (-k[i]) * la[i]);
        }
    }
    void vjp_x_D(const StepTape& t, const std::vector<V3<T>>& u, const std::vector<T>& lf,
                 std::vector<V3<T>>& lx) const {
        const long n = (long)t.x.size();
        for (long i = 0; i < n; ++i) {
            for (unsigned jj = 0; jj < t.nb.counts[i]; ++jj) {
                const unsigned j = t.nb.ids[t.nb.offsets[i] + jj];
                const V3<T> g = vjp_grad(t.x[i] - t.x[j], (info.V * lf[i]) * (u[i] - u[j]));
                lx[i] = lx[i] + g;
                lx[j] = lx[j] - g;
            }
            lx[i] = lx[i] + contact_vjp(t, i, lf[i] * u[i]);
        }
    }
    // reverse of one Jacobi solve (divergence: sign = -1, source D; pressure: sign = +1, source 1 - rhoadv):
    // kappa^{k+1} = kappa^k - s o r^k, r^k = M^k o (src - c f^k), f^k = D(A(kappa^k)).
    // Accumulates lk0 (cotangent of kappa^0), lsrc (of the source) and ls (of the scale s) and positions.
    void reverse_jacobi(const StepTape& t, const std::vector<std::vector<T>>& K, const std::vector<std::vector<uint8_t>>& M,
                        const std::vector<T>& scale, const std::vector<T>& src, T c, std::vector<T>& lk,
                        std::vector<T>& lsrc, std::vector<T>& ls, std::vector<V3<T>>& lx) const {
        const long n = (long)t.x.size();
        std::vector<V3<T>> a, la;
        std::vector<T> f, lf(n), tmp;
        for (long k = (long)M.size() - 1; k >= 0; --k) {
            op_A(t, K[k], a);
            op_D(t, a, f);
            for (long i = 0; i < n; ++i) {
                const T r = M[k][i] ? src[i] - c * f[i] : T(0);
                const T lr = -scale[i] * lk[i];
                ls[i] += -r * lk[i];
                lsrc[i] += M[k][i] ? lr : T(0);
                lf[i] = M[k][i] ? -c * lr : T(0);
            }
            // f = D(a): la = D^T lf = -A(lf)
            op_A(t, lf, la);
            for (auto& q : la) q = -q;
            vjp_x_D(t, a, lf, lx);
            // a = A(kappa): lk += A^T la = -D(la)
            op_D(t, la, tmp);
            for (long i = 0; i < n; ++i) lk[i] -= tmp[i];
            vjp_x_A(t, K[k], la, lx);
        }
    }
    // Boundary viscosity samples of a contact (tangential(), DFSPHKernels.cu:636-681, 793-839): the offsets
    // pd_k = r + td t1, r - td t1, r + td t2, r - td t2 from the particle to the four samples around xb, with
    // t1 = normalize(N x v), t2 = normalize(N x (N x v)), N = -r / |r| and v the axis orthogonal_vectors picks (held),
    // and their derivatives in the particle's position: d pd_k / dx = dr/dx +- td dt/dN dN/dx, dN/dx = -dn/dx.
    struct TanSamples {
        bool ok = false;
        V3<T> pd[4];
        M3<T> J[4];
    };
    TanSamples tan_samples(const typename StepTape::Contact& c) const {
        TanSamples o;
        V3<T> N = -c.r;
        const T nl = length(N);
        if (!(nl > T(0.0001))) return o;
        N = N / nl;
        V3<T> v = v3<T>(1, 0, 0);
        if ((double)std::fabs(dot3(v, N)) > 0.999) v = v3<T>(0, 1, 0);
        const V3<T> x = cross(N, v), y = cross(N, x);
        const T xl = std::sqrt(dot3(x, x)), yl = std::sqrt(dot3(y, y));
        const V3<T> t1 = normalize(x), t2 = normalize(y);
        const M3<T> dxdN = mlin(T(-1), skew(v), T(0), skew(v));
        const M3<T> dydN = mlin(T(-1), skew(x), T(-1), mmul(skew(N), skew(v)));
        const M3<T> dNdx = mlin(T(-1), c.Jn, T(0), c.Jn);
        const M3<T> dt1 = mmul(mmul(normalize_jacobian(t1, xl), dxdN), dNdx);
        const M3<T> dt2 = mmul(mmul(normalize_jacobian(t2, yl), dydN), dNdx);
        const T td = info.td;
        o.pd[0] = c.r + td * t1;
        o.pd[1] = c.r - td * t1;
        o.pd[2] = c.r + td * t2;
        o.pd[3] = c.r - td * t2;
        o.J[0] = mlin(T(1), c.Jr, td, dt1);
        o.J[1] = mlin(T(1), c.Jr, -td, dt1);
        o.J[2] = mlin(T(1), c.Jr, td, dt2);
        o.J[3] = mlin(T(1), c.Jr, -td, dt2);
        o.ok = true;
        return o;
    }
    // C w = sum_k c_k G(|pd_k|) pd_k (pd_k . w), c_k = 10 mu_b (vb / 4) / (|pd_k|^2 + 0.01 R^2): the boundary block of
    // the viscosity matrix of particle i (symmetric), summed over the bodies it samples
    V3<T> visc_boundary_apply(const StepTape& t, long i, V3<T> w) const {
        V3<T> o = v3<T>(0, 0, 0);
        if (info.muB == T(0)) return o;
        for (const auto& cb : t.contact) {
            const auto& c = cb[i];
            if (!(c.vb > T(0))) continue;
            const TanSamples s = tan_samples(c);
            if (!s.ok) continue;
            const T vol = T(0.25) * c.vb;
            for (int k = 0; k < 4; ++k) {
                const T q = length2(s.pd[k]) + T(0.01) * info.R2;
                const T ck = T(10) * info.muB * vol / q;
                o = o + (ck * kernel.gval(std::sqrt(dot3(s.pd[k], s.pd[k]))) * dot3(s.pd[k], w)) * s.pd[k];
            }
        }
        return o;
    }
    // viscosity system of the tape: (A y)_i = y_i - sum_j T_ij,
    // T_ij = (h / rho_i) 10 mu m / (rho_j (|r|^2 + 0.01 R^2)) G(|r|) ((y_i - y_j) . r) r
    T visc_k(const StepTape& t, const std::vector<T>& rho, long i, unsigned j, V3<T> r) const {
        return (t.h / rho[i]) * (T(10) * info.mu * info.m / rho[j]) * kernel.gval(std::sqrt(dot3(r, r))) /
               (length2(r) + T(0.01) * info.R2);
    }
    void visc_apply_T(const StepTape& t, const std::vector<T>& rho, const std::vector<V3<T>>& w,
                      std::vector<V3<T>>& out) const {  // A^T w
        const long n = (long)t.x.size();
        out = w;
        for (long i = 0; i < n; ++i)
            for (unsigned jj = 0; jj < t.nb.counts[i]; ++jj) {
                const unsigned j = t.nb.ids[t.nb.offsets[i] + jj];
                const V3<T> r = t.x[i] - t.x[j];
                // (A^T w)_i -= r r^T (k_ij w_i - k_ji w_j)
                const V3<T> u = visc_k(t, rho, i, j, r) * w[i] - visc_k(t, rho, (long)j, (unsigned)i, -r) * w[j];
                out[i] = out[i] - dot3(r, u) * r;
            }
        for (long i = 0; i < n; ++i)  // boundary block: (A^T w)_i -= (h / rho_i) C_i w_i (C_i symmetric)
            out[i] = out[i] - (t.h / rho[i]) * visc_boundary_apply(t, i, w[i]);
    }
    // BiCGSTAB for A^T x = b (double precision, relative residual 1e-12)
    void solve_visc_T(const StepTape& t, const std::vector<T>& rho, const std::vector<V3<T>>& b,
                      std::vector<V3<T>>& x) const {
        const long n = (long)b.size();
        auto dot = [&](const std::vector<V3<T>>& a, const std::vector<V3<T>>& c) {
            double s = 0;
            for (long i = 0; i < n; ++i) s += (double)dot3(a[i], c[i]);
            return s;
        };
        x = b;
        std::vector<V3<T>> r(n), rh, pv(n, v3<T>(0, 0, 0)), v(n, v3<T>(0, 0, 0)), sv(n), tv(n), ax;
        visc_apply_T(t, rho, x, ax);
        for (long i = 0; i < n; ++i) r[i] = b[i] - ax[i];
        rh = r;
        double rho_old = 1, alpha = 1, omega = 1;
        const double bn = std::sqrt(dot(b, b));
        if (bn == 0) return;
        for (int it = 0; it < 1000 && std::sqrt(dot(r, r)) > 1e-12 * bn; ++it) {
            const double rho_new = dot(rh, r);
            const double beta = (rho_new / rho_old) * (alpha / omega);
            for (long i = 0; i < n; ++i) pv[i] = r[i] + (T)beta * (pv[i] - (T)omega * v[i]);
            visc_apply_T(t, rho, pv, v);
            alpha = rho_new / dot(rh, v);
            for (long i = 0; i < n; ++i) sv[i] = r[i] - (T)alpha * v[i];
            visc_apply_T(t, rho, sv, tv);
            omega = dot(tv, sv) / dot(tv, tv);
            for (long i = 0; i < n; ++i) {
                x[i] = x[i] + (T)alpha * pv[i] + (T)omega * sv[i];
                r[i] = sv[i] - (T)omega * tv[i];
            }
            rho_old = rho_new;
        }
    }
    // Surface tension of a taped step as a function of the positions, with nMC, kMC and the classification held:
    // for a particle with nMC_i != 0 (ComputeSurfaceTensionNormalsAndCurvatureKernel :960-1006, blending :1008-1046)
    //   w_ij = 1 - |x_i - x_j| / R over neighbours with nMC_j != 0,  nc = sum w nMC_j,  cc = sum w kMC_j,  cf = sum w,
    //   nS = normalize((1 - sf) nMC_i + sf normalize(nc)),  kS = ((1 - sf) kMC_i + sf cc) / (1 - sf + sf cf),
    //   a_i = -(sigma / m) E_i nS with E_i = sum over the P blend passes of the pass's curvature fk_p.
    // Without temporal smoothing fk_p = kS (E = P kS). With it fk_p = 0.05 kS + 0.95 fk_(p-1) from fk_0 = dK_i, the
    // curvature the particle carried in, so E = A dK_i + (P - A) kS with A = sum_(p=1..P) 0.95^p, and the particle
    // carries out dK' = 0.95^P dK_i + (1 - 0.95^P) kS (a particle with nMC_i = 0: no force, dK' = 0.95^P dK_i).
    // Without temporal smoothing dK' = kS (0 for nMC_i = 0) whatever dK was. Accumulates the position cotangent of a
    // against la and maps ldK, the cotangent of the carried curvature, from dK' to dK.
    void st_vjp(const StepTape& t, const std::vector<V3<T>>& la, std::vector<V3<T>>& lx, std::vector<T>& ldK) const {
        const long n = (long)t.x.size();
        const T sf = info.smoothing, R = info.R;
        const T c = -info.sigma * info.mInv;
        const T P = (T)t.stPasses;
        T A = 0, aP = 1;
        for (unsigned q = 0; q < t.stPasses; ++q) {
            aP *= T(0.95);
            A += aP;
        }
        std::vector<V3<T>> lnc(n, v3<T>(0, 0, 0));
        std::vector<T> lcc(n, T(0)), lcf(n, T(0));
        auto nrm_vjp = [](V3<T> a, V3<T> l) {  // d normalize(a) ^T l (identity where a = 0, as the forward keeps a)
            const T a2 = length2(a);
            if (!(a2 > T(0))) return l;
            const V3<T> u = normalize(a);
            return (l - dot3(u, l) * u) / std::sqrt(a2);
        };
        for (long i = 0; i < n; ++i) {
            if (is_zero(t.nMC[i])) continue;
            V3<T> nc = v3<T>(0, 0, 0);
            T cc = 0, cf = 0;
            for (unsigned jj = 0; jj < t.nb.counts[i]; ++jj) {
                const unsigned j = t.nb.ids[t.nb.offsets[i] + jj];
                if (is_zero(t.nMC[j])) continue;
                const T w = T(1) - length(t.x[j] - t.x[i]) / R;
                nc = nc + t.nMC[j] * w;
                cc += t.kMC[j] * w;
                cf += w;
            }
            const V3<T> m = length2(nc) > T(0) ? normalize(nc) : nc;
            const V3<T> u = (T(1) - sf) * t.nMC[i] + sf * m;
            const V3<T> ns = length2(u) > T(0) ? normalize(u) : u;
            const T den = T(1) - sf + sf * cf;
            const T ks = ((T(1) - sf) * t.kMC[i] + sf * cc) / den;
            const T lan = dot3(la[i], ns);
            T E, lks;
            if (t.temporal) {
                E = A * t.dK0[i] + (P - A) * ks;
                lks = c * (P - A) * lan + (T(1) - aP) * ldK[i];
                ldK[i] = c * A * lan + aP * ldK[i];
            } else {  // dK' = kS: the carried curvature in is overwritten
                E = P * ks;
                lks = c * P * lan + ldK[i];
                ldK[i] = T(0);
            }
            lnc[i] = nrm_vjp(nc, sf * nrm_vjp(u, (c * E) * la[i]));
            lcc[i] = lks * sf / den;
            lcf[i] = -lks * ks * sf / den;
        }
        for (long i = 0; i < n; ++i)
            if (is_zero(t.nMC[i])) ldK[i] = t.temporal ? aP * ldK[i] : T(0);  // dK' = 0.95^P dK, or 0
        for (long i = 0; i < n; ++i) {
            if (is_zero(t.nMC[i])) continue;
            for (unsigned jj = 0; jj < t.nb.counts[i]; ++jj) {
                const unsigned j = t.nb.ids[t.nb.offsets[i] + jj];
                if (is_zero(t.nMC[j])) continue;
                const V3<T> d = t.x[i] - t.x[j];
                const T dl = length(d);
                if (!(dl > T(0))) continue;
                const T lw = dot3(lnc[i], t.nMC[j]) + lcc[i] * t.kMC[j] + lcf[i];
                const V3<T> g = (-lw / (R * dl)) * d;
                lx[i] = lx[i] + g;
                lx[j] = lx[j] - g;
            }
        }
    }
    // Reverse of one taped step: (lx, lv, ldK) are the cotangents of the step's output (x', v', dK') on entry and of
    // its input (x, v, dK) on return; lmu accumulates d/dmu.
    void reverse_step(const StepTape& t, std::vector<V3<T>>& lx, std::vector<V3<T>>& lv, std::vector<T>& ldK,
                      double& lmu) const {
        const long n = (long)t.x.size();
        const T h = t.h, H = t.H;
        std::vector<T> lrho(n, T(0)), lalpha(n, T(0)), tmp;
        std::vector<V3<T>> la(n), tmpv;
        // x' = x + H v3
        std::vector<V3<T>> lv3(n);
        for (long i = 0; i < n; ++i) lv3[i] = lv[i] + H * lx[i];  // lx keeps flowing to x
        // pressure final: v3 = v2 + H A(kp^K)
        std::vector<V3<T>> lv2 = lv3;
        for (long i = 0; i < n; ++i) la[i] = H * lv3[i];
        std::vector<T> lk(n);
        op_D(t, la, tmp);
        for (long i = 0; i < n; ++i) lk[i] = -tmp[i];
        vjp_x_A(t, t.kp.back(), la, lx);
        // pressure iterations: scale gamma = alpha / H^2, source 1 - rhoadv, c = H^2
        std::vector<T> gamma(n), srcp(n), lsrc(n, T(0)), lgamma(n, T(0));
        for (long i = 0; i < n; ++i) {
            gamma[i] = t.alpha[i] / (H * H);
            srcp[i] = T(1) - t.rhoadv[i];
        }
        reverse_jacobi(t, t.kp, t.mp, gamma, srcp, H * H, lk, lsrc, lgamma, lx);
        // pressure init: kp^0 = mp0 (rhoadv - 1) gamma
        std::vector<T> lrhoadv(n);
        for (long i = 0; i < n; ++i) {
            lrhoadv[i] = -lsrc[i] + (t.mp0[i] ? gamma[i] * lk[i] : T(0));
            lgamma[i] += t.mp0[i] ? (t.rhoadv[i] - T(1)) * lk[i] : T(0);
            lalpha[i] += lgamma[i] / (H * H);
        }
        // rhoadv = rho / rho0 + H D(v2)
        std::vector<T> ld2(n);
        for (long i = 0; i < n; ++i) {
            lrho[i] += lrhoadv[i] / info.rho0;
            ld2[i] = H * lrhoadv[i];
        }
        op_A(t, ld2, tmpv);  // D^T ld2 = -A(ld2)
        for (long i = 0; i < n; ++i) lv2[i] = lv2[i] - tmpv[i];
        vjp_x_D(t, t.v2, ld2, lx);
        // v2 includes H a_st(x) (surface tension)
        if (t.st && t.stPasses) {
            std::vector<V3<T>> las(n);
            for (long i = 0; i < n; ++i) las[i] = H * lv2[i];
            st_vjp(t, las, lx, ldK);
        }
        // v2 = v1 + H g + (H / h)(y - v1)
        std::vector<V3<T>> ly(n), lv1(n);
        for (long i = 0; i < n; ++i) {
            ly[i] = t.visc ? (H / h) * lv2[i] : v3<T>(0, 0, 0);
            lv1[i] = t.visc ? (T(1) - H / h) * lv2[i] : lv2[i];
        }
        // densities (for the viscosity terms)
        std::vector<T> rho(n);
        for (long i = 0; i < n; ++i) {
            T d = info.V * kernel.W0;
            for (unsigned jj = 0; jj < t.nb.counts[i]; ++jj)
                d += info.V * kernel.getW(t.x[i] - t.x[t.nb.ids[t.nb.offsets[i] + jj]]);
            d += contact_density(t, i, nullptr);
            rho[i] = d * info.rho0;
        }
        if (t.visc) {
            // A y = v1: lb = A^-T ly; parameters through -lb^T (dA) y = + lb . dT
            std::vector<V3<T>> lb;
            solve_visc_T(t, rho, ly, lb);
            for (long i = 0; i < n; ++i) lv1[i] = lv1[i] + lb[i];
            const T e = T(0.01) * info.R2;
            for (long i = 0; i < n; ++i)
                for (unsigned jj = 0; jj < t.nb.counts[i]; ++jj) {
                    const unsigned j = t.nb.ids[t.nb.offsets[i] + jj];
                    const V3<T> r = t.x[i] - t.x[j];
                    const V3<T> u = t.y[i] - t.y[j];
                    const T k = visc_k(t, rho, i, j, r);
                    const T ur = dot3(u, r), sr = dot3(lb[i], r);
                    const T lT = k * ur * sr;  // lb_i . T_ij
                    lmu += (double)(lT / info.mu);
                    lrho[i] -= lT / rho[i];
                    lrho[j] -= lT / rho[(long)j];
                    const T rl = std::sqrt(dot3(r, r));
                    const T C = (h / rho[i]) * (T(10) * info.mu * info.m / rho[j]);
                    const T g = kernel.gval(rl), d2 = rl * rl + e;
                    const T dk = C * (kernel.dG(rl) * d2 - g * T(2) * rl) / (d2 * d2);
                    V3<T> gr = k * (ur * lb[i] + sr * u);
                    if (rl > T(0)) gr = gr + (ur * sr * dk / rl) * r;
                    lx[i] = lx[i] + gr;
                    lx[j] = lx[j] - gr;
                }
            // boundary samples: lb_i . (h / rho_i) C_i(x_i) y_i, through rho_i, the sample offsets and vb
            if (info.muB != T(0))
                for (long i = 0; i < n; ++i) {
                    const T hr = h / rho[i];
                    lrho[i] -= hr * dot3(lb[i], visc_boundary_apply(t, i, t.y[i])) / rho[i];
                    for (const auto& cb : t.contact) {
                        const auto& c = cb[i];
                        if (!(c.vb > T(0))) continue;
                        const TanSamples s = tan_samples(c);
                        if (!s.ok) continue;
                        const T vol = T(0.25) * c.vb;
                        for (int k = 0; k < 4; ++k) {
                            const V3<T> pd = s.pd[k];
                            const T pl = std::sqrt(dot3(pd, pd)), q = pl * pl + T(0.01) * info.R2;
                            const T ck = T(10) * info.muB * vol / q, g = kernel.gval(pl);
                            const T lp = dot3(lb[i], pd), py = dot3(pd, t.y[i]);
                            const T fk = ck * g * lp * py;
                            V3<T> gp = (ck * g) * (py * lb[i] + lp * t.y[i]) + (lp * py * g * (T(-2) * ck / q)) * pd;
                            if (pl > T(0)) gp = gp + (lp * py * ck * kernel.dG(pl) / pl) * pd;
                            lx[i] = lx[i] + hr * (mulmtv(s.J[k], gp) + (fk / c.vb) * c.dvb);
                        }
                    }
                }
        }
        // divergence final: v1 = v0 + h A(kv^K)
        for (long i = 0; i < n; ++i) {
            lv[i] = lv1[i];
            la[i] = h * lv1[i];
        }
        op_D(t, la, tmp);
        for (long i = 0; i < n; ++i) lk[i] = -tmp[i];
        vjp_x_A(t, t.kv.back(), la, lx);
        // divergence iterations: scale beta = alpha / h, source -D, c = h
        std::vector<T> beta(n), srcd(n), lsrcd(n, T(0)), lbeta(n, T(0));
        for (long i = 0; i < n; ++i) {
            beta[i] = t.alpha[i] / h;
            srcd[i] = -t.D[i];
        }
        reverse_jacobi(t, t.kv, t.mv, beta, srcd, h, lk, lsrcd, lbeta, lx);
        // divergence init: kv^0 = D beta, D = m0 D(v0)
        std::vector<T> ld(n);
        for (long i = 0; i < n; ++i) {
            const T lD = -lsrcd[i] + beta[i] * lk[i];
            lbeta[i] += t.D[i] * lk[i];
            ld[i] = t.m0[i] ? lD : T(0);
            lalpha[i] += lbeta[i] / h;
        }
        op_A(t, ld, tmpv);
        for (long i = 0; i < n; ++i) lv[i] = lv[i] - tmpv[i];
        vjp_x_D(t, t.v0, ld, lx);
        // factor: alpha = 1 / S, S_i = V^2 sum |gradW_ij|^2 + |V sum_j gradW_ij|^2
        for (long i = 0; i < n; ++i) {
            if (t.alpha[i] == T(0)) continue;
            const T lS = -t.alpha[i] * t.alpha[i] * lalpha[i];
            V3<T> gs = v3<T>(0, 0, 0);
            for (unsigned jj = 0; jj < t.nb.counts[i]; ++jj) {
                const V3<T> r = t.x[i] - t.x[t.nb.ids[t.nb.offsets[i] + jj]];
                gs = gs + (info.V * kernel.gval(std::sqrt(dot3(r, r)))) * r;
            }
            gs = gs + contact_sum(t, i);
            lx[i] = lx[i] + contact_vjp(t, i, (T(2) * lS) * gs);
            for (unsigned jj = 0; jj < t.nb.counts[i]; ++jj) {
                const unsigned j = t.nb.ids[t.nb.offsets[i] + jj];
                const V3<T> r = t.x[i] - t.x[j];
                const V3<T> gw = kernel.gval(std::sqrt(dot3(r, r))) * r;
                const V3<T> g = vjp_grad(r, (T(2) * info.V * lS) * (info.V * gw + gs));
                lx[i] = lx[i] + g;
                lx[j] = lx[j] - g;
            }
        }
        // density: rho_i = rho0 (V (W0 + sum_j W_ij) + sum_b vb W(x_i - xb))
        for (long i = 0; i < n; ++i) {
            for (unsigned jj = 0; jj < t.nb.counts[i]; ++jj) {
                const unsigned j = t.nb.ids[t.nb.offsets[i] + jj];
                const V3<T> r = t.x[i] - t.x[j];
                const T rl = std::sqrt(dot3(r, r));
                if (!(rl > T(0))) continue;
                const V3<T> g = (info.rho0 * info.V * lrho[i] * kernel.dW(rl) / rl) * r;
                lx[i] = lx[i] + g;
                lx[j] = lx[j] - g;
            }
            V3<T> gb = v3<T>(0, 0, 0);
            contact_density(t, i, &gb);
            lx[i] = lx[i] + (info.rho0 * lrho[i]) * gb;
        }
        // boundary push-out: x = x_before + delta n(x_before), v = 0
        for (long i = 0; i < n; ++i)
            if (t.pushed[i]) {
                lx[i] = mulmtv(t.Jpush[i], lx[i]);
                lv[i] = v3<T>(0, 0, 0);
            }
    }

    // Dense viscosity system of the current state: A (3n x 3n, row-major) built by applying the matvec to unit
    // vectors, and b (the right-hand side, = v). Test aid for tiny n.
    void assemble_viscosity(double* A, double* b) {
        const long n = (long)p.size();
        std::vector<V3<T>> in(n, v3<T>(0, 0, 0)), out(n);
        for (long k = 0; k < n; ++k) {
            for (int c = 0; c < 3; ++c) {
                in[k] = v3<T>(c == 0 ? T(1) : T(0), c == 1 ? T(1) : T(0), c == 2 ? T(1) : T(0));
                matvec(in, out);
                for (long i = 0; i < n; ++i) {
                    A[(3 * i + 0) * 3 * n + 3 * k + c] = (double)out[i].x;
                    A[(3 * i + 1) * 3 * n + 3 * k + c] = (double)out[i].y;
                    A[(3 * i + 2) * 3 * n + 3 * k + c] = (double)out[i].z;
                }
                in[k] = v3<T>(0, 0, 0);
            }
        }
        for (long i = 0; i < n; ++i) {
            b[3 * i] = (double)p[i].v.x;
            b[3 * i + 1] = (double)p[i].v.y;
            b[3 * i + 2] = (double)p[i].v.z;
        }
    }

    // field access -------------------------------------------------------------------------------------------
    int get_field(int field, unsigned body, double* out) const {
        const size_t n = p.size();
        auto put3 = [&](size_t i, V3<T> v) {
            out[3 * i] = (double)v.x;
            out[3 * i + 1] = (double)v.y;
            out[3 * i + 2] = (double)v.z;
        };
        for (size_t i = 0; i < n; ++i) {
            const Particle<T>& q = p[i];
            switch (field) {
                case VFD_FIELD_POSITION: put3(i, q.x); break;
                case VFD_FIELD_VELOCITY: put3(i, q.v); break;
                case VFD_FIELD_ACCELERATION: put3(i, q.a); break;
                case VFD_FIELD_PRESSURE_ACCELERATION: put3(i, q.ap); break;
                case VFD_FIELD_PRESSURE_RESIDUUM: out[i] = (double)q.residuum; break;
                case VFD_FIELD_DENSITY: out[i] = (double)q.density; break;
                case VFD_FIELD_DENSITY_ADVECTION: out[i] = (double)q.densityAdv; break;
                case VFD_FIELD_PRESSURE_RHO2: out[i] = (double)q.kappa; break;
                case VFD_FIELD_PRESSURE_RHO2_V: out[i] = (double)q.kappaV; break;
                case VFD_FIELD_FACTOR: out[i] = (double)q.factor; break;
                case VFD_FIELD_VELOCITY_DIFFERENCE: put3(i, q.dv); break;
                case VFD_FIELD_MC_SURFACE_NORMAL: put3(i, q.nMC); break;
                case VFD_FIELD_MC_SURFACE_NORMAL_SMOOTH: put3(i, q.nS); break;
                case VFD_FIELD_MC_SURFACE_CURVATURE: out[i] = (double)q.kMC; break;
                case VFD_FIELD_MC_SURFACE_CURVATURE_SMOOTH: out[i] = (double)q.kS; break;
                case VFD_FIELD_DELTA_FINAL_CURVATURE: out[i] = (double)q.dK; break;
                case VFD_FIELD_BOUNDARY_XJ:
                    if (body >= bodies.size()) return VFD_ERROR_INVALID_ARGUMENT;
                    put3(i, bodies[body].xb[i]);
                    break;
                case VFD_FIELD_BOUNDARY_VOLUME:
                    if (body >= bodies.size()) return VFD_ERROR_INVALID_ARGUMENT;
                    out[i] = (double)bodies[body].vb[i];
                    break;
                case VFD_FIELD_VISCOSITY_SOLUTION: put3(i, vb_g[i]); break;
                case VFD_FIELD_VISCOSITY_PRECONDITIONER:
                    for (int e = 0; e < 9; ++e) out[9 * i + e] = (double)minv[i].m[e];
                    break;
                default: return VFD_ERROR_INVALID_ARGUMENT;
            }
        }
        return VFD_OK;
    }
    int set_field(int field, const double* in) {
        const size_t n = p.size();
        if (field == VFD_FIELD_DELTA_FINAL_CURVATURE) {
            for (size_t i = 0; i < n; ++i) p[i].dK = (T)in[i];
            return VFD_OK;
        }
        for (size_t i = 0; i < n; ++i) {
            const V3<T> v = v3<T>((T)in[3 * i], (T)in[3 * i + 1], (T)in[3 * i + 2]);
            switch (field) {
                case VFD_FIELD_POSITION: p[i].x = v; break;
                case VFD_FIELD_VELOCITY: p[i].v = v; break;
                case VFD_FIELD_VELOCITY_DIFFERENCE: p[i].dv = v; break;
                default: return VFD_ERROR_INVALID_ARGUMENT;
            }
        }
        return VFD_OK;
    }
};

struct Handle {
    int fp64;
    Sim<float>* f;
    Sim<double>* d;
};

}  // namespace orc

using namespace orc;

#define DISPATCH(h, ...)                            \
    do {                                            \
        Handle* hh = (Handle*)(h);                  \
        if (!hh) return VFD_ERROR_INVALID_ARGUMENT; \
        if (hh->fp64) {                             \
            auto* s = hh->d;                        \
            __VA_ARGS__;                            \
        } else {                                    \
            auto* s = hh->f;                        \
            __VA_ARGS__;                            \
        }                                           \
    } while (0)

extern "C" {

void* orc_create(const vfd_desc* desc, int fp64) {
    Handle* h = new Handle{fp64, nullptr, nullptr};
    if (fp64)
        h->d = new Sim<double>(*desc);
    else
        h->f = new Sim<float>(*desc);
    return h;
}
void orc_destroy(void* hv) {
    Handle* h = (Handle*)hv;
    if (!h) return;
    delete h->f;
    delete h->d;
    delete h;
}
int orc_set_description(void* h, const vfd_desc* d) { DISPATCH(h, s->set_description(*d)); return VFD_OK; }
int orc_set_fluid(void* h, const float* x, const float* v, uint32_t n) {
    DISPATCH(h, s->set_fluid(x, v, n));
    return VFD_OK;
}
int orc_add_boundary(void* h, const vfd_boundary_desc* b) { DISPATCH(h, s->add_boundary(*b)); return VFD_OK; }
// RigidBody ctor for a mesh body (RigidBody.cu:10-73); the mesh is the caller's (indices checked by the caller)
int orc_add_boundary_mesh(void* h, const vfd_mesh_boundary_desc* m, const float* verts, uint32_t nv,
                          const uint32_t* tris, uint32_t nt) {
    vfd_boundary_desc d{};
    d.Shape = -1;
    d.Inverted = m->Inverted;
    d.Padding = m->Padding;
    for (int a = 0; a < 3; ++a) d.CollisionMapResolution[a] = m->CollisionMapResolution[a];
    DISPATCH(h, {
        using T = typename std::remove_reference<decltype(*s)>::type::Real;
        TriMesh<T> mesh;
        mesh.build(verts, nv, tris, nt, m->Transform);
        s->add_boundary(d, &mesh);
    });
    return VFD_OK;
}
// FluidObject ctor + SampleMeshVolume (FluidObject.cpp:6-24): transform (column-major, NULL = identity), sample.
int orc_sample_mesh_volume(int fp64, const float* verts, uint32_t nv, const uint32_t* tris, uint32_t nt,
                           const float* transform, float radius, const uint32_t* res, int inverted, int mode,
                           double* out, uint32_t capacity, uint32_t* count) {
    auto run = [&](auto tag) {
        using T = decltype(tag);
        TriMesh<T> mesh;
        mesh.build(verts, nv, tris, nt, transform);
        std::vector<T> pts;
        const unsigned r[3] = {res[0], res[1], res[2]};
        sample_mesh_volume<T>(mesh, (T)radius, r, inverted != 0, mode, pts);
        *count = (uint32_t)(pts.size() / 3);
        for (size_t k = 0; out && k < std::min<size_t>(pts.size(), 3 * (size_t)capacity); ++k) out[k] = (double)pts[k];
    };
    if (fp64) run(double()); else run(float());
    return VFD_OK;
}
// MeshDistance::SignedDistance of the transformed mesh at n points (float points for the FP32 mirror)
int orc_mesh_signed_distance(int fp64, const float* verts, uint32_t nv, const uint32_t* tris, uint32_t nt,
                             const float* transform, const double* pts, uint32_t n, double* out) {
    auto run = [&](auto tag) {
        using T = decltype(tag);
        TriMesh<T> mesh;
        mesh.build(verts, nv, tris, nt, transform);
#pragma omp parallel for schedule(dynamic, 64)
        for (long i = 0; i < (long)n; ++i)
            out[i] = (double)mesh.signed_distance(v3<T>((T)pts[3 * i], (T)pts[3 * i + 1], (T)pts[3 * i + 2]));
    };
    if (fp64) run(double()); else run(float());
    return VFD_OK;
}
int orc_step(void* h) { DISPATCH(h, s->step()); return VFD_OK; }
int orc_step_until(void* h, int phase) { DISPATCH(h, s->step(phase)); return VFD_OK; }
int orc_set_time_step(void* h, double dt) {
    DISPATCH(h, { s->set_dt(dt); s->update_sample_count(); });
    return VFD_OK;
}
int orc_get_field(void* h, int field, uint32_t body, double* out) {
    DISPATCH(h, return s->get_field(field, body, out));
    return VFD_OK;
}
int orc_set_field(void* h, int field, const double* in) {
    DISPATCH(h, return s->set_field(field, in));
    return VFD_OK;
}
// ---- adjoint (reference gradient) -------------------------------------------------------------------------------
int orc_set_smooth_kernel(void* h, int on) { DISPATCH(h, s->kernel.smooth = on != 0); return VFD_OK; }
int orc_tape(void* h, int on) {
    DISPATCH(h, {
        s->taping = on != 0;
        if (on) {
            s->tapes.clear();
            s->tape_error.clear();
        }
    });
    return VFD_OK;
}
// cotangents of the last taped step's output (lx_out, lv_out; 3n each) -> of the first taped step's input
// (lx_in, lv_in) and d/dmu; returns the number of reversed steps, or -1 with a message in err (256 bytes)
int orc_adjoint(void* h, const double* lx_out, const double* lv_out, double* lx_in, double* lv_in, double* lmu,
                char* err) {
    int steps = -1;
    DISPATCH(h, {
        if (!s->tape_error.empty()) {
            std::strncpy(err, s->tape_error.c_str(), 255);
            err[255] = 0;
        } else {
            const size_t n = s->p.size();
            using VT = typename std::remove_reference<decltype(s->p[0].x)>::type;
            using TT = decltype(VT{}.x);
            std::vector<VT> lx(n), lv(n);
            for (size_t i = 0; i < n; ++i) {
                lx[i] = VT{(TT)lx_out[3 * i], (TT)lx_out[3 * i + 1], (TT)lx_out[3 * i + 2]};
                lv[i] = VT{(TT)lv_out[3 * i], (TT)lv_out[3 * i + 1], (TT)lv_out[3 * i + 2]};
            }
            double m = 0;
            std::vector<TT> ldK(n, TT(0));
            for (long k = (long)s->tapes.size() - 1; k >= 0; --k) s->reverse_step(s->tapes[k], lx, lv, ldK, m);
            for (size_t i = 0; i < n; ++i) {
                lx_in[3 * i] = lx[i].x, lx_in[3 * i + 1] = lx[i].y, lx_in[3 * i + 2] = lx[i].z;
                lv_in[3 * i] = lv[i].x, lv_in[3 * i + 1] = lv[i].y, lv_in[3 * i + 2] = lv[i].z;
            }
            *lmu = m;
            steps = (int)s->tapes.size();
        }
    });
    return steps;
}
int orc_particle_count(void* h, uint32_t* n) { DISPATCH(h, *n = (uint32_t)s->p.size()); return VFD_OK; }
int orc_get_neighbors(void* h, uint32_t* counts, uint32_t* offsets, uint32_t* ids, uint64_t* total) {
    DISPATCH(h, {
        const NeighborLists& L = s->nb;
        if (total) *total = L.ids.size();
        if (counts) std::memcpy(counts, L.counts.data(), L.counts.size() * 4);
        if (offsets) std::memcpy(offsets, L.offsets.data(), L.offsets.size() * 4);
        if (ids) std::memcpy(ids, L.ids.data(), L.ids.size() * 4);
    });
    return VFD_OK;
}
int orc_get_debug_info(void* h, vfd_debug_info* out) {
    DISPATCH(h, {
        *out = s->dbg;
        out->TimeStepSize = (float)s->info.dt;
        out->MaxVelocityMagnitude = (float)s->maxVel;
    });
    return VFD_OK;
}
// info as doubles in vfd_info member order (numeric members only, gravity last)
int orc_get_info(void* h, double* out) {
    DISPATCH(h, {
        const auto& I = s->info;
        const double v[] = {(double)I.particleCount, (double)I.bodyCount, (double)I.R,  (double)I.R2,
                            (double)I.r,  (double)I.diam, (double)I.dt, (double)I.dt2,
                            (double)I.dtInv, (double)I.dt2Inv, (double)I.V,  (double)I.rho0,
                            (double)I.m,  (double)I.mInv, (double)I.nu, (double)I.nuB,
                            (double)I.mu, (double)I.muB, (double)I.tdf, (double)I.td,
                            (double)I.sigma, (double)I.S, (double)I.slope, (double)I.constant,
                            (double)I.temporal, (double)I.smoothing, (double)I.factor, (double)I.NR,
                            (double)I.mcFactor, (double)I.g.x, (double)I.g.y, (double)I.g.z};
        std::memcpy(out, v, sizeof(v));
    });
    return VFD_OK;
}
int orc_get_kernel_table(void* h, double* w, double* g, double* scalars) {
    DISPATCH(h, {
        const auto& K = s->kernel;
        for (unsigned i = 0; i < kTableResolution; ++i) w[i] = (double)K.W[i];
        for (unsigned i = 0; i <= kTableResolution; ++i) g[i] = (double)K.G[i];
        scalars[0] = K.R;
        scalars[1] = K.R2;
        scalars[2] = K.invStep;
        scalars[3] = K.W0;
        scalars[4] = K.K;
        scalars[5] = K.L;
    });
    return VFD_OK;
}
// kernel evaluation for tests: which = 0 W(vec), 1 W(scalar |r|), 2 gradW(vec)
int orc_kernel_eval(void* h, int which, const double* r, uint32_t n, double* out) {
    DISPATCH(h, {
        using TT = std::remove_reference<decltype(s->kernel.R)>::type;
        for (uint32_t i = 0; i < n; ++i) {
            const V3<TT> v = v3<TT>((TT)r[3 * i], (TT)r[3 * i + 1], (TT)r[3 * i + 2]);
            if (which == 0)
                out[3 * i] = (double)s->kernel.getW(v);
            else if (which == 1)
                out[3 * i] = (double)s->kernel.getW((TT)r[3 * i]);
            else {
                const V3<TT> g = s->kernel.gradW(v);
                out[3 * i] = g.x;
                out[3 * i + 1] = g.y;
                out[3 * i + 2] = g.z;
            }
        }
    });
    return VFD_OK;
}
int orc_get_boundary_map(void* h, uint32_t body, uint32_t field, double* nodes, uint32_t* count, double* geom) {
    DISPATCH(h, {
        if (body >= s->bodies.size() || field > 1) return VFD_ERROR_INVALID_ARGUMENT;
        const auto& M = s->bodies[body].map;
        if (count) *count = M.nodeCount;
        if (nodes)
            for (unsigned i = 0; i < M.nodeCount; ++i) nodes[i] = (double)M.nodes[field][i];
        if (geom) {
            const double gv[] = {(double)M.domain.min.x, (double)M.domain.min.y, (double)M.domain.min.z,
                                 (double)M.domain.max.x, (double)M.domain.max.y, (double)M.domain.max.z,
                                 (double)M.res[0], (double)M.res[1], (double)M.res[2]};
            std::memcpy(geom, gv, sizeof(gv));
        }
    });
    return VFD_OK;
}
// interpolate a body's map field at points (field 0 also returns the gradient)
int orc_map_interpolate(void* h, uint32_t body, uint32_t field, const double* pts, uint32_t n, double* out) {
    DISPATCH(h, {
        using TT = std::remove_reference<decltype(s->kernel.R)>::type;
        if (body >= s->bodies.size()) return VFD_ERROR_INVALID_ARGUMENT;
        const auto& M = s->bodies[body].map;
        for (uint32_t i = 0; i < n; ++i) {
            const V3<TT> x = v3<TT>((TT)pts[3 * i], (TT)pts[3 * i + 1], (TT)pts[3 * i + 2]);
            unsigned cell[32];
            V3<TT> c0, grad = v3<TT>(0, 0, 0);
            TT N[32];
            V3<TT> dN[32];
            double phi = (double)FLT_MAX;
            if (M.locate(x, cell, c0, N, dN))
                phi = field == 0 ? (double)M.interp_grad(0, cell, c0, N, dN, grad) : (double)M.interp(1, cell, N);
            out[4 * i] = phi;
            out[4 * i + 1] = grad.x;
            out[4 * i + 2] = grad.y;
            out[4 * i + 3] = grad.z;
        }
    });
    return VFD_OK;
}

int orc_viscosity_system(void* h, double* A, double* b) {
    DISPATCH(h, s->assemble_viscosity(A, b));
    return VFD_OK;
}

// ----- unit-level entry points ---------------------------------------------------------------------------------
void orc_halton(int fp64, double* out) {
    if (fp64) {
        auto h = halton_table<double>();
        for (unsigned i = 0; i < kHaltonMod; ++i) out[i] = h[i];
    } else {
        auto h = halton_table<float>();
        for (unsigned i = 0; i < kHaltonMod; ++i) out[i] = h[i];
    }
}
void orc_gauss_legendre(int n, double* x, double* w) {
    std::vector<double> gx, gw;
    gauss_legendre(n, gx, gw);
    for (int i = 0; i < n; ++i) {
        x[i] = gx[i];
        w[i] = gw[i];
    }
}
void orc_shape_function(const double* xi, double* N, double* dN) {
    double n[32];
    V3<double> d[32];
    shape_function(v3<double>(xi[0], xi[1], xi[2]), n, d);
    for (int k = 0; k < 32; ++k) {
        N[k] = n[k];
        dN[3 * k] = d[k].x;
        dN[3 * k + 1] = d[k].y;
        dN[3 * k + 2] = d[k].z;
    }
}
void orc_shape_function_hess(const double* xi, double* H) {
    double h[32][6];
    shape_function_hess(v3<double>(xi[0], xi[1], xi[2]), h);
    for (int k = 0; k < 32; ++k)
        for (int e = 0; e < 6; ++e) H[6 * k + e] = h[k][e];
}
uint32_t orc_morton3(uint32_t x, uint32_t y, uint32_t z) { return morton3(x, y, z); }
// cell node indices of a map of resolution res (SDF.cu:81-131)
void orc_cell_nodes(const uint32_t* res, uint32_t i, uint32_t j, uint32_t k, uint32_t* out) {
    VolumeMap<double> m;
    Box<double> b{v3<double>(0, 0, 0), v3<double>(1, 1, 1)};
    m.init(b, res);
    m.cell_nodes(i, j, k, out);
}
void orc_node_position(const double* dmin, const double* dmax, const uint32_t* res, uint32_t i, double* out) {
    VolumeMap<double> m;
    Box<double> b{v3<double>(dmin[0], dmin[1], dmin[2]), v3<double>(dmax[0], dmax[1], dmax[2])};
    m.init(b, res);
    const V3<double> p = m.node_position(i);
    out[0] = p.x;
    out[1] = p.y;
    out[2] = p.z;
}
uint32_t orc_node_count(const uint32_t* res) {
    VolumeMap<double> m;
    Box<double> b{v3<double>(0, 0, 0), v3<double>(1, 1, 1)};
    m.init(b, res);
    return m.nodeCount;
}
// O(N^2) neighbour sets (no grid) with the same distance expression; lists ascending by index, capped at cap.
int orc_brute_force_neighbors(const float* x, uint32_t n, float R, uint32_t cap, uint32_t* counts, uint32_t* ids) {
    const float R2 = R * R;
#pragma omp parallel for schedule(dynamic, 64)
    for (long il = 0; il < (long)n; ++il) {
        const uint32_t i = (uint32_t)il;
        uint32_t c = 0;
        for (uint32_t j = 0; j < n; ++j) {
            const float dx = x[3 * j] - x[3 * i], dy = x[3 * j + 1] - x[3 * i + 1], dz = x[3 * j + 2] - x[3 * i + 2];
            const float d2 = std::fma(dz, dz, std::fma(dy, dy, dx * dx));
            if (d2 < R2 && d2 > 0.0f) {
                if (c < cap) ids[(size_t)i * cap + c] = j;
                ++c;
            }
        }
        counts[i] = c;
    }
    return VFD_OK;
}
// grid geometry of the search for positions x (tests): floats gridMin[3], gridDelta[3]; uints dims[3], meta[3]
int orc_search_grid(const float* x, uint32_t n, float R, float* fout, uint32_t* uout) {
    std::vector<V3<float>> xs(n);
    for (uint32_t i = 0; i < n; ++i) xs[i] = v3<float>(x[3 * i], x[3 * i + 1], x[3 * i + 2]);
    const SearchGrid<float> g = make_grid(xs, R);
    fout[0] = g.gridMin.x;
    fout[1] = g.gridMin.y;
    fout[2] = g.gridMin.z;
    fout[3] = g.gridDelta.x;
    fout[4] = g.gridDelta.y;
    fout[5] = g.gridDelta.z;
    for (int a = 0; a < 3; ++a) {
        uout[a] = g.dims[a];
        uout[3 + a] = g.meta[a];
    }
    uout[6] = g.cellCount;
    return VFD_OK;
}

}  // extern "C"
