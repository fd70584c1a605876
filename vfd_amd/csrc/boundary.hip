// boundary.hip - SDF volume-map boundaries.
//
// Build (setup, once per body): RigidBody ctor (VFD/Source/Simulation/DFSPH/RigidBody/RigidBody.cu:10-73) builds a
// two-field cubic-Lagrange volume map (Utility/SDF/SDF.cu:45-139): field 0 = sign * (d(x) - (Padding - r)),
// field 1 = 0.8 * integral over the support ball of {1 if phi <= 0; W(phi)/W(0) if phi < R} by 16^3-point
// Gauss-Legendre quadrature (Core/Math/GaussQuadrature.cpp:5621-5655). The reference evaluates this on the host
// with OpenMP; here one GPU thread builds one node, keeping the reference's summation order (sequential in
// i, j, k) so that the nodes reproduce the CPU result.
//
// Step: ComputeVolumeAndBoundaryKernel (DFSPHKernels.cu:69-150) per particle and body: map lookup, boundary
// neighbour position x_b and volume V_b, push-out of particles inside the solid.
#include "device.h"
#include "kernels.h"

namespace vfd {

// ---- 32-node serendipity shape functions on [-1,1]^3 (SDFDeviceData.cuh:36-370) ---------------------------------
// Node order: 8 corners (bit0 x, bit1 y, bit2 z), x-edge nodes 8..15 (pair index iz + 2*iy), y-edge nodes 16..23
// (pair index ix + 2*iz), z-edge nodes 24..31 (pair index iy + 2*ix); in each pair the node at -1/3 comes first.
__device__ __forceinline__ void edge1d(float s, int side, float& f, float& df) {
    const float t = side == 0 ? 1.0f - 3.0f * s : 1.0f + 3.0f * s;
    const float ds = side == 0 ? -3.0f : 3.0f;
    f = (1.0f - s * s) * t;
    df = -2.0f * s * t + (1.0f - s * s) * ds;
}

template <bool kGrad>
__device__ __forceinline__ void shape_fn(f3 xi, float N[32], f3 dN[32]) {
    const float x = xi.x, y = xi.y, z = xi.z;
    const float x2 = x * x, y2 = y * y, z2 = z * z;
    const float lx[2] = {1.0f - x, 1.0f + x}, ly[2] = {1.0f - y, 1.0f + y}, lz[2] = {1.0f - z, 1.0f + z};
    const float sg[2] = {-1.0f, 1.0f};
    const float q = 9.0f * (x2 + y2 + z2) - 19.0f;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const int ix = c & 1, iy = (c >> 1) & 1, iz = (c >> 2) & 1;
        const float a = lx[ix], b = ly[iy], d = lz[iz];
        N[c] = 1.0f / 64.0f * q * a * b * d;
        if (kGrad) {
            dN[c].x = (sg[ix] * q + 18.0f * x * a) * b * d / 64.0f;
            dN[c].y = (sg[iy] * q + 18.0f * y * b) * a * d / 64.0f;
            dN[c].z = (sg[iz] * q + 18.0f * z * d) * a * b / 64.0f;
        }
    }
    const float k = 9.0f / 64.0f;
#pragma unroll
    for (int iy = 0; iy < 2; ++iy)
#pragma unroll
        for (int iz = 0; iz < 2; ++iz)
#pragma unroll
            for (int side = 0; side < 2; ++side) {
                float f, df;
                edge1d(x, side, f, df);
                const int n = 8 + 2 * (iz + 2 * iy) + side;
                N[n] = k * f * ly[iy] * lz[iz];
                if (kGrad) dN[n] = mk3(k * df * ly[iy] * lz[iz], k * f * sg[iy] * lz[iz], k * f * ly[iy] * sg[iz]);
            }
#pragma unroll
    for (int iz = 0; iz < 2; ++iz)
#pragma unroll
        for (int ix = 0; ix < 2; ++ix)
#pragma unroll
            for (int side = 0; side < 2; ++side) {
                float f, df;
                edge1d(y, side, f, df);
                const int n = 16 + 2 * (ix + 2 * iz) + side;
                N[n] = k * f * lx[ix] * lz[iz];
                if (kGrad) dN[n] = mk3(k * f * sg[ix] * lz[iz], k * df * lx[ix] * lz[iz], k * f * lx[ix] * sg[iz]);
            }
#pragma unroll
    for (int ix = 0; ix < 2; ++ix)
#pragma unroll
        for (int iy = 0; iy < 2; ++iy)
#pragma unroll
            for (int side = 0; side < 2; ++side) {
                float f, df;
                edge1d(z, side, f, df);
                const int n = 24 + 2 * (iy + 2 * ix) + side;
                N[n] = k * f * lx[ix] * ly[iy];
                if (kGrad) dN[n] = mk3(k * f * sg[ix] * ly[iy], k * f * lx[ix] * sg[iy], k * df * lx[ix] * ly[iy]);
            }
}

// 32 node indices of cell (i,j,k) in the node layout of SDF::AddFunction (SDF.cu:81-131).
__device__ __forceinline__ void cell_nodes(const MapDev& m, uint32_t i, uint32_t j, uint32_t k, uint32_t out[32]) {
    const uint32_t nx = m.res[0], ny = m.res[1], nz = m.res[2];
    const uint32_t nv = (nx + 1) * (ny + 1) * (nz + 1);
    const uint32_t nex = nx * (ny + 1) * (nz + 1), ney = (nx + 1) * ny * (nz + 1);
#pragma unroll
    for (uint32_t c = 0; c < 8; ++c)
        out[c] = (nx + 1) * (ny + 1) * (k + ((c >> 2) & 1u)) + (nx + 1) * (j + ((c >> 1) & 1u)) + (i + (c & 1u));
#pragma unroll
    for (uint32_t e = 0; e < 4; ++e) {
        const uint32_t b = nv + 2 * (nx * (ny + 1) * (k + (e & 1u)) + nx * (j + (e >> 1)) + i);
        out[8 + 2 * e] = b;
        out[9 + 2 * e] = b + 1;
    }
#pragma unroll
    for (uint32_t e = 0; e < 4; ++e) {
        const uint32_t b = nv + 2 * nex + 2 * (ny * (nz + 1) * (i + (e & 1u)) + ny * (k + (e >> 1)) + j);
        out[16 + 2 * e] = b;
        out[17 + 2 * e] = b + 1;
    }
#pragma unroll
    for (uint32_t e = 0; e < 4; ++e) {
        const uint32_t b = nv + 2 * nex + 2 * ney + 2 * (nz * (nx + 1) * (j + (e & 1u)) + nz * (i + (e >> 1)) + k);
        out[24 + 2 * e] = b;
        out[25 + 2 * e] = b + 1;
    }
}

// DetermineShapeFunction (SDFDeviceData.cuh:372-419): false outside the map domain.
template <bool kGrad>
__device__ __forceinline__ bool map_locate(const MapDev& m, f3 x, uint32_t cell[32], f3& c0, float N[32], f3 dN[32],
                                           f3* xiOut = nullptr) {
    if (!(m.domMin[0] <= x.x && m.domMin[1] <= x.y && m.domMin[2] <= x.z && m.domMax[0] >= x.x &&
          m.domMax[1] >= x.y && m.domMax[2] >= x.z))
        return false;
    uint32_t mi[3] = {(uint32_t)((x.x - m.domMin[0]) * m.cellSizeInv[0]),
                      (uint32_t)((x.y - m.domMin[1]) * m.cellSizeInv[1]),
                      (uint32_t)((x.z - m.domMin[2]) * m.cellSizeInv[2])};
#pragma unroll
    for (int a = 0; a < 3; ++a)
        if (mi[a] >= m.res[a]) mi[a] = m.res[a] - 1;
    const f3 origin = mk3(m.domMin[0] + (float)mi[0] * m.cellSize[0], m.domMin[1] + (float)mi[1] * m.cellSize[1],
                          m.domMin[2] + (float)mi[2] * m.cellSize[2]);
    const f3 top = origin + mk3(m.cellSize[0], m.cellSize[1], m.cellSize[2]);
    const f3 den = top - origin;
    c0 = mk3(2.0f / den.x, 2.0f / den.y, 2.0f / den.z);
    const f3 sum = top + origin;
    const f3 c1 = mk3(sum.x / den.x, sum.y / den.y, sum.z / den.z);
    const f3 xi = cmul(c0, x) - c1;
    cell_nodes(m, mi[0], mi[1], mi[2], cell);
    shape_fn<kGrad>(xi, N, dN);
    if (xiOut) *xiOut = xi;
    return true;
}

// sum_k c_k Hess(N_k)(xi) in the local coordinates, as (xx, yy, zz, xy, xz, yz), c_k = nodes[cell[k]]: the second
// derivatives of the interpolant, for the adjoint's derivative of the boundary normal (no reference counterpart).
// Corners N = q a b d / 64 (q = 9 |xi|^2 - 19, a = 1 + sx x, ...): N_xx = (18 a + 36 sx x) b d / 64,
// N_xy = (18 x a sy + 18 y sx b + q sx sy) d / 64. Edges N = (9/64) e(s) u v with e = (1 - s^2)(1 + 3 t s):
// e'' = -2 (1 + 3 t s) - 12 t s.
__device__ __forceinline__ void shape_hess_sum(f3 xi, const float* __restrict__ nodes, const uint32_t cell[32],
                                               float hs[6]) {
    const float x = xi.x, y = xi.y, z = xi.z;
    const float lx[2] = {1.0f - x, 1.0f + x}, ly[2] = {1.0f - y, 1.0f + y}, lz[2] = {1.0f - z, 1.0f + z};
    const float sg[2] = {-1.0f, 1.0f};
    const float q = 9.0f * (x * x + y * y + z * z) - 19.0f;
    const float c = 1.0f / 64.0f;
#pragma unroll
    for (int e = 0; e < 6; ++e) hs[e] = 0.0f;
    for (int n = 0; n < 8; ++n) {
        const int ix = n & 1, iy = (n >> 1) & 1, iz = (n >> 2) & 1;
        const float a = lx[ix], b = ly[iy], d = lz[iz], sx = sg[ix], sy = sg[iy], sz = sg[iz];
        const float cv = nodes[cell[n]] * c;
        hs[0] += cv * (18.0f * a + 36.0f * sx * x) * b * d;
        hs[1] += cv * (18.0f * b + 36.0f * sy * y) * a * d;
        hs[2] += cv * (18.0f * d + 36.0f * sz * z) * a * b;
        hs[3] += cv * (18.0f * x * a * sy + 18.0f * y * sx * b + q * sx * sy) * d;
        hs[4] += cv * (18.0f * x * a * sz + 18.0f * z * sx * d + q * sx * sz) * b;
        hs[5] += cv * (18.0f * y * b * sz + 18.0f * z * sy * d + q * sy * sz) * a;
    }
    const float k = 9.0f / 64.0f;
    auto edge = [](float s, int side, float& f, float& df, float& ddf) {
        const float t = side == 0 ? -1.0f : 1.0f;
        const float u = 1.0f + 3.0f * t * s;
        f = (1.0f - s * s) * u;
        df = -2.0f * s * u + 3.0f * t * (1.0f - s * s);
        ddf = -2.0f * u - 12.0f * t * s;
    };
    for (int iy = 0; iy < 2; ++iy)
        for (int iz = 0; iz < 2; ++iz)
            for (int side = 0; side < 2; ++side) {
                float f, df, ddf;
                edge(x, side, f, df, ddf);
                const float cv = nodes[cell[8 + 2 * (iz + 2 * iy) + side]] * k;
                hs[0] += cv * ddf * ly[iy] * lz[iz];
                hs[3] += cv * df * sg[iy] * lz[iz];
                hs[4] += cv * df * ly[iy] * sg[iz];
                hs[5] += cv * f * sg[iy] * sg[iz];
            }
    for (int iz = 0; iz < 2; ++iz)
        for (int ix = 0; ix < 2; ++ix)
            for (int side = 0; side < 2; ++side) {
                float f, df, ddf;
                edge(y, side, f, df, ddf);
                const float cv = nodes[cell[16 + 2 * (ix + 2 * iz) + side]] * k;
                hs[1] += cv * ddf * lx[ix] * lz[iz];
                hs[3] += cv * df * sg[ix] * lz[iz];
                hs[4] += cv * f * sg[ix] * sg[iz];
                hs[5] += cv * df * lx[ix] * sg[iz];
            }
    for (int ix = 0; ix < 2; ++ix)
        for (int iy = 0; iy < 2; ++iy)
            for (int side = 0; side < 2; ++side) {
                float f, df, ddf;
                edge(z, side, f, df, ddf);
                const float cv = nodes[cell[24 + 2 * (iy + 2 * ix) + side]] * k;
                hs[2] += cv * ddf * lx[ix] * ly[iy];
                hs[3] += cv * f * sg[ix] * sg[iy];
                hs[4] += cv * df * sg[ix] * ly[iy];
                hs[5] += cv * df * lx[ix] * sg[iy];
            }
}

__device__ __forceinline__ f3 node_position(const MapDev& m, uint32_t i) {
    const uint32_t nx = m.res[0], ny = m.res[1], nz = m.res[2];
    const uint32_t nv = (nx + 1) * (ny + 1) * (nz + 1);
    const uint32_t nex = nx * (ny + 1) * (nz + 1), ney = (nx + 1) * ny * (nz + 1);
    const f3 cs = mk3(m.cellSize[0], m.cellSize[1], m.cellSize[2]);
    const f3 mn = mk3(m.domMin[0], m.domMin[1], m.domMin[2]);
    f3 idx;
    f3 r;
    if (i < nv) {
        idx.z = (float)(i / ((ny + 1) * (nx + 1)));
        const uint32_t t = i % ((ny + 1) * (nx + 1));
        idx.y = (float)(t / (nx + 1));
        idx.x = (float)(t % (nx + 1));
        r = mn + cmul(cs, idx);
    } else if (i < nv + 2 * nex) {
        i -= nv;
        const uint32_t e = i / 2;
        idx.z = (float)(e / ((ny + 1) * nx));
        const uint32_t t = e % ((ny + 1) * nx);
        idx.y = (float)(t / nx);
        idx.x = (float)(t % nx);
        r = mn + cmul(cs, idx);
        r.x += (1.0f + (float)(i % 2)) / 3.0f * cs.x;
    } else if (i < nv + 2 * (nex + ney)) {
        i -= nv + 2 * nex;
        const uint32_t e = i / 2;
        idx.x = (float)(e / ((nz + 1) * ny));
        const uint32_t t = e % ((nz + 1) * ny);
        idx.z = (float)(t / ny);
        idx.y = (float)(t % ny);
        r = mn + cmul(cs, idx);
        r.y += (1.0f + (float)(i % 2)) / 3.0f * cs.y;
    } else {
        i -= nv + 2 * (nex + ney);
        const uint32_t e = i / 2;
        idx.y = (float)(e / ((nx + 1) * nz));
        const uint32_t t = e % ((nx + 1) * nz);
        idx.x = (float)(t / nz);
        idx.z = (float)(t % nz);
        r = mn + cmul(cs, idx);
        r.z += (1.0f + (float)(i % 2)) / 3.0f * cs.z;
    }
    return r;
}

// exact signed distance of the analytic shape; d < 0 inside
__device__ __forceinline__ float shape_distance(const ShapeDev& s, f3 p) {
    const f3 d = p - mk3(s.c[0], s.c[1], s.c[2]);
    switch (s.shape) {
        case 0: {  // box
            const f3 q = mk3(fabsf(d.x) - s.h[0], fabsf(d.y) - s.h[1], fabsf(d.z) - s.h[2]);
            const f3 qp = mk3(fmaxf(q.x, 0.0f), fmaxf(q.y, 0.0f), fmaxf(q.z, 0.0f));
            return sqrtf(dot3(qp, qp)) + fminf(fmaxf(q.x, fmaxf(q.y, q.z)), 0.0f);
        }
        case 1:  // sphere
            return sqrtf(dot3(d, d)) - s.radius;
        case 2:  // plane
            return dot3(d, mk3(s.n[0], s.n[1], s.n[2]));
        default: {  // cylinder, axis y
            const float rxz = sqrtf(fmaf(d.z, d.z, d.x * d.x)) - s.radius;
            const float ry = fabsf(d.y) - s.h[1];
            const float ox = fmaxf(rxz, 0.0f), oy = fmaxf(ry, 0.0f);
            return sqrtf(fmaf(oy, oy, ox * ox)) + fminf(fmaxf(rxz, ry), 0.0f);
        }
    }
}

__device__ __forceinline__ float map_interp_at0(const MapDev& m, const float* nodes0, f3 x) {
    uint32_t cell[32];
    f3 c0;
    float N[32];
    f3 dN[32];
    if (!map_locate<false>(m, x, cell, c0, N, dN)) return FLT_MAX;
    float phi = 0.0f;
#pragma unroll
    for (int k = 0; k < 32; ++k) phi += nodes0[cell[k]] * N[k];
    return phi;
}

__global__ void k_map_field0(ShapeDev shape, MapDev map, float* nodes0) {
    const uint32_t l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= map.nodeCount) return;
    const f3 x = node_position(map, l);
    nodes0[l] = shape.sign * (shape_distance(shape, x) - shape.tol);
}

__global__ void __launch_bounds__(64) k_map_field1(MapDev map, const float* __restrict__ nodes0, float* nodes1,
                                                   KTabHost kt, const float* __restrict__ glx,
                                                   const float* __restrict__ glw) {
    const uint32_t l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= map.nodeCount) return;
    const KTab k{kt.tab, kt.R, kt.R2, kt.invStep, kt.W0};
    const float R = kt.R;
    const f3 x = node_position(map, l);
    const float dX = map_interp_at0(map, nodes0, x);
    if (dX > 2.0f * R) {
        nodes1[l] = 0.0f;
        return;
    }
    const double c0 = (double)R;
    float res = 0.0f;
    for (int i = 0; i < 16; ++i) {
        const float wi = glw[i];
        for (int j = 0; j < 16; ++j) {
            const float wij = wi * glw[j];
            for (int kk = 0; kk < 16; ++kk) {
                const float wijk = wij * glw[kk];
                const f3 xi = mk3((float)(c0 * (double)glx[i]), (float)(c0 * (double)glx[j]), (float)(c0 * (double)glx[kk]));
                float f = 0.0f;
                if (!(dot3(xi, xi) > R * R)) {
                    const float dist = map_interp_at0(map, nodes0, x + xi);
                    if (dist <= 0.0f)
                        f = 1.0f;
                    else if (dist < R)
                        f = getW_scalar(k, dist) / k.W0;
                }
                res += wijk * f;
            }
        }
    }
    res = (float)((double)res * (c0 * c0 * c0));
    nodes1[l] = 0.8f * res;
}

void launch_map_field0(hipStream_t st, const ShapeDev& shape, const MapDev& map, float* nodes0) {
    hipLaunchKernelGGL(k_map_field0, dim3((map.nodeCount + 255) / 256), dim3(256), 0, st, shape, map, nodes0);
}

void launch_map_field1(hipStream_t st, const MapDev& map, const float* nodes0, float* nodes1, const KTabHost& kt,
                       const float* glx, const float* glw) {
    hipLaunchKernelGGL(k_map_field1, dim3((map.nodeCount + 63) / 64), dim3(64), 0, st, map, nodes0, nodes1, kt, glx,
                       glw);
}

// ---- triangle-mesh signed distance (MeshDistance::SignedDistance, Utility/SDF/MeshDistance.cpp:186-225) ----------
// The reference's host code evaluates glm::dot as (x x' + y y') + z z' without fused terms; so does this code.
__device__ __forceinline__ float gdot(f3 a, f3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }

// Squared distance from p to one triangle (PointTriangleDistanceSquared, MeshDistance.cpp:303-607): the minimiser
// (s, t) of |x0 + s e0 + t e1 - p|^2 over the triangle, found by classifying the unconstrained minimum into the seven
// regions around the triangle (Eberly, "Distance Between Point and Triangle in 3D", Geometric Tools), with the
// reference's branch order and float expressions. kWhere also returns the closest point and its feature:
// 0..2 = vertex x0, x1, x2; 3..5 = edge x0x1, x1x2, x2x0; 6 = interior.
template <bool kWhere>
__device__ __forceinline__ float tri_d2(float4 r0, float4 r1, float4 r2, float4 r3, f3 p, f3* cp, int* feature) {
    const f3 x0 = mk3(r0.x, r0.y, r0.z), e0 = mk3(r1.x, r1.y, r1.z), e1 = mk3(r2.x, r2.y, r2.z);
    const float a00 = r0.w, a01 = r1.w, a11 = r2.w, det = r3.x, denom = r3.y, invDet = r3.z;
    const f3 diff = x0 - p;
    const float b0 = gdot(diff, e0), b1 = gdot(diff, e1), c = gdot(diff, diff);
    float s = a01 * b1 - a11 * b0;
    float t = a01 * b0 - a00 * b1;
    float d2;
    int f;
    // the five boundary outcomes shared by several regions
    auto at_x0 = [&] { s = 0.0f; t = 0.0f; d2 = c; f = 0; };
    auto at_x1 = [&] { s = 1.0f; t = 0.0f; d2 = (a00 + 2.0f * b0) + c; f = 1; };
    auto at_x2 = [&] { s = 0.0f; t = 1.0f; d2 = (a11 + 2.0f * b1) + c; f = 2; };
    auto on_e0 = [&] { t = 0.0f; s = -b0 / a00; d2 = b0 * s + c; f = 3; };
    auto on_e2 = [&] { s = 0.0f; t = -b1 / a11; d2 = b1 * t + c; f = 5; };
    auto quadratic = [&] { return (s * ((a00 * s + a01 * t) + 2.0f * b0) + t * ((a01 * s + a11 * t) + 2.0f * b1)) + c; };
    // edge x1x2 (s + t = 1) at s = u
    auto on_e1 = [&](float u, bool uIsS) {
        if (uIsS) { s = u; t = 1.0f - s; } else { t = u; s = 1.0f - t; }
        d2 = quadratic();
        f = 4;
    };
    // s = 0 edge, vertices x0/x2 (regions 3 and 4 with b0 >= 0)
    auto side_s0 = [&] {
        if (b1 >= 0.0f) at_x0();
        else if (-b1 >= a11) at_x2();
        else on_e2();
    };
    if (s + t <= det) {
        if (s < 0.0f) {
            if (t < 0.0f) {  // region 4
                if (b0 < 0.0f) {
                    if (-b0 >= a00) at_x1();
                    else on_e0();
                } else {
                    side_s0();
                }
            } else {  // region 3
                side_s0();
            }
        } else if (t < 0.0f) {  // region 5
            if (b0 >= 0.0f) at_x0();
            else if (-b0 >= a00) at_x1();
            else on_e0();
        } else {  // region 0: interior (invDet = 1 / det, the same float the reference divides out per query)
            s = s * invDet;
            t = t * invDet;
            d2 = quadratic();
            f = 6;
        }
    } else if (s < 0.0f) {  // region 2
        const float tmp0 = a01 + b0, tmp1 = a11 + b1;
        if (tmp1 > tmp0) {
            const float numer = tmp1 - tmp0;
            if (numer >= denom) at_x1();
            else on_e1(numer / denom, true);
        } else if (tmp1 <= 0.0f) at_x2();
        else if (b1 >= 0.0f) at_x0();
        else on_e2();
    } else if (t < 0.0f) {  // region 6
        const float tmp0 = a01 + b1, tmp1 = a00 + b0;
        if (tmp1 > tmp0) {
            const float numer = tmp1 - tmp0;
            if (numer >= denom) at_x2();
            else on_e1(numer / denom, false);
        } else if (tmp1 <= 0.0f) at_x1();
        else if (b0 >= 0.0f) at_x0();
        else on_e0();
    } else {  // region 1
        const float numer = ((a11 + b1) - a01) - b0;
        if (numer <= 0.0f) at_x2();
        else if (numer >= denom) at_x1();
        else on_e1(numer / denom, true);
    }
    if (d2 < 0.0f) d2 = 0.0f;
    if (kWhere) {
        *cp = (x0 + e0 * s) + e1 * t;
        *feature = f;
    }
    return d2;
}

constexpr uint32_t kMeshTile = 256;
static_assert(kMeshTile % kMeshCluster == 0 && kMeshCluster == 64, "a wave decides one cluster at a time");

// The map's nodes as four grids (SDF.cu:45-139 numbering): vertices (nx+1, ny+1, nz+1), x-edge nodes (2nx, ny+1,
// nz+1), y-edge nodes (2ny, nz+1, nx+1) and z-edge nodes (2nz, nx+1, ny+1), each linear with its first dimension
// fastest. The distance kernel gives every wave one 4 x 4 x 4 brick of one grid, so that a wave's nodes are close
// together and its cluster test can skip most of a large mesh.
struct NodeBricks {
    uint32_t dim[4][3], nb[4][3], first[5], base[4];
    __host__ __device__ explicit NodeBricks(const MapDev& m) {
        const uint32_t nx = m.res[0], ny = m.res[1], nz = m.res[2];
        const uint32_t d[4][3] = {{nx + 1, ny + 1, nz + 1}, {2 * nx, ny + 1, nz + 1}, {2 * ny, nz + 1, nx + 1},
                                  {2 * nz, nx + 1, ny + 1}};
        first[0] = 0;
        base[0] = 0;
        for (int c = 0; c < 4; ++c) {
            for (int a = 0; a < 3; ++a) {
                dim[c][a] = d[c][a];
                nb[c][a] = (d[c][a] + 3) / 4;
            }
            first[c + 1] = first[c] + nb[c][0] * nb[c][1] * nb[c][2];
            if (c < 3) base[c + 1] = base[c] + d[c][0] * d[c][1] * d[c][2];
        }
    }
    __host__ __device__ uint32_t waves() const { return first[4]; }
    // node of lane `lane` of brick-wave w, or ~0u past the grid's edge
    __device__ uint32_t node(uint32_t w, uint32_t lane) const {
        if (w >= first[4]) return ~0u;
        int c = 0;
        while (w >= first[c + 1]) ++c;
        const uint32_t b = w - first[c];
        const uint32_t i0 = (b % nb[c][0]) * 4 + (lane & 3u);
        const uint32_t i1 = ((b / nb[c][0]) % nb[c][1]) * 4 + ((lane >> 2) & 3u);
        const uint32_t i2 = (b / (nb[c][0] * nb[c][1])) * 4 + (lane >> 4);
        if (i0 >= dim[c][0] || i1 >= dim[c][1] || i2 >= dim[c][2]) return ~0u;
        return base[c] + (i2 * dim[c][1] + i1) * dim[c][0] + i0;
    }
};

// Field 0 of a mesh body: one thread per map node, a wave per brick of nodes (NodeBricks). A block stages
// kMeshTile triangles at a time in LDS (read as broadcasts); each wave skips a cluster of kMeshCluster triangles when
// every lane's node is provably farther from the cluster's bounding sphere than from a triangle already seen (or than
// the farthest point of some cluster, the bound before any triangle is seen). The winner is the least squared distance, ties to the smaller mesh index - the
// first minimum of the reference's order - so the skipped clusters cannot change it: the test keeps a margin of
// 4e-5 (|p - c| + r)^2 over the rounding of the squared distances, which stays below 1e-6 of that scale. The
// winner's closest point and feature give the sign: negative where (p - closest) points against the feature's normal.
__global__ void __launch_bounds__(kMeshTile) k_mesh_field0(MeshDev mesh, MapDev map, float sign, float tol,
                                                          float* nodes0) {
    __shared__ float4 sTri[4 * kMeshTile];
    const NodeBricks bricks(map);
    const uint32_t t = blockIdx.x * kMeshTile + threadIdx.x;
    const uint32_t l = bricks.node(t / 64u, t % 64u);
    const bool live = l < map.nodeCount;
    const f3 p = live ? node_position(map, l) : mk3(0.0f, 0.0f, 0.0f);
    const uint32_t clusters = (mesh.triCount + kMeshCluster - 1) / kMeshCluster;
    // |p - c| and the margin of one cluster
    auto reach = [&](uint32_t c, float& dc, float& r) {
        const float4 sp = mesh.spheres[c];
        const f3 d = p - mk3(sp.x, sp.y, sp.z);
        dc = sqrtf(gdot(d, d));
        r = sp.w;
    };
    float bound = INFINITY;  // no triangle is farther than (|p - c| + r)^2 from p, for every cluster c
    for (uint32_t c = 0; c < clusters; ++c) {
        float dc, r;
        reach(c, dc, r);
        const float u = dc + r;
        bound = fminf(bound, u * u);
    }
    float best = INFINITY;
    uint32_t bestIdx = 0xffffffffu, bestPos = mesh.firstPos;
    for (uint32_t base = 0; base < mesh.triCount; base += kMeshTile) {
        __syncthreads();
        const uint32_t tl = base + threadIdx.x;
        if (tl < mesh.triCount)
            for (int q = 0; q < 4; ++q) sTri[4 * threadIdx.x + q] = mesh.tri[4 * (size_t)tl + q];
        __syncthreads();
        for (uint32_t cb = 0; cb < kMeshTile && base + cb < mesh.triCount; cb += kMeshCluster) {
            float dc, r;
            reach((base + cb) / kMeshCluster, dc, r);
            const float lb = dc - r, u = dc + r;
            const float lim = fminf(best, bound) + 4e-5f * (u * u);
            const bool skip = !live || (lb > 0.0f && lb * lb > lim);
            if (__all(skip)) continue;  // wave-uniform
            const uint32_t n = min(kMeshCluster, mesh.triCount - (base + cb));
            for (uint32_t k = cb; k < cb + n; ++k) {
                const float4 r3 = sTri[4 * k + 3];
                const float d2 = tri_d2<false>(sTri[4 * k], sTri[4 * k + 1], sTri[4 * k + 2], r3, p, nullptr, nullptr);
                const uint32_t idx = __float_as_uint(r3.w);
                if (d2 < best || (d2 == best && idx < bestIdx)) {
                    best = d2;
                    bestIdx = idx;
                    bestPos = base + k;
                }
            }
        }
    }
    if (!live) return;
    const float4* rp = mesh.tri + 4 * (size_t)bestPos;
    f3 cp;
    int feature;
    float d = sqrtf(tri_d2<true>(rp[0], rp[1], rp[2], rp[3], p, &cp, &feature));
    const float* nrm = mesh.normals + 21 * (size_t)bestPos + 3 * feature;
    if (gdot(p - cp, mk3(nrm[0], nrm[1], nrm[2])) < 0.0f) d = -d;
    nodes0[l] = sign * (d - tol);
}

void launch_mesh_field0(hipStream_t st, const MeshDev& mesh, const MapDev& map, float sign, float tol, float* nodes0) {
    const uint32_t waves = NodeBricks(map).waves();
    hipLaunchKernelGGL(k_mesh_field0, dim3((waves * 64u + kMeshTile - 1) / kMeshTile), dim3(kMeshTile), 0, st, mesh,
                       map, sign, tol, nodes0);
}

// ---- ParticleSampler::SampleMeshVolume (Utility/Sampler/ParticleSampler.cpp:7-91) ---------------------------------
// Candidate (i, j, k) of the x, y, z loops; i and j are the reference's counterX and counterY.
__device__ __forceinline__ f3 sample_point(const SampleGrid& g, uint32_t i, uint32_t j, uint32_t k) {
    const float x = g.xs[i], y = g.ys[j], z = g.zs[k], r = g.radius;
    if (g.mode == 0) return mk3(x + r, y + r, z + r);
    if (g.mode == 1) return (j % 2 == 0) ? mk3(x, y + r, z + r) : mk3(x + r, y + r, z);
    f3 shift = mk3(0.0f, 0.0f, 0.0f);
    if (i % 2) shift.z = shift.z + g.diameter / (2.0f * ((j % 2) ? -1.0f : 1.0f));
    if (j % 2) {
        shift.x = shift.x + g.shiftX / 2.0f;
        shift.z = shift.z + g.diameter / 2.0f;
    }
    return mk3(x, y + r, z + r) + shift;
}

__global__ void k_sample_flags(SampleGrid g, MapDev map, const float* __restrict__ nodes0, uint32_t* flags) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t nxy = g.nx * g.ny;
    if (c >= nxy * g.nz) return;
    const uint32_t k = c / nxy, rem = c % nxy;
    const f3 p = sample_point(g, rem % g.nx, rem / g.nx, k);
    flags[c] = map_interp_at0(map, nodes0, p) < 0.0f ? 1u : 0u;  // SDF::GetDistance(p, 0) < 0 (SDF.cu:217-225)
}

__global__ void k_sample_scatter(SampleGrid g, const uint32_t* __restrict__ flags,
                                 const uint32_t* __restrict__ offsets, float* out) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t nxy = g.nx * g.ny;
    if (c >= nxy * g.nz || !flags[c]) return;
    const uint32_t k = c / nxy, rem = c % nxy;
    const f3 p = sample_point(g, rem % g.nx, rem / g.nx, k);
    float* o = out + 3 * (size_t)offsets[c];
    o[0] = p.x;
    o[1] = p.y;
    o[2] = p.z;
}

void launch_sample_flags(hipStream_t st, const SampleGrid& g, const MapDev& map, const float* nodes0, uint32_t* flags) {
    const uint32_t n = g.nx * g.ny * g.nz;
    hipLaunchKernelGGL(k_sample_flags, dim3((n + 255) / 256), dim3(256), 0, st, g, map, nodes0, flags);
}

void launch_sample_scatter(hipStream_t st, const SampleGrid& g, const uint32_t* flags, const uint32_t* offsets,
                           float* out) {
    const uint32_t n = g.nx * g.ny * g.nz;
    hipLaunchKernelGGL(k_sample_scatter, dim3((n + 255) / 256), dim3(256), 0, st, g, flags, offsets, out);
}

// ComputeVolumeAndBoundaryKernel (DFSPHKernels.cu:69-150).
__global__ void __launch_bounds__(kTile) k_boundary(const SceneParams* __restrict__ P,
                                                    const uint32_t* __restrict__ tileScene,
                                                    const MapDev* __restrict__ maps, float4* pos, float4* vel,
                                                    float4* xbOut, float* vbOut, uint32_t capacity) {
    const TileCtx c = tile_ctx(tileScene, P);
    if (!c.valid) return;
    const SceneParams& p = P[c.s];
    f3 x = ld3(pos[c.i]);
    bool moved = false;
    for (uint32_t b = 0; b < p.bodyCount; ++b) {
        const MapDev& m = maps[c.s * kMaxBodies + b];
        f3 xb = mk3(0.0f, 0.0f, 0.0f);
        float vb = 0.0f;
        f3 normal = mk3(0.0f, 0.0f, 0.0f);
        float dist = FLT_MAX;
        uint32_t cell[32];
        f3 c0;
        float N[32];
        f3 dN[32];
        if (map_locate<true>(m, x, cell, c0, N, dN)) {
            float phi = 0.0f;
            f3 g = mk3(0.0f, 0.0f, 0.0f);
#pragma unroll
            for (int k = 0; k < 32; ++k) {
                const float cv = m.nodes0[cell[k]];
                phi += cv * N[k];
                g = g + cv * dN[k];
            }
            normal = cmul(g, c0);
            dist = phi;
        }
        if (dist > 0.0f && dist < p.R) {
            float vol = 0.0f;
#pragma unroll
            for (int k = 0; k < 32; ++k) vol += m.nodes1[cell[k]] * N[k];
            if (vol > 0.0f && vol != FLT_MAX) {
                vb = vol;
                const float nl = length(normal);
                if (nl > 1.0e-9f) {
                    normal = normal / nl;
                    const float d = fmaxf(dist + 0.5f * p.r, p.diam);
                    xb = x - d * normal;
                } else {
                    vb = 0.0f;
                }
            }
        } else if (dist <= 0.0f) {
            const float nl = length(normal);
            if (nl > 1.0e-5f) {
                normal = normal / nl;
                const float delta = fminf(p.diam - dist, 0.1f * p.r);
                x = x + delta * normal;
                moved = true;
            }
        }
        xbOut[(size_t)b * capacity + c.i] = st4(xb);
        vbOut[(size_t)b * capacity + c.i] = vb;
    }
    if (moved) {
        pos[c.i] = st4(x);
        vel[c.i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
}

void launch_boundary(const Launch& L, const MapDev* maps, float4* pos, float4* vel, float4* xb, float* vb) {
    hipLaunchKernelGGL(k_boundary, dim3(L.tiles), dim3(kTile), 0, L.stream, L.P, L.tileScene, maps, pos, vel, xb, vb,
                       L.capacity);
}

// (I - n n^T) Hs / |g|: the derivative of the unit normal n = g / |g|, Hs = Hess(phi) as (xx, yy, zz, xy, xz, yz)
__device__ __forceinline__ void normal_jacobian(f3 n, float gl, const float hs[6], float J[9]) {
    const float Hm[9] = {hs[0], hs[3], hs[4], hs[3], hs[1], hs[5], hs[4], hs[5], hs[2]};
    const float nv[3] = {n.x, n.y, n.z};
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            float s = Hm[3 * r + c];
#pragma unroll
            for (int k = 0; k < 3; ++k) s -= nv[r] * nv[k] * Hm[3 * k + c];
            J[3 * r + c] = s / gl;
        }
}

// Adjoint recording of the boundary pass of one scene (no reference counterpart): replays k_boundary from the
// positions it started from, with the same arithmetic and branches, and stores per particle and body the contact
// sample's derivatives (con[(kContactRecs b + q) n + i]: q = 0 {r = x - xb, vb}, 1 {grad vb, 0}, 2..4 the rows of
// dr/dx, with r = s n, s = max(phi + r/2, 2r), 5..7 the rows of dn/dx) and per particle the push-out's Jacobian (push[q n + i], q = 0..2 rows; row 0's w
// = 1 for a pushed particle). flags[0] / flags[1]: some particle has a contact sample / was pushed; flags[2]: a
// particle was pushed twice or after a contact sample, or the replay's position differs from the forward's.
__global__ void __launch_bounds__(kTile) k_adj_boundary_tape(const SceneParams* __restrict__ P,
                                                             const uint32_t* __restrict__ tileScene,
                                                             const MapDev* __restrict__ mapsAll, uint32_t n,
                                                             uint32_t capacity, const float4* __restrict__ xpre,
                                                             const float4* __restrict__ posAfter,
                                                             const float* __restrict__ vbF,
                                                             const float4* __restrict__ xbF, float4* con,
                                                             float4* push, uint32_t* flagsAll) {
    const uint32_t i = blockIdx.x * kTile + threadIdx.x;
    if (i >= n) return;
    const uint32_t scene = tileScene[blockIdx.x];
    const SceneParams& p = P[scene];
    if (i >= p.base + p.count) return;  // a slot past the scene's last particle
    const MapDev* maps = mapsAll + (size_t)scene * kMaxBodies;
    uint32_t* flags = flagsAll + 3 * (size_t)scene;
    f3 x = ld3(xpre[i]);
    bool sampled = false, pushed = false, bad = false;
    float Jp[9] = {1.0f, 0.0f, 0.0f, 0.0f, 1.0f, 0.0f, 0.0f, 0.0f, 1.0f};
    for (uint32_t b = 0; b < p.bodyCount; ++b) {
        const MapDev& m = maps[b];
        float4* cr = con + (size_t)kContactRecs * b * n + i;
        cr[0] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        f3 normal = mk3(0.0f, 0.0f, 0.0f);
        float dist = FLT_MAX;
        uint32_t cell[32];
        f3 c0, xi;
        float N[32];
        f3 dN[32];
        if (map_locate<true>(m, x, cell, c0, N, dN, &xi)) {
            float phi = 0.0f;
            f3 g = mk3(0.0f, 0.0f, 0.0f);
#pragma unroll
            for (int k = 0; k < 32; ++k) {
                const float cv = m.nodes0[cell[k]];
                phi += cv * N[k];
                g = g + cv * dN[k];
            }
            normal = cmul(g, c0);
            dist = phi;
        } else {
            continue;
        }
        const float vb = vbF[(size_t)b * capacity + i];
        const bool contact = vb > 0.0f;
        const bool push = !contact && dist <= 0.0f && length(normal) > 1.0e-5f;
        if (!contact && !push) continue;
        float hs[6];
        shape_hess_sum(xi, m.nodes0, cell, hs);
        const float sc[6] = {c0.x * c0.x, c0.y * c0.y, c0.z * c0.z, c0.x * c0.y, c0.x * c0.z, c0.y * c0.z};
#pragma unroll
        for (int e = 0; e < 6; ++e) hs[e] *= sc[e];
        const float nl = length(normal);
        const f3 nn = normal / nl;
        float Jn[9];
        normal_jacobian(nn, nl, hs, Jn);
        const float nv[3] = {nn.x, nn.y, nn.z};
        if (contact) {
            const float s0 = dist + 0.5f * p.r;
            const float s = fmaxf(s0, p.diam);
            const f3 ds = s0 < p.diam ? mk3(0.0f, 0.0f, 0.0f) : normal;
            f3 dv = mk3(0.0f, 0.0f, 0.0f);
#pragma unroll
            for (int k = 0; k < 32; ++k) dv = dv + m.nodes1[cell[k]] * dN[k];
            dv = cmul(dv, c0);
            const f3 r = x - ld3(xbF[(size_t)b * capacity + i]);
            const float dsv[3] = {ds.x, ds.y, ds.z};
            cr[0] = make_float4(r.x, r.y, r.z, vb);
            cr[n] = make_float4(dv.x, dv.y, dv.z, 0.0f);
#pragma unroll
            for (int q = 0; q < 3; ++q)
                cr[(size_t)(2 + q) * n] = make_float4(nv[q] * dsv[0] + s * Jn[3 * q], nv[q] * dsv[1] + s * Jn[3 * q + 1],
                                                      nv[q] * dsv[2] + s * Jn[3 * q + 2], 0.0f);
#pragma unroll
            for (int q = 0; q < 3; ++q)
                cr[(size_t)(5 + q) * n] = make_float4(Jn[3 * q], Jn[3 * q + 1], Jn[3 * q + 2], 0.0f);
            sampled = true;
        } else {
            if (sampled || pushed) bad = true;
            const float d0 = p.diam - dist;
            const float delta = fminf(d0, 0.1f * p.r);
            const f3 dd = 0.1f * p.r < d0 ? mk3(0.0f, 0.0f, 0.0f) : -1.0f * normal;
            const float ddv[3] = {dd.x, dd.y, dd.z};
#pragma unroll
            for (int r = 0; r < 3; ++r)
#pragma unroll
                for (int k = 0; k < 3; ++k)
                    Jp[3 * r + k] = (r == k ? 1.0f : 0.0f) + nv[r] * ddv[k] + delta * Jn[3 * r + k];
            x = x + delta * nn;
            pushed = true;
        }
    }
    const f3 xa = ld3(posAfter[i]);
    if (x.x != xa.x || x.y != xa.y || x.z != xa.z) bad = true;
    push[i] = make_float4(Jp[0], Jp[1], Jp[2], pushed ? 1.0f : 0.0f);
    push[n + i] = make_float4(Jp[3], Jp[4], Jp[5], 0.0f);
    push[2 * (size_t)n + i] = make_float4(Jp[6], Jp[7], Jp[8], 0.0f);
    if (sampled) flags[0] = 1u;
    if (pushed) flags[1] = 1u;
    if (bad) flags[2] = 1u;
}

void adj_launch_boundary_tape(hipStream_t s, const SceneParams* P, const uint32_t* tileScene, const MapDev* maps,
                              uint32_t n, uint32_t capacity, const float4* xpre, const float4* posAfter, const float* vb,
                              const float4* xb, float4* con, float4* push, uint32_t* flags) {
    if (n)
        hipLaunchKernelGGL(k_adj_boundary_tape, dim3((n + kTile - 1) / kTile), dim3(kTile), 0, s, P, tileScene, maps,
                           n, capacity, xpre, posAfter, vb, xb, con, push, flags);
}

}  // namespace vfd
