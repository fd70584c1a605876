// mesh.cpp - triangle-mesh inputs: the host preparation of a boundary mesh for the distance kernel and the particle
// sampler (FluidObject ctor + ParticleSampler::SampleMeshVolume).
//
// The reference transforms the vertices (RigidBody.cu:18-22, FluidObject.cpp:9-14), builds a half-edge mesh
// (Renderer/Mesh/EdgeMesh.cpp:132-166) and precomputes face and angle-weighted vertex normals
// (Utility/SDF/MeshDistance.cpp:13-57). The same float arithmetic runs here, on the host, once per mesh; the
// per-node distance search runs on the GPU (boundary.hip: k_mesh_field0). Compiled with -ffp-contract=off, so every
// expression below rounds as written.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/vfd.h"
#include "kernels.h"
#include "mesh.h"

namespace vfd {

int set_error(int code, const std::string& msg);  // sim.cpp: vfd_last_error's message

namespace {

struct V {
    float x, y, z;
    float operator[](int a) const { return a == 0 ? x : (a == 1 ? y : z); }
};
V operator+(V a, V b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
V operator-(V a, V b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
V operator-(V a) { return {-a.x, -a.y, -a.z}; }
V operator*(float s, V a) { return {s * a.x, s * a.y, s * a.z}; }
// glm::dot, glm::cross, glm::normalize (v * inversesqrt(dot(v, v)), inversesqrt(x) = 1 / sqrt(x))
float gdot(V a, V b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
V gcross(V a, V b) { return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y}; }
V gnormalize(V a) { return (1.0f / std::sqrt(gdot(a, a))) * a; }

// glm::mat4 * vec4(v, 1): (m0 x + m1 y) + (m2 z + m3 w), column by column
V transform_point(const float* m, V v) {
    V r;
    float* o = &r.x;
    for (int i = 0; i < 3; ++i) o[i] = (m[i] * v.x + m[4 + i] * v.y) + (m[8 + i] * v.z + m[12 + i] * 1.0f);
    return r;
}

}  // namespace

int prepare_mesh(const vfd_mesh& mesh, const float* transform, MeshPrep& out) {
    if (!mesh.Vertices || !mesh.Triangles || mesh.VertexCount == 0 || mesh.TriangleCount == 0)
        return set_error(VFD_ERROR_INVALID_ARGUMENT, "a mesh needs vertices and at least one triangle");
    if (mesh.TriangleCount > (1u << 26))
        return set_error(VFD_ERROR_INVALID_ARGUMENT, "at most 2^26 triangles per mesh");
    const uint32_t nv = mesh.VertexCount, nt = mesh.TriangleCount;
    for (size_t k = 0; k < 3 * (size_t)nt; ++k)
        if (mesh.Triangles[k] >= nv)
            return set_error(VFD_ERROR_INVALID_ARGUMENT, "triangle " + std::to_string(k / 3) +
                                                             " names vertex " + std::to_string(mesh.Triangles[k]) +
                                                             " of " + std::to_string(nv));
    std::vector<V> x(nv);
    for (uint32_t i = 0; i < nv; ++i) {
        const V v{mesh.Vertices[3 * (size_t)i], mesh.Vertices[3 * (size_t)i + 1], mesh.Vertices[3 * (size_t)i + 2]};
        x[i] = transform ? transform_point(transform, v) : v;
    }
    // BoundingBox(vertices) (Core/Structures/BoundingBox.h:16-31) starts from the zero box and extends it with
    // glm::min / glm::max, so the box always holds the origin.
    float lo[3] = {0.0f, 0.0f, 0.0f}, hi[3] = {0.0f, 0.0f, 0.0f};
    for (const V& v : x) {
        const float c[3] = {v.x, v.y, v.z};
        for (int a = 0; a < 3; ++a) {
            lo[a] = c[a] < lo[a] ? c[a] : lo[a];
            hi[a] = hi[a] < c[a] ? c[a] : hi[a];
        }
    }
    std::memcpy(out.lo, lo, sizeof(lo));
    std::memcpy(out.hi, hi, sizeof(hi));

    // face normals and angle-weighted vertex normals, summed in face order (MeshDistance.cpp:28-56)
    const uint32_t* F = mesh.Triangles;
    std::vector<V> fn(nt), vn(nv, V{0.0f, 0.0f, 0.0f});
    for (uint32_t f = 0; f < nt; ++f) {
        const V x0 = x[F[3 * f]], x1 = x[F[3 * f + 1]], x2 = x[F[3 * f + 2]];
        const V n = gnormalize(gcross(x1 - x0, x2 - x0));
        const V e1 = gnormalize(x1 - x0), e2 = gnormalize(x2 - x1), e3 = gnormalize(x0 - x2);
        const float alpha[3] = {std::acos(gdot(e1, -e3)), std::acos(gdot(e2, -e1)), std::acos(gdot(e3, -e2))};
        for (int k = 0; k < 3; ++k) vn[F[3 * f + k]] = vn[F[3 * f + k]] + alpha[k] * n;
        fn[f] = n;
    }
    // opposite half-edges (EdgeMesh.cpp:138-165): half-edge k of face f runs from corner k to corner k+1 and pairs
    // with a pending half-edge running the other way; half-edges left unpaired are border edges. On a manifold mesh
    // the pairing is unique; at an edge shared by more than two faces the earliest pending one is taken.
    std::vector<int64_t> opposite(3 * (size_t)nt, -1);
    std::unordered_map<uint64_t, std::vector<uint32_t>> pending;
    pending.reserve(3 * (size_t)nt / 2 + 1);
    for (uint32_t f = 0; f < nt; ++f)
        for (uint32_t k = 0; k < 3; ++k) {
            const uint32_t a = F[3 * f + k], b = F[3 * f + (k + 1) % 3];
            const uint32_t he = 3 * f + k;
            auto it = pending.find(((uint64_t)b << 32) | a);
            if (it != pending.end() && !it->second.empty()) {
                const uint32_t o = it->second.front();
                it->second.erase(it->second.begin());
                opposite[he] = o;
                opposite[o] = he;
            } else {
                pending[((uint64_t)a << 32) | b].push_back(he);
            }
        }

    // storage order: Morton code of the centroid over the bounding box (10 bits per axis), then the mesh index, so
    // that each cluster of kMeshCluster stored triangles is compact in space
    std::vector<uint64_t> key(nt);
    for (uint32_t f = 0; f < nt; ++f) {
        uint64_t code = 0;
        for (int a = 0; a < 3; ++a) {
            const double c = ((double)x[F[3 * f]][a] + x[F[3 * f + 1]][a] + x[F[3 * f + 2]][a]) / 3.0;
            const double ext = (double)hi[a] - lo[a];
            double u = ext > 0.0 ? (c - lo[a]) / ext : 0.0;
            u = std::isfinite(u) ? std::min(std::max(u, 0.0), 1.0) : 0.0;
            const uint64_t q = (uint64_t)(u * 1023.0);
            for (int b = 0; b < 10; ++b) code |= ((q >> b) & 1u) << (3 * b + a);
        }
        key[f] = (code << 32) | f;
    }
    std::sort(key.begin(), key.end());
    out.tri.assign(16 * (size_t)nt, 0.0f);
    out.normals.assign(21 * (size_t)nt, 0.0f);
    const uint32_t clusters = (nt + kMeshCluster - 1) / kMeshCluster;
    out.spheres.assign(4 * (size_t)clusters, 0.0f);
    for (uint32_t c = 0; c < clusters; ++c) {
        // bounding sphere of the cluster's vertices: centre of their box, radius rounded up so that every vertex
        // (hence every triangle) lies inside it in exact arithmetic
        double blo[3] = {INFINITY, INFINITY, INFINITY}, bhi[3] = {-INFINITY, -INFINITY, -INFINITY};
        const uint32_t end = std::min(nt, (c + 1) * kMeshCluster);
        for (uint32_t k = c * kMeshCluster; k < end; ++k)
            for (int v = 0; v < 3; ++v)
                for (int a = 0; a < 3; ++a) {
                    const double w = x[F[3 * (key[k] & 0xffffffffu) + v]][a];
                    blo[a] = std::min(blo[a], w);
                    bhi[a] = std::max(bhi[a], w);
                }
        const double cc[3] = {0.5 * (blo[0] + bhi[0]), 0.5 * (blo[1] + bhi[1]), 0.5 * (blo[2] + bhi[2])};
        const float cf[3] = {(float)cc[0], (float)cc[1], (float)cc[2]};
        double r2 = 0.0;
        for (uint32_t k = c * kMeshCluster; k < end; ++k)
            for (int v = 0; v < 3; ++v) {
                double d2 = 0.0;
                for (int a = 0; a < 3; ++a) {
                    const double d = (double)x[F[3 * (key[k] & 0xffffffffu) + v]][a] - (double)cf[a];
                    d2 += d * d;
                }
                r2 = std::max(r2, d2);
            }
        const float rf = (float)(std::sqrt(r2) * (1.0 + 1e-6)) * (1.0f + 1e-6f);
        const float sph[4] = {cf[0], cf[1], cf[2], std::isfinite(rf) ? rf : INFINITY};
        std::memcpy(&out.spheres[4 * (size_t)c], sph, sizeof(sph));
    }
    for (uint32_t pos = 0; pos < nt; ++pos) {
        const uint32_t f = (uint32_t)(key[pos] & 0xffffffffu);
        if (f == 0) out.firstPos = pos;
        const V x0 = x[F[3 * f]], x1 = x[F[3 * f + 1]], x2 = x[F[3 * f + 2]];
        // the point-independent terms of PointTriangleDistanceSquared (MeshDistance.cpp:303-315)
        const V e0 = x1 - x0, e1 = x2 - x0;
        const float a00 = gdot(e0, e0), a01 = gdot(e0, e1), a11 = gdot(e1, e1);
        const float det = std::fabs(a00 * a11 - a01 * a01);
        const float denom = (a00 - 2.0f * a01) + a11;
        float* r = &out.tri[16 * (size_t)pos];
        const float invDet = 1.0f / det;
        float index;
        std::memcpy(&index, &f, sizeof(index));
        const float rec[16] = {x0.x, x0.y, x0.z, a00, e0.x, e0.y, e0.z, a01,
                               e1.x, e1.y, e1.z, a11, det, denom, invDet, index};
        std::memcpy(r, rec, sizeof(rec));
        V nrm[7];
        for (int k = 0; k < 3; ++k) nrm[k] = vn[F[3 * f + k]];
        // CalculateEdgeNormal (MeshDistance.cpp:255-269): own face normal plus the opposite face's, or own alone
        for (int k = 0; k < 3; ++k) {
            const int64_t o = opposite[3 * (size_t)f + k];
            nrm[3 + k] = o < 0 ? fn[f] : fn[f] + fn[(size_t)o / 3];
        }
        nrm[6] = fn[f];
        std::memcpy(&out.normals[21 * (size_t)pos], nrm, sizeof(nrm));
    }
    return VFD_OK;
}

}  // namespace vfd

using namespace vfd;

namespace {

template <class T> struct Dev {
    T* p = nullptr;
    hipError_t alloc(size_t n) { return hipMalloc((void**)&p, std::max<size_t>(n, 1) * sizeof(T)); }
    ~Dev() {
        if (p) (void)hipFree(p);
    }
};

// the reference's float loop "for (v = lo; v <= hi; v += step)", its values in order
int loop_values(float lo, float hi, float step, std::vector<float>& out) {
    out.clear();
    for (float v = lo; v <= hi; v += step) {
        if (out.size() >= (1u << 20))
            return set_error(VFD_ERROR_INVALID_ARGUMENT, "sampling lattice too large (radius too small for the mesh)");
        out.push_back(v);
    }
    return VFD_OK;
}

}  // namespace

extern "C" int vfd_sample_mesh_volume(const vfd_mesh* mesh, const float* transform, float radius,
                                      const uint32_t resolution[3], int32_t inverted, int32_t mode, float* out,
                                      uint32_t capacity, uint32_t* count) {
    if (!mesh || !resolution || !count) return set_error(VFD_ERROR_INVALID_ARGUMENT, "mesh, resolution and count must not be NULL");
    if (!(radius > 0.0f) || !std::isfinite(radius)) return set_error(VFD_ERROR_INVALID_ARGUMENT, "radius must be positive");
    if (mode < VFD_SAMPLE_MIN_DENSITY || mode > VFD_SAMPLE_MAX_DENSITY)
        return set_error(VFD_ERROR_INVALID_ARGUMENT, "unknown sample mode");
    for (int a = 0; a < 3; ++a)
        if (resolution[a] == 0) return set_error(VFD_ERROR_INVALID_ARGUMENT, "resolution must be positive");
    if ((uint64_t)resolution[0] * resolution[1] * resolution[2] > (1u << 26))
        return set_error(VFD_ERROR_INVALID_ARGUMENT, "resolution too large");
    MeshPrep prep;
    int rc = prepare_mesh(*mesh, transform, prep);
    if (rc != VFD_OK) return rc;

    // SDF(mesh, bounds, resolution, inverted) (SDF.cu:16-35): the bounds grown by 0.001 of the diagonal, the upper
    // corner first and the lower one with the diagonal of the grown box
    MapDev map{};
    for (int a = 0; a < 3; ++a) {
        map.domMin[a] = prep.lo[a];
        map.domMax[a] = prep.hi[a];
    }
    auto grow = [&] {
        const float d[3] = {map.domMax[0] - map.domMin[0], map.domMax[1] - map.domMin[1],
                            map.domMax[2] - map.domMin[2]};
        return 0.001f * std::sqrt((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]);
    };
    const float gmax = grow();
    for (int a = 0; a < 3; ++a) map.domMax[a] = map.domMax[a] + gmax;
    const float gmin = grow();
    for (int a = 0; a < 3; ++a) map.domMin[a] = map.domMin[a] - gmin;
    for (int a = 0; a < 3; ++a) {
        map.res[a] = resolution[a];
        map.cellSize[a] = (map.domMax[a] - map.domMin[a]) / (float)map.res[a];
        map.cellSizeInv[a] = 1.0f / map.cellSize[a];
    }
    const uint32_t nx = map.res[0], ny = map.res[1], nz = map.res[2];
    map.nodeCount = (nx + 1) * (ny + 1) * (nz + 1) +
                    2 * (nx * (ny + 1) * (nz + 1) + (nx + 1) * ny * (nz + 1) + (nx + 1) * (ny + 1) * nz);

    // the lattice (ParticleSampler.cpp:19-34)
    const float diameter = 2.0f * radius;
    float shiftX = diameter, shiftY = diameter;
    if (mode == VFD_SAMPLE_MEDIUM_DENSITY) {
        shiftY = std::sqrt(3.0f) * radius;
    } else if (mode == VFD_SAMPLE_MAX_DENSITY) {
        shiftX = std::sqrt(3.0f) * radius;
        shiftY = std::sqrt(6.0f) * diameter / 3.0f;
    }
    std::vector<float> xs, ys, zs;
    if ((rc = loop_values(prep.lo[0], prep.hi[0], shiftX, xs)) != VFD_OK) return rc;
    if ((rc = loop_values(prep.lo[1], prep.hi[1], shiftY, ys)) != VFD_OK) return rc;
    if ((rc = loop_values(prep.lo[2], prep.hi[2], diameter, zs)) != VFD_OK) return rc;
    const uint64_t candidates = (uint64_t)xs.size() * ys.size() * zs.size();
    if (candidates > (1u << 30)) return set_error(VFD_ERROR_INVALID_ARGUMENT, "sampling lattice too large");
    if (candidates == 0) {
        *count = 0;
        return VFD_OK;
    }
    const uint32_t n = (uint32_t)candidates, nt = mesh->TriangleCount;

    hipStream_t st = nullptr;
    hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    if (e != hipSuccess) return set_error(VFD_ERROR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() { (void)hipStreamDestroy(s); }
    } guard{st};
    Dev<float> dTri, dNrm, dSph, dNodes, dX, dY, dZ, dOut;
    Dev<uint32_t> dFlags, dOffs, dScratch;
    // declared after the buffers, so destroyed before them: every return drains the stream before the buffers and
    // the host vectors that queued copies read are released
    struct DrainGuard {
        hipStream_t s;
        ~DrainGuard() { (void)hipStreamSynchronize(s); }
    } drain{st};
    if ((e = dTri.alloc(prep.tri.size())) != hipSuccess || (e = dNrm.alloc(prep.normals.size())) != hipSuccess ||
        (e = dSph.alloc(prep.spheres.size())) != hipSuccess ||
        (e = dNodes.alloc(map.nodeCount)) != hipSuccess || (e = dX.alloc(xs.size())) != hipSuccess ||
        (e = dY.alloc(ys.size())) != hipSuccess || (e = dZ.alloc(zs.size())) != hipSuccess ||
        (e = dFlags.alloc(n)) != hipSuccess || (e = dOffs.alloc(n)) != hipSuccess ||
        (e = dScratch.alloc(scan_scratch_words(n))) != hipSuccess || (e = dOut.alloc(3 * (size_t)n)) != hipSuccess)
        return set_error(VFD_ERROR_OUT_OF_MEMORY, std::string("sampler buffers: ") + hipGetErrorString(e));
    const struct {
        void* d;
        const void* h;
        size_t bytes;
    } up[] = {{dTri.p, prep.tri.data(), prep.tri.size() * 4}, {dNrm.p, prep.normals.data(), prep.normals.size() * 4},
              {dSph.p, prep.spheres.data(), prep.spheres.size() * 4},
              {dX.p, xs.data(), xs.size() * 4}, {dY.p, ys.data(), ys.size() * 4}, {dZ.p, zs.data(), zs.size() * 4}};
    for (const auto& u : up)
        if ((e = hipMemcpyAsync(u.d, u.h, u.bytes, hipMemcpyHostToDevice, st)) != hipSuccess)
            return set_error(VFD_ERROR_HIP, std::string("sampler upload: ") + hipGetErrorString(e));
    const MeshDev md{(const float4*)dTri.p, dNrm.p, (const float4*)dSph.p, nt, prep.firstPos};
    launch_mesh_field0(st, md, map, inverted ? -1.0f : 1.0f, 0.0f, dNodes.p);
    map.nodes0 = dNodes.p;
    const SampleGrid g{dX.p, dY.p, dZ.p, (uint32_t)xs.size(), (uint32_t)ys.size(), (uint32_t)zs.size(), mode, radius,
                       diameter, shiftX};
    launch_sample_flags(st, g, map, dNodes.p, dFlags.p);
    exclusive_scan(dFlags.p, dOffs.p, n, dScratch.p, st);
    launch_sample_scatter(st, g, dFlags.p, dOffs.p, dOut.p);
    uint32_t last[2] = {0, 0};
    e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(&last[0], dOffs.p + (n - 1), 4, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(&last[1], dFlags.p + (n - 1), 4, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return set_error(VFD_ERROR_HIP, std::string("mesh sampling: ") + hipGetErrorString(e));
    const uint32_t total = last[0] + last[1];
    *count = total;
    if (out && capacity) {
        const size_t k = std::min<uint32_t>(capacity, total);
        if (k && (e = hipMemcpy(out, dOut.p, 3 * k * sizeof(float), hipMemcpyDeviceToHost)) != hipSuccess)
            return set_error(VFD_ERROR_HIP, std::string("sampler download: ") + hipGetErrorString(e));
    }
    return VFD_OK;
}
