// kernels.h - host-side launchers of the HIP kernels (declarations shared by the .hip files and sim.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include "common.h"

namespace vfd {

struct TileGrid;

// Launch geometry shared by per-particle kernels: kTile threads per block, one tile (kTile particles) at a time.
struct Launch {
    hipStream_t stream;
    const SceneParams* P;        // device [sceneCount]
    const uint32_t* tileScene;   // device [tiles]
    uint32_t tiles;
    uint32_t sceneCount;
    uint32_t capacity;           // tiles * kTile
    uint32_t tilesPerXcd;        // ceil(tiles / 8)
    uint32_t gridAll;            // 8 * tilesPerXcd: one block per tile, XCD-aware order
    uint32_t grid4;              // persistent grid, 4 blocks per CU (40 KB LDS table)
    uint32_t grid2;              // persistent grid, 2 blocks per CU (80 KB LDS table)
    uint32_t gridGroups;         // persistent grid of kGroupSize-thread blocks (one group at a time), 1 per CU
    uint32_t groups;             // 8 * per / kGroupTiles
    struct { uint32_t tiles, per; } tg;  // per = tiles of one XCD chunk, a multiple of 4
};

// ---- adjoint.hip -----------------------------------------------------------------------------------------------
struct AdjK;
constexpr uint32_t kContactRecs = 8;  // float4 records per particle and body of a recorded contact sample
// Geometry and constants of one recorded step and one scene for the adjoint kernels (spline constants
// K = 8/(pi R^3), L = 48/(pi R^3)); arrays in the step's slot order, the scene's n particles in slots
// [base, base + n).
struct AdjKHost {
    const float2* tab;
    float R, invStep, K, L, V, rho0, m, mu, R2;
    uint32_t n, cap;
    const uint32_t* count;
    const uint32_t* list;
    const float4* x;
    const float4* con = nullptr;  // boundary contact records of the step (adj_launch_boundary_tape); null: none
    uint32_t bodies = 0;
    float muB = 0.0f, td = 0.0f;  // boundary viscosity (dynamic) and tangential sample distance
    uint32_t base = 0, stride = 0;  // the scene's first slot; slots per record of the step (contact records; 0: n)
    // group loops of the transposed Jacobi and viscosity kernels: the step's halos (launch_build_halo over its lists)
    // and the persistent grid; grid 0: one thread per particle with global gathers
    struct {
        uint32_t grid = 0, tiles = 0, per = 0;
        const uint32_t* haloCount = nullptr;
        const uint32_t* haloIds = nullptr;
        const uint32_t* local = nullptr;
    } groups;
    AdjK dev() const;
};
int configure_adjoint_kernels();  // dynamic LDS of the group kernels
void adj_launch_accel_T(hipStream_t s, const AdjKHost& h, const float* kap, const float4* la, float* lk, float4* lx);
void adj_launch_div_T(hipStream_t s, const AdjKHost& h, const float* lf, const float4* u, float4* lu, float4* lx);
void adj_launch_jacobi_point(hipStream_t s, uint32_t n, const float* lk, const float* alpha, float div,
                             const float* res, float c, float* lsrc, float* ls, float* lf);
void adj_launch_pres_init(hipStream_t s, uint32_t n, float H, float rho0, const float* alpha, const float* rhoadv,
                          const float* lk, const float* lsrc, const float* lgam, float* lalpha, float* lrho,
                          float* ld2);
void adj_launch_div_init(hipStream_t s, uint32_t n, float h, const float* alpha, const float* D, const float* lk,
                         const float* lsrc, const float* lbet, float* lalpha, float* ld);
void adj_launch_scale4(hipStream_t s, uint32_t n, const float4* in, float f, float4* out);
void adj_launch_axpy4(hipStream_t s, uint32_t n, float f, const float4* in, float4* out);
void adj_launch_fill(hipStream_t s, float* p, float v, uint32_t n);
// scalars of a device-driven BiCGSTAB (adjoint viscosity solve); kernels given a state return at once once `done`
struct BicgState {
    double rhoOld, rhoNew, alpha, omega, beta, rr, tol2;  // tol2: continue while rr > tol2
    uint32_t it, maxIt, done, pad;
};
enum { kBicgRho = 0, kBicgRv = 1, kBicgTsTt = 2, kBicgRr = 3, kBicgBegin = 4 };
struct BicgVectors {
    float4 *x, *r, *rh, *p, *v, *s, *t;
};
void adj_launch_visc_T(hipStream_t s, const AdjKHost& h, float dt, const float* rho, const float4* w, float4* out,
                       const BicgState* st = nullptr);
// the solve's start on the device: state reset, |b|^2 and |r0|^2, the stopping threshold 1e-12 |b|^2, and done
// when b = 0 or r0 already meets it (b, r0 slot-indexed; part holds 2 * ceil(n / kTile) doubles)
void adj_launch_bicg_begin(hipStream_t s, const AdjKHost& h, const float4* b, const float4* r, double* part,
                           BicgState* st, uint32_t maxIt);
// acc[0] += the sum of part[0, blocks) (per-block d/dmu partials of one step), in a fixed order
void adj_launch_mu_acc(hipStream_t s, const double* part, uint32_t blocks, double* acc);
// one BiCGSTAB iteration for A^T x = b on the device; part holds 2 * ceil(n / kTile) doubles
void adj_launch_bicg_iteration(hipStream_t s, const AdjKHost& h, float dt, const float* rho, const BicgVectors& v,
                               double* part, BicgState* st);
void adj_launch_visc_vjp(hipStream_t s, const AdjKHost& h, float dt, const float* rho, const float4* y,
                         const float4* lb, float4* lx, float* lrho, double* muPart);
void adj_launch_factor_density(hipStream_t s, const AdjKHost& h, const float* alpha, const float* lalpha, float* lS,
                               float4* gs, const float* lrho, float4* lx);
void adj_launch_copy3(hipStream_t s, uint32_t n, const float* a, float* da, const float4* b, float4* db,
                      const float* c, float* dc);
void adj_launch_to_slots(hipStream_t s, uint32_t n, const uint32_t* id, const float* orig, float4* slot);
void adj_launch_to_orig(hipStream_t s, uint32_t n, const uint32_t* id, const float4* slot, float* orig);
void adj_launch_to_slots1(hipStream_t s, uint32_t n, const uint32_t* id, const float* orig, float* slot);
void adj_launch_to_orig1(hipStream_t s, uint32_t n, const uint32_t* id, const float* slot, float* orig);
void adj_launch_push(hipStream_t s, uint32_t n, uint32_t stride, const float4* push, float4* lx, float4* lv);
// position cotangent of the step's surface tension acceleration -(sigma / m) E nS (c = -sigma / m) against
// scale * lv over `passes` blend passes (nMC, kMC held); with temporal smoothing dK0 is the blended curvature the
// particles carried into the step (null: none); ldK (null: not tracked) maps the carried curvature's cotangent from
// the step's output to its input; lnc, lcw are n-element scratch
void adj_launch_surface_tension(hipStream_t s, const AdjKHost& h, float scale, const float4* lv, const float4* nMC,
                                const float* kMC, float sf, float c, uint32_t passes, const float* dK0, float* ldK,
                                float4* lnc, float2* lcw, float4* lx);
struct MapDev;
// boundary.hip: replay of the boundary pass of every scene from its starting positions over slots [0, n), recording
// contact samples (con: kContactRecs * kMaxBodies * n records, [(kContactRecs b + q) n + i]) and push-outs (push:
// 3 * n) with their derivatives; flags[3 * scene + k] (contact, push, out of scope or replay mismatch) are set to 1
// where they occur and must be cleared beforehand
void adj_launch_boundary_tape(hipStream_t s, const SceneParams* P, const uint32_t* tileScene, const MapDev* maps,
                              uint32_t n, uint32_t capacity, const float4* xpre, const float4* posAfter, const float* vb,
                              const float4* xb, float4* con, float4* push, uint32_t* flags);

// one-time kernel attributes (dynamic LDS above 64 KB)
int configure_kernels();
int configure_search_kernels();

// Analytic shape + volume-map geometry of one boundary body (device copies live in an array indexed by
// scene * kMaxBodies + body).
struct ShapeDev {
    int32_t shape;
    float c[3], h[3], radius, n[3];
    float sign, tol;
};
struct MapDev {
    float domMin[3], domMax[3];
    float cellSize[3], cellSizeInv[3];
    uint32_t res[3];
    uint32_t nodeCount;
    const float* nodes0;  // signed distance
    const float* nodes1;  // boundary volume
};

// A boundary mesh prepared for the distance kernel. The triangles are stored in the Morton order of their centroids
// (the mesh's own index travels with each record and decides ties). Per triangle, four float4 records hold the
// point-independent terms of the closest-point computation: {x0, a00}, {e0 = x1 - x0, a01}, {e1 = x2 - x0, a11},
// {det, denom, 1/det, index as bits} with a00 = e0.e0, a01 = e0.e1, a11 = e1.e1, det = |a00 a11 - a01^2|,
// denom = a00 - 2 a01 + a11; and 21 floats of normals: the angle-weighted vertex normals of x0, x1, x2, the edge
// normals of x0x1, x1x2, x2x0 and the face normal (MeshDistance.cpp:13-57, 255-284). spheres[c] = {centre, radius}
// bounds the triangles [kMeshCluster c, kMeshCluster (c + 1)) of that order.
constexpr uint32_t kMeshCluster = 64;
struct MeshDev {
    const float4* tri;
    const float* normals;
    const float4* spheres;
    uint32_t triCount;
    uint32_t firstPos;  // stored position of the mesh's triangle 0
};

// The candidate lattice of ParticleSampler::SampleMeshVolume: xs/ys/zs hold the loop variables as the reference's
// float loops accumulate them.
struct SampleGrid {
    const float* xs;
    const float* ys;
    const float* zs;
    uint32_t nx, ny, nz;
    int32_t mode;
    float radius, diameter, shiftX;
};

// Kernel table view (device pointer + scalars).
struct KTabHost {
    const float2* tab;
    float R, R2, invStep, W0;
};

// ---- search.hip -----------------------------------------------------------------------------------------------
uint32_t scan_scratch_words(uint32_t n);
void exclusive_scan(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* scratch, hipStream_t st);
void launch_minmax(const Launch& L, const float4* pos, int* part, int* mm);
void launch_sort(const Launch& L, const SceneGrid* G, uint32_t totalCells, const float4* posIn, const float4* velIn,
                 const float4* dvIn, const float* dkIn, const uint32_t* idIn, float4* posOut, float4* velOut,
                 float4* dvOut, float* dkOut, uint32_t* idOut, uint32_t* cellOf, uint32_t* rank, uint32_t* tmpId,
                 uint32_t* cellCount, uint32_t* cellStart, uint32_t* scanScratch);
void launch_build_halo(const Launch& L, const uint32_t* nbrCount, const uint32_t* nbrList, uint32_t haloLimit,
                       uint32_t* haloCount, uint32_t* haloIds, uint32_t* nbrLocal, uint32_t* haloStats,
                       const float4* pos, const KTabHost& kt);
void launch_build_lists(const Launch& L, const SceneGrid* G, const float4* pos, const uint32_t* cellStart,
                        const uint32_t* cellCount, uint32_t* nbrCount, uint32_t* nbrList,
                        unsigned long long* nbStats);

// Static scratch of the allocation-free step, in 32-bit words: the grid's cell counts and cell starts (kStaticCells
// each), the scan's chunk sums, the surface-tension curvature tables (kStaticMcFloats floats) and their per-scene
// offsets (kStaticMcScenes words). 32 MiB + 272 KiB per device.
constexpr uint32_t kStaticCells = 1u << 22;
constexpr uint32_t kStaticScanWords = kStaticCells / 2048u + 1u;  // scan_scratch_words(kStaticCells)
constexpr uint32_t kStaticMcFloats = 1u << 16;
constexpr uint32_t kStaticMcScenes = 4096;
constexpr size_t kStaticCellCountAt = 0;
constexpr size_t kStaticCellStartAt = kStaticCells;
constexpr size_t kStaticScanAt = 2 * (size_t)kStaticCells;
constexpr size_t kStaticMcTabAt = (kStaticScanAt + kStaticScanWords + 63) / 64 * 64;
constexpr size_t kStaticMcOffAt = kStaticMcTabAt + kStaticMcFloats;
constexpr size_t kStaticScratchWords = kStaticMcOffAt + kStaticMcScenes;
constexpr uint32_t kMaxStaticScratchDevices = 64;
// Device address of that device's static scratch (kStaticScratchWords words); looked up once per device, which must
// be the current device on the first call. Returns 0 on success.
int static_scratch_address(int device, uint32_t** out);

// ---- boundary.hip ---------------------------------------------------------------------------------------------
void launch_map_field0(hipStream_t st, const ShapeDev& shape, const MapDev& map, float* nodes0);
void launch_mesh_field0(hipStream_t st, const MeshDev& mesh, const MapDev& map, float sign, float tol, float* nodes0);
void launch_map_field1(hipStream_t st, const MapDev& map, const float* nodes0, float* nodes1, const KTabHost& kt,
                       const float* glx, const float* glw);
// flags[c] = 1 where candidate c of the lattice lies inside (field 0 < 0), c = (k * ny + j) * nx + i
void launch_sample_flags(hipStream_t st, const SampleGrid& g, const MapDev& map, const float* nodes0, uint32_t* flags);
void launch_sample_scatter(hipStream_t st, const SampleGrid& g, const uint32_t* flags, const uint32_t* offsets,
                           float* out);
void launch_boundary(const Launch& L, const MapDev* maps, float4* pos, float4* vel, float4* xb, float* vb);

// ---- dfsph.hip ------------------------------------------------------------------------------------------------
struct Fields {
    float4* pos;      // working positions (sorted)
    float4* posNext;  // end-of-step positions
    float4* vel;
    float4* acc;
    float4* ap;       // pressure acceleration
    float4* dv;       // VelocityDifference
    float* dk;        // DeltaFinalCurvature
    const uint32_t* id;
    float* density;
    float* factor;
    float* densAdv;
    float* kappa;
    float* kappaV;
    float* residuum;
    float4* nMC;
    float4* nS;
    float* kMC;
    float* kS;
    float4* xb;       // [kMaxBodies][capacity]
    float* vb;        // [kMaxBodies][capacity]
    const uint32_t* nbrCount;
    const uint32_t* nbrList;
    const uint32_t* nbrLocal;    // [j][capacity] halo slot | table code << 16 (staged groups)
    const uint32_t* haloCount;   // [groups] kHaloNone: global gathers
    const uint32_t* haloIds;     // [groups][kHaloMax]
    // viscosity solver
    float* minv;      // [9][capacity]
    float4* cgX;      // solution (ViscosityGradientG)
    float4* cgR;
    float4* cgP;
    float4* cgZ;
    float4* cgAp;
    // reductions
    double* partials;      // [tiles][kPartialSlots]
    float* partialsF;      // [tiles]
    uint32_t* live;        // [3] scenes still iterating in the divergence, pressure and viscosity solves
    SceneState* state;     // [sceneCount]
    const float* halton;   // [49152]
    // per scene s, mcTab[mcOff[s] + count] = cos(acos(1 - 2 count / S) + MonteCarloFactor) for count = 0..S (the
    // curvature factor of the surface-tension classification, evaluated on the host for this step's S)
    const float* mcTab;
    const uint32_t* mcOff;
};

void launch_density_factor(const Launch& L, const KTabHost& kt, const Fields& F);
void launch_scene_init_div(const Launch& L, const Fields& F);
void launch_div_init(const Launch& L, const KTabHost& kt, const Fields& F);
void launch_div_accel(const Launch& L, const KTabHost& kt, const Fields& F);
void launch_div_iter(const Launch& L, const KTabHost& kt, const Fields& F);
void launch_jacobi_finish(const Launch& L, const Fields& F, int which);
void launch_div_final(const Launch& L, const KTabHost& kt, const Fields& F);
void launch_surface_tension(const Launch& L, const Fields& F);
void launch_visc_precond(const Launch& L, const KTabHost& kt, const Fields& F);
void launch_cg_begin(const Launch& L, const KTabHost& kt, const Fields& F);
void launch_cg_begin_finish(const Launch& L, const Fields& F);
void launch_cg_matvec(const Launch& L, const KTabHost& kt, const Fields& F);
void launch_cg_alpha(const Launch& L, const Fields& F);
void launch_cg_update(const Launch& L, const Fields& F);
void launch_cg_update_finish(const Launch& L, const Fields& F);
void launch_cg_direction(const Launch& L, const Fields& F);
void launch_visc_apply(const Launch& L, const Fields& F);
void launch_time_step(const Launch& L, const Fields& F);
void launch_time_step_finish(const Launch& L, const Fields& F);
void launch_velocity(const Launch& L, const Fields& F);
void launch_frame_pack(hipStream_t s, uint32_t base, uint32_t count, const uint32_t* id, const float4* x,
                       const float4* v, const float4* a, float* out);
void launch_scene_init_pres(const Launch& L, const Fields& F);
void launch_pres_init(const Launch& L, const KTabHost& kt, const Fields& F);
void launch_pres_accel(const Launch& L, const KTabHost& kt, const Fields& F);
void launch_pres_iter(const Launch& L, const KTabHost& kt, const Fields& F);
void launch_pres_final(const Launch& L, const KTabHost& kt, const Fields& F, int writeNext);
void launch_scene_state_init(const Launch& L, SceneState* state, const float* dt0);

}  // namespace vfd
