// common.h - shared host/device definitions of the MI355X DFSPH step.
//
// Layout (HBM, structure of arrays, particles of scene s in slots [base_s, base_s + count_s), base_s a multiple of
// kTile, slots in Morton cell order and re-sorted every step):
//   positions / velocities / vectors: float4 per particle (w unused), one 16-byte load per neighbour gather
//   scalars: float per particle
//   neighbour lists: uint32 slot indices, transposed list[j * capacity + i] so that the j-th neighbour of 64
//   consecutive particles is one coalesced 256-byte read; at most kMaxNeighbors per particle; for the particles of a
//   staged group also as uint16 halo slots and uint16 kernel-table codes of the pairs, in the same transposed layout;
//   halos: per group, haloCount[g] and the halo's particle slots haloIds[g * kHaloMax + k].
//
// All HIP sources are compiled with -ffp-contract=off: every fused multiply-add is an explicit fmaf, so the
// arithmetic that decides discrete outcomes (neighbour membership, kernel-table index, cell index) is reproducible on
// the CPU. dot3 is the fma chain fma(z,z', fma(y,y', x*x')) used by every length in the step.
#pragma once

#include <stdint.h>

#ifndef __HIPCC__
#define __host__
#define __device__
#define __forceinline__ inline
#endif

namespace vfd {

constexpr unsigned kTile = 256;              // particles per tile = threads per block
constexpr unsigned kMaxNeighbors = 70;       // ParticleSearchKernels.cuh:7
constexpr unsigned kMetaGroup = 8;           // ParticleSearchKernels.cuh:8
constexpr unsigned kMetaBlock = 512;         // ParticleSearchKernels.cuh:9
constexpr unsigned kTableResolution = 10000; // Kernel/DFSPHKernels.h:143
constexpr unsigned kTableEntries = kTableResolution - 1;  // averaged entries 0.5*(T[p]+T[p+1]), p <= 9998
constexpr unsigned kHaltonFloats = 49152;    // HaltonVec323.cuh
constexpr unsigned kMaxBodies = 4;           // boundary bodies per scene
constexpr double kEps = 1.0e-5;              // Core/Math/Math.h:6 (double literal)

// Neighbour groups: kGroupTiles consecutive tiles (one 1024-thread block) whose distinct neighbours - the halo - are
// staged in LDS by the neighbour-loop kernels. A group's halo holds at most kHaloMax particles; a group with more
// keeps gathering from the global arrays (haloCount = kHaloNone). The lists of a staged group are also stored as
// 16-bit halo slots.
// VFD_GROUP_TILES, VFD_HALO_MAX and VFD_GROUPS_PER_CU (the persistent grid of the group kernels, in blocks per CU)
// are compile-time launch shapes the autotuner may vary.
#ifndef VFD_GROUP_TILES
#define VFD_GROUP_TILES 4
#endif
#ifndef VFD_HALO_MAX
#define VFD_HALO_MAX 3808
#endif
#ifndef VFD_GROUPS_PER_CU
#define VFD_GROUPS_PER_CU 1
#endif
constexpr unsigned kGroupTiles = VFD_GROUP_TILES;
constexpr unsigned kGroupSize = kTile * kGroupTiles;
constexpr unsigned kHaloMax = VFD_HALO_MAX;
constexpr unsigned kGroupsPerCU = VFD_GROUPS_PER_CU;
static_assert(kGroupTiles == 1 || kGroupTiles == 2 || kGroupTiles == 4, "a group is 1, 2 or 4 tiles");
static_assert(kHaloMax <= 4 * kGroupSize && kHaloMax <= 65535, "halo slots are 16-bit, filled four per thread");
constexpr uint32_t kHaloNone = 0xFFFFFFFFu;

// Constant per-scene parameters (DFSPHSimulationInfo + DFSPHSimulationDescription, DFSPHImplementation.cu:318-368).
struct SceneParams {
    uint32_t base, count, tileBegin, tileEnd;
    uint32_t bodyCount, pad0, pad1, pad2;
    float R, R2, r, diam;
    float V, rho0, m, mInv;
    float mu, muB, td, sigma;
    float slope, constant, smoothing, NR;
    float mcFactor, gx, gy, gz;
    float divTol, presTol, viscTol, dtMin;
    float dtMax, frameLength, pad3, pad4;
    uint32_t minDiv, maxDiv, minPres, maxPres;
    uint32_t minVisc, maxVisc, stPasses, temporal;
    int32_t enableDiv, enableVisc, enableST, csdFix;
    int32_t csd, pad5, pad6, pad7;
};

// Per-scene neighbour-search grid (ParticleSearch.cu:57-93), recomputed on the host every step.
struct SceneGrid {
    float gridMin[3];
    float gridDelta[3];
    uint32_t dims[3];
    uint32_t meta[3];
    uint32_t cellBase;   // first global cell of the scene
    uint32_t cellCount;  // meta volume * 512
    int32_t startCorr;   // base - (particles of earlier scenes): added to the global exclusive cell scan
    uint32_t pad;
};

// Jacobi solve control (DFSPHImplementation.cu:455-493, 526-564).
struct JacobiState {
    uint32_t active, it, pad0, pad1;
    float err, eta, pad2, pad3;
};

// PCG control (DFSPHImplementation.cu:617-808).
struct CgState {
    uint32_t active, it, zero, loops;  // loops: iterations that ran (profiling attribution)
    float rhs2, thr, res2, absNew;
    float alpha, beta, err, pad1;
};

// Mutable per-scene state (time step, solvers). Device resident; copied to the host at sync points.
struct SceneState {
    float dt, dt2, dtInv, dt2Inv;
    uint32_t S;  // SurfaceTensionSampleCount
    float maxVel;
    float frameTime;
    uint32_t nonfinite;  // set by the step's last kernel when a particle's new position or velocity is not finite
    JacobiState div, pres;
    CgState cg;
};

// Reduction partials: per-tile partial sums (double), kPartialSlots values per tile.
constexpr unsigned kPartialSlots = 4;  // values reduced at once

}  // namespace vfd
