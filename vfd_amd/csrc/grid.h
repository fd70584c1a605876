// grid.h - host-side sizing of one scene's neighbour-search grid (ParticleSearch::ComputeCellInformation,
// VFD/Source/Simulation/DFSPH/ParticleSearch/ParticleSearch.cu:57-93) from the integer cell bounds the min/max pass
// returns (ComputeMinMaxKernel, ParticleSearchKernels.cu:40-62).
//
// The float arithmetic is the reference's, so that gridMin and gridDelta - which decide every particle's cell and
// therefore the neighbour order - are the same bits. What is added is range checking: the bounds come from float ->
// int conversions of positions divided by R, which saturate for a runaway particle, so every intermediate is formed
// in 64-bit integers or checked before it is narrowed, and a grid beyond kMaxGridCells cells is refused instead of
// wrapping around.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>

#include "common.h"

namespace vfd {

// Most cells of all scenes' grids together (cellCount, cellStart: 4 bytes each per cell, i.e. 8 GiB at the limit).
// The 2 x 1 x 1 m tank at r = 0.004 needs 0.7 million cells; a scene needs more than a billion only when particles
// are hundreds of metres apart.
constexpr uint64_t kMaxGridCells = 1ull << 30;

// Fills g (all fields but cellBase and startCorr) for integer cell bounds lo..hi (inclusive, floor(x / R) of the
// scene's particles) and cell size R. Returns the grid's cell count, or 0 when that count (with cellsBefore cells of
// earlier scenes) would exceed kMaxGridCells; g is then left zeroed.
inline uint64_t size_scene_grid(const int lo[3], const int hi[3], float R, uint64_t cellsBefore, SceneGrid& g) {
    std::memset(&g, 0, sizeof(g));
    if (!(R > 0.0f) || !std::isfinite(R)) return 0;
    uint32_t dims[3], meta[3];
    float gmin[3], delta[3];
    uint64_t cells = kMetaBlock;
    for (int a = 0; a < 3; ++a) {
        if (hi[a] < lo[a]) return 0;
        // bounds min = lo * R, max = (hi + 1) * R (the reference's upper bound is the cell after the last one)
        const float bmin = (float)lo[a] * R;
        const float bmax = (float)((int64_t)hi[a] + 1) * R;
        const float size = bmax - bmin;
        const float span = std::ceil(size / R);
        // a grid axis of more than kMaxGridCells cells can never fit; reject before narrowing to uint32
        if (!std::isfinite(span) || !(span >= 0.0f) || span > (float)kMaxGridCells) return 0;
        dims[a] = (uint32_t)span + 4u;
        gmin[a] = bmin - R * 2.0f;
        meta[a] = (uint32_t)std::ceil((float)dims[a] / (float)kMetaGroup);
        const float gs = (float)dims[a] * R;
        delta[a] = (float)dims[a] / gs;
        cells *= meta[a];
        if (cells + cellsBefore > kMaxGridCells) return 0;
    }
    for (int a = 0; a < 3; ++a) {
        g.dims[a] = dims[a];
        g.meta[a] = meta[a];
        g.gridMin[a] = gmin[a];
        g.gridDelta[a] = delta[a];
    }
    g.cellCount = (uint32_t)cells;
    return cells;
}

// Capacity to allocate for `cells` cells: a quarter more, so that a slowly growing grid does not reallocate every
// step, but never above kMaxGridCells.
inline uint64_t grid_capacity(uint64_t cells) {
    const uint64_t want = cells + cells / 4;
    return want > kMaxGridCells ? kMaxGridCells : want;
}

}  // namespace vfd
