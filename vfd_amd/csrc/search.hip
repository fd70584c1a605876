// search.hip - compact-grid neighbour search with a Morton-ordered structure-of-arrays reorder.
//
// Reference: ParticleSearch::FindNeighbors (VFD/Source/Simulation/DFSPH/ParticleSearch/ParticleSearch.cu:29-195,
// ParticleSearchKernels.cu:4-193). Same grid (cell = support radius, origin two cells below the particles' minimum
// cell, 8^3 Morton meta blocks), same strict test 0 < d^2 < R^2, same cap of 70 and the same traversal order (27 cells
// z,y,x). Different from the reference by design:
//   - particles are physically reordered by (cell, original id) every step, so every later neighbour loop reads
//     contiguous cells; the reference keeps particles in input order and only sorts index arrays;
//   - the order inside a cell is ascending original id (deterministic) instead of atomicAdd order;
//   - lists hold slot indices in a transposed [j][i] layout with a fixed capacity of 70 per particle.
#include <mutex>

#include "device.h"
#include "kernels.h"

namespace vfd {

__device__ __forceinline__ unsigned spread3(unsigned x) {  // low 10 bits of x to every third bit
    x &= 0x3ffu;
    x = (x | (x << 16)) & 0x030000ffu;
    x = (x | (x << 8)) & 0x0300f00fu;
    x = (x | (x << 4)) & 0x030c30c3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}
// cell (cx,cy,cz) -> meta-block linear index * 512 + Morton code inside the 8^3 block (ParticleSearchKernels.cu:24-38)
__device__ __forceinline__ uint32_t cell_key(const SceneGrid& g, uint32_t cx, uint32_t cy, uint32_t cz) {
    const uint32_t mx = cx >> 3, my = cy >> 3, mz = cz >> 3;
    const uint32_t lin = (mz * g.meta[1] + my) * g.meta[0] + mx;
    const uint32_t m = (spread3(cz & 7u) << 2) | (spread3(cy & 7u) << 1) | spread3(cx & 7u);
    return g.cellBase + lin * kMetaBlock + m;
}

// ComputeMinMaxKernel (ParticleSearchKernels.cu:40-62): integer cell bounds floor(x / R). Each tile writes its
// bounds to part[tile * 6 + a]; k_minmax_finish reduces a scene's tiles (no atomics on a single address).
__global__ void __launch_bounds__(kTile) k_minmax(const float4* __restrict__ pos, const SceneParams* __restrict__ P,
                                                  const uint32_t* __restrict__ tileScene, int* part) {
    __shared__ int sh[kTile / kWave][6];
    const TileCtx c = tile_ctx(tileScene, P);
    const float R = P[c.s].R;
    int v[6] = {INT_MAX, INT_MAX, INT_MAX, INT_MIN, INT_MIN, INT_MIN};
    if (c.valid) {
        const float4 x = pos[c.i];
        v[0] = v[3] = (int)floorf(x.x / R);
        v[1] = v[4] = (int)floorf(x.y / R);
        v[2] = v[5] = (int)floorf(x.z / R);
    }
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        v[a] = wave_min_i(v[a]);
        v[3 + a] = wave_max_i(v[3 + a]);
    }
    if (lane == 0)
#pragma unroll
        for (int a = 0; a < 6; ++a) sh[w][a] = v[a];
    __syncthreads();
    if (threadIdx.x < 6) {
        const int a = threadIdx.x;
        int m = sh[0][a];
        for (int j = 1; j < (int)(kTile / kWave); ++j) m = a < 3 ? min(m, sh[j][a]) : max(m, sh[j][a]);
        part[(size_t)blockIdx.x * 6 + a] = m;
    }
}

__global__ void __launch_bounds__(kTile) k_minmax_finish(const SceneParams* __restrict__ P,
                                                         const int* __restrict__ part, int* mm) {
    __shared__ int sh[kTile / kWave][6];
    const uint32_t s = blockIdx.x;
    const SceneParams& p = P[s];
    int v[6] = {INT_MAX, INT_MAX, INT_MAX, INT_MIN, INT_MIN, INT_MIN};
    for (uint32_t t = p.tileBegin + threadIdx.x; t < p.tileEnd; t += kTile)
#pragma unroll
        for (int a = 0; a < 6; ++a) v[a] = a < 3 ? min(v[a], part[(size_t)t * 6 + a]) : max(v[a], part[(size_t)t * 6 + a]);
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        v[a] = wave_min_i(v[a]);
        v[3 + a] = wave_max_i(v[3 + a]);
    }
    if (lane == 0)
#pragma unroll
        for (int a = 0; a < 6; ++a) sh[w][a] = v[a];
    __syncthreads();
    if (threadIdx.x < 6) {
        const int a = threadIdx.x;
        int m = sh[0][a];
        for (int j = 1; j < (int)(kTile / kWave); ++j) m = a < 3 ? min(m, sh[j][a]) : max(m, sh[j][a]);
        mm[6 * s + a] = m;
    }
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__device__ __forceinline__ uint32_t clampu(int v, uint32_t hi) { return v < 0 ? 0u : ((uint32_t)v >= hi ? hi - 1u : (uint32_t)v); }

// InsertParticlesMortonKernel (ParticleSearchKernels.cu:64-78): cell key and arbitrary rank inside the cell.
// Coordinates are clamped into the grid so that a non-finite position cannot index outside the cell arrays.
__global__ void __launch_bounds__(kTile) k_insert(const float4* __restrict__ pos, const SceneParams* __restrict__ P,
                                                  const SceneGrid* __restrict__ G,
                                                  const uint32_t* __restrict__ tileScene, uint32_t* cellOf,
                                                  uint32_t* rank, uint32_t* cellCount) {
    const TileCtx c = tile_ctx(tileScene, P);
    if (!c.valid) return;
    const SceneGrid& g = G[c.s];
    const float4 x = pos[c.i];
    const float fx = (x.x - g.gridMin[0]) * g.gridDelta[0];
    const float fy = (x.y - g.gridMin[1]) * g.gridDelta[1];
    const float fz = (x.z - g.gridMin[2]) * g.gridDelta[2];
    const uint32_t key = cell_key(g, clampu((int)fx, g.dims[0]), clampu((int)fy, g.dims[1]), clampu((int)fz, g.dims[2]));
    cellOf[c.i] = key;
    rank[c.i] = atomicAdd(&cellCount[key], 1u);
}

// Places every particle's original id at its provisional slot (CountingSortIndicesKernel, :80-91).
__global__ void __launch_bounds__(kTile) k_scatter_ids(const SceneParams* __restrict__ P,
                                                       const SceneGrid* __restrict__ G,
                                                       const uint32_t* __restrict__ tileScene,
                                                       const uint32_t* __restrict__ cellOf,
                                                       const uint32_t* __restrict__ rank,
                                                       const uint32_t* __restrict__ cellStart,
                                                       const uint32_t* __restrict__ idOld, uint32_t* tmpId) {
    const TileCtx c = tile_ctx(tileScene, P);
    if (!c.valid) return;
    const uint32_t key = cellOf[c.i];
    tmpId[(uint32_t)((int32_t)cellStart[key] + G[c.s].startCorr) + rank[c.i]] = idOld[c.i];
}

// Final slot = cell start + number of cell members with a smaller original id. Moves the persistent state there.
__global__ void __launch_bounds__(kTile) k_reorder(const SceneParams* __restrict__ P, const SceneGrid* __restrict__ G,
                                                   const uint32_t* __restrict__ tileScene,
                                                   const uint32_t* __restrict__ cellOf,
                                                   const uint32_t* __restrict__ cellStart,
                                                   const uint32_t* __restrict__ cellCount,
                                                   const uint32_t* __restrict__ tmpId, const float4* __restrict__ posIn,
                                                   const float4* __restrict__ velIn, const float4* __restrict__ dvIn,
                                                   const float* __restrict__ dkIn, const uint32_t* __restrict__ idIn,
                                                   float4* posOut, float4* velOut, float4* dvOut, float* dkOut,
                                                   uint32_t* idOut) {
    const TileCtx c = tile_ctx(tileScene, P);
    if (!c.valid) return;
    const uint32_t key = cellOf[c.i];
    const uint32_t start = (uint32_t)((int32_t)cellStart[key] + G[c.s].startCorr);
    const uint32_t n = cellCount[key];
    const uint32_t me = idIn[c.i];
    uint32_t r = 0;
    for (uint32_t t = 0; t < n; ++t) r += tmpId[start + t] < me ? 1u : 0u;
    const uint32_t dst = start + r;
    posOut[dst] = posIn[c.i];
    velOut[dst] = velIn[c.i];
    dvOut[dst] = dvIn[c.i];
    dkOut[dst] = dkIn[c.i];
    idOut[dst] = me;
}

#ifndef VFD_SEARCH_UNROLL
#define VFD_SEARCH_UNROLL 8
#endif
#ifndef VFD_STAGE_LISTS
#define VFD_STAGE_LISTS 0
#endif
constexpr uint32_t kSearchUnroll = VFD_SEARCH_UNROLL;

// FindNeighbors (ParticleSearch.cu:95-195; kernels ParticleSearchKernels.cu:93-193) in one traversal per particle:
// the 27 cells in z, y, x order, each cell's particles in slot order (= ascending original id), strict
// 0 < d^2 < R^2, at most kMaxNeighbors; out(cnt, t) receives the cnt-th neighbour. The ranges of all 27 cells are
// requested before the first candidate is loaded. The three cells of a row (x - 1, x, x + 1 at fixed y, z) form one
// stream of candidates, loaded kSearchUnroll at a time from clamped positions (every load stays inside a range) with
// no branch around the loads; a candidate past the row's end or past the cap is simply not accepted, so the first
// kMaxNeighbors in the reference's order are kept. Cells whose box lies farther than the support radius from the
// particle (by a 0.1% margin: no candidate there can pass d^2 < R^2) are skipped: about a quarter of the 27 on
// average, corner and edge cells, without changing the list.
template <class Out>
__device__ __forceinline__ uint32_t neighbor_list(const TileCtx& c, const SceneParams* __restrict__ P,
                                                  const SceneGrid* __restrict__ G, const float4* __restrict__ pos,
                                                  const uint32_t* __restrict__ cellStart,
                                                  const uint32_t* __restrict__ cellCount, Out&& out) {
    const SceneGrid& g = G[c.s];
    const float R2 = P[c.s].R * P[c.s].R;
    const float4 xi = pos[c.i];
    const float ux = (xi.x - g.gridMin[0]) * g.gridDelta[0], uy = (xi.y - g.gridMin[1]) * g.gridDelta[1];
    const float uz = (xi.z - g.gridMin[2]) * g.gridDelta[2];
    const int gx = (int)floorf(ux), gy = (int)floorf(uy), gz = (int)floorf(uz);
    // squared distances (world units) from the particle to the slabs of the cells below / above along each axis
    const float csx = 1.0f / g.gridDelta[0], csy = 1.0f / g.gridDelta[1], csz = 1.0f / g.gridDelta[2];
    const float fx = ux - (float)gx, fy = uy - (float)gy, fz = uz - (float)gz;
    const float bx[3] = {fx * csx * fx * csx, 0.0f, (1.0f - fx) * csx * (1.0f - fx) * csx};
    const float by[3] = {fy * csy * fy * csy, 0.0f, (1.0f - fy) * csy * (1.0f - fy) * csy};
    const float bz[3] = {fz * csz * fz * csz, 0.0f, (1.0f - fz) * csz * (1.0f - fz) * csz};
    const float lim = R2 * 1.002f;
    uint32_t cs[27], cn[27];
#pragma unroll
    for (int k = 0; k < 27; ++k) {
        const int ox = k % 3, oy = (k / 3) % 3, oz = k / 9;
        const int cx = gx + ox - 1, cy = gy + oy - 1, cz = gz + oz - 1;
        const bool in = cx >= 0 && cx < (int)g.dims[0] && cy >= 0 && cy < (int)g.dims[1] && cz >= 0 &&
                        cz < (int)g.dims[2] && bx[ox] + by[oy] + bz[oz] <= lim;
        const uint32_t key = in ? cell_key(g, (uint32_t)cx, (uint32_t)cy, (uint32_t)cz) : 0u;
        cn[k] = in ? cellCount[key] : 0u;
        cs[k] = in ? (uint32_t)((int32_t)cellStart[key] + g.startCorr) : 0u;
    }
    uint32_t cnt = 0;
#pragma unroll
    for (int row = 0; row < 9; ++row) {
        const uint32_t sa = cs[3 * row], sb = cs[3 * row + 1], sc = cs[3 * row + 2];
        const uint32_t a = cn[3 * row], ab = a + cn[3 * row + 1], L = ab + cn[3 * row + 2];
        for (uint32_t q0 = 0; q0 < L; q0 += kSearchUnroll) {
            uint32_t t[kSearchUnroll];
            float4 xj[kSearchUnroll];
#pragma unroll
            for (uint32_t u = 0; u < kSearchUnroll; ++u) {
                const uint32_t q = min(q0 + u, L - 1u);
                t[u] = q < a ? sa + q : (q < ab ? sb + (q - a) : sc + (q - ab));
                xj[u] = pos[t[u]];
            }
#pragma unroll
            for (uint32_t u = 0; u < kSearchUnroll; ++u) {
                const float ddx = xj[u].x - xi.x, ddy = xj[u].y - xi.y, ddz = xj[u].z - xi.z;
                const float d2 = fmaf(ddz, ddz, fmaf(ddy, ddy, ddx * ddx));
                const bool acc = q0 + u < L && d2 < R2 && d2 > 0.0f && cnt < kMaxNeighbors;
                if (acc) out(cnt, t[u]);
                cnt += acc ? 1u : 0u;
            }
        }
    }
    return cnt;
}

// With VFD_STAGE_LISTS the lists are staged in LDS ([kMaxNeighbors][kTile] per block) while the traversals run,
// then written row by row, so every global store of a row is one coalesced 1 KB write of the tile; otherwise each
// entry is stored as it is found.
constexpr unsigned kLdsListStage = VFD_STAGE_LISTS ? kMaxNeighbors * kTile * sizeof(uint32_t) : 0u;
__global__ void __launch_bounds__(kTile) k_build_lists(const float4* __restrict__ pos, const SceneParams* __restrict__ P,
                                                       const SceneGrid* __restrict__ G,
                                                       const uint32_t* __restrict__ tileScene,
                                                       const uint32_t* __restrict__ cellStart,
                                                       const uint32_t* __restrict__ cellCount, uint32_t* nbrCount,
                                                       uint32_t* nbrList, uint32_t capacity,
                                                       unsigned long long* nbStats) {
    extern __shared__ uint32_t stage[];  // [kMaxNeighbors][kTile]
    __shared__ uint32_t rows;
    const TileCtx c = tile_ctx(tileScene, P);
    if (threadIdx.x == 0) rows = 0;
    __syncthreads();
    uint32_t cnt = 0;
    if (c.valid) {
        cnt = neighbor_list(c, P, G, pos, cellStart, cellCount, [&](uint32_t j, uint32_t t) {
            if (VFD_STAGE_LISTS)
                stage[j * kTile + threadIdx.x] = t;
            else
                nbrList[(size_t)j * capacity + c.i] = t;
        });
        nbrCount[c.i] = cnt;
        if (cnt) atomicMax(&rows, cnt);
    }
    if (VFD_STAGE_LISTS) {
        __syncthreads();
        const uint32_t nr = rows;
        for (uint32_t j = 0; j < nr; ++j)
            if (j < cnt) nbrList[(size_t)j * capacity + c.i] = stage[j * kTile + threadIdx.x];
    }
    // per-scene total and maximum neighbour count (debug statistics)
    int tot = (int)cnt, mx = (int)cnt;
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) {
        tot += __shfl_xor(tot, o, kWave);
        mx = max(mx, __shfl_xor(mx, o, kWave));
    }
    if ((threadIdx.x & (kWave - 1)) == 0 && tot > 0) {
        atomicAdd(&nbStats[2 * c.s], (unsigned long long)tot);
        atomicMax(&nbStats[2 * c.s + 1], (unsigned long long)mx);
    }
}

// ---- cell-parallel list build (VFD_SEARCH_CELLS) ---------------------------------------------------------------------
#ifndef VFD_SEARCH_CELLS
#define VFD_SEARCH_CELLS 0
#endif
constexpr uint32_t kCellWaves = kTile / kWave;

__device__ __forceinline__ uint32_t slot_cell(const SceneGrid& g, const float4& x) {  // k_insert's key
    const float fx = (x.x - g.gridMin[0]) * g.gridDelta[0];
    const float fy = (x.y - g.gridMin[1]) * g.gridDelta[1];
    const float fz = (x.z - g.gridMin[2]) * g.gridDelta[2];
    return cell_key(g, clampu((int)fx, g.dims[0]), clampu((int)fy, g.dims[1]), clampu((int)fz, g.dims[2]));
}

// Every wave owns the 64 slots of its window and builds exactly their lists. The window's particles fall into runs of
// one grid cell (segments: a cell may begin before the window or end after it; only the window's own particles are
// handled). A segment's particles share the 27 candidate ranges, so the wave loads the candidates once, 64 per chunk
// with a lane per candidate (consecutive slots of a range are consecutive lanes: few, coalesced loads instead of one
// load per particle and candidate), and tests them against each of the segment's particles in turn; a ballot gives
// each accepted candidate its rank in that particle's list (candidates are visited in list order: ranges in z, y, x
// order, slots ascending), so the first kMaxNeighbors are kept exactly as the per-particle traversal keeps them.
// Entries are staged in LDS as 16-bit {range, offset} codes in [row][lane] (rows padded so that the ranks of one
// particle fall into different banks), decoded with the segment's range starts and written at the end as whole
// 256-byte rows of 64 consecutive slots. A segment that does not fit the codes (a range of 2048 or more particles,
// or a window of more than kWinSegs cells) stores its entries directly instead.
constexpr uint32_t kWinPitch = kWave + 2;  // u16 entries per staged row
constexpr uint32_t kWinSegs = 16;          // segments per window with staged entries
constexpr uint32_t kWinDirect = 0xFFu;     // segment index of a lane whose entries went straight to HBM
// lane q's value (q wave-uniform), read into a scalar register
__device__ __forceinline__ float rdlane(float v, uint32_t q) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), (int)q));
}
__device__ __forceinline__ uint32_t rdlane(uint32_t v, uint32_t q) {
    return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)q);
}
__global__ void __launch_bounds__(kTile) k_build_lists_cells(const float4* __restrict__ pos,
                                                           const SceneParams* __restrict__ P,
                                                           const SceneGrid* __restrict__ G,
                                                           const uint32_t* __restrict__ tileScene, uint32_t tiles,
                                                           uint32_t per, const uint32_t* __restrict__ cellStart,
                                                           const uint32_t* __restrict__ cellCount, uint32_t* nbrCount,
                                                           uint32_t* nbrList, uint32_t capacity,
                                                           unsigned long long* nbStats) {
    __shared__ uint16_t stage[kCellWaves][kMaxNeighbors * kWinPitch];
    __shared__ uint32_t tabStart[kCellWaves][kWinSegs][28];  // range starts of the staged segments
    __shared__ uint32_t pre[kCellWaves][32];  // exclusive prefix of the 27 range lengths of the current segment
    __shared__ uint32_t cst[kCellWaves][32];  // first slot of each range of the current segment
    const uint32_t tile = (blockIdx.x & 7u) * per + (blockIdx.x >> 3);
    if (tile >= tiles) return;  // block-uniform
    const uint32_t w = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave), lane = threadIdx.x & (kWave - 1);
    const uint32_t s = tileScene[tile];
    const SceneParams& p = P[s];
    const SceneGrid& g = G[s];
    const uint32_t end = p.base + p.count;
    const uint32_t win0 = tile * kTile + w * kWave;  // first slot of the window
    const uint32_t i = win0 + lane;
    const bool valid = i < end;
    const float4 x = valid ? pos[i] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const uint32_t key = valid ? slot_cell(g, x) : 0xFFFFFFFFu;
    const uint32_t prevKey = __shfl_up(key, 1, kWave);
    unsigned long long segs = __ballot(valid && (lane == 0 || key != prevKey));  // first lane of each segment
    const uint32_t nvalid = (uint32_t)__popcll(__ballot(valid));  // the valid lanes are the first nvalid
    const float R2 = p.R * p.R;
    uint16_t* st = stage[w];
    uint32_t cnt = 0;            // accepted neighbours of this lane's particle (only the first kMaxNeighbors are kept)
    uint32_t mySeg = kWinDirect;  // this lane's staged segment
    for (uint32_t seg = 0; segs; ++seg) {
        const uint32_t b = (uint32_t)__ffsll(segs) - 1u;
        segs &= segs - 1ull;
        const uint32_t e = segs ? (uint32_t)__ffsll(segs) - 1u : nvalid;  // segment lanes [b, e)
        const float xlx = rdlane(x.x, b), xly = rdlane(x.y, b), xlz = rdlane(x.z, b);
        // the cell of the key (k_insert's clamped cell; for a particle inside the grid the cell the per-particle
        // traversal takes with floorf)
        const int gx = (int)clampu((int)((xlx - g.gridMin[0]) * g.gridDelta[0]), g.dims[0]);
        const int gy = (int)clampu((int)((xly - g.gridMin[1]) * g.gridDelta[1]), g.dims[1]);
        const int gz = (int)clampu((int)((xlz - g.gridMin[2]) * g.gridDelta[2]), g.dims[2]);
        uint32_t cn = 0, cs = 0;
        if (lane < 27) {
            const int cx = gx + (int)(lane % 3) - 1, cy = gy + (int)((lane / 3) % 3) - 1, cz = gz + (int)(lane / 9) - 1;
            if (cx >= 0 && cx < (int)g.dims[0] && cy >= 0 && cy < (int)g.dims[1] && cz >= 0 && cz < (int)g.dims[2]) {
                const uint32_t k = cell_key(g, (uint32_t)cx, (uint32_t)cy, (uint32_t)cz);
                cn = cellCount[k];
                cs = (uint32_t)((int32_t)cellStart[k] + g.startCorr);
            }
        }
        const bool staged = seg < kWinSegs && __ballot(cn >= 2048u) == 0ull;  // wave-uniform
        uint32_t incl = cn;
#pragma unroll
        for (int o = 1; o < 32; o <<= 1) {
            const uint32_t t = __shfl_up(incl, o, kWave);
            if ((int)lane >= o) incl += t;
        }
        const uint32_t T = rdlane(incl, 26);  // candidates in the 27 ranges
        if (lane < 27) {
            pre[w][lane] = incl - cn;
            cst[w][lane] = cs;
            if (staged) tabStart[w][seg][lane] = cs;
        }
        if (lane >= b && lane < e) mySeg = staged ? seg : kWinDirect;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // candidate t of the segment's stream: its range (the last one whose prefix is <= t) and offset in it
        auto locate = [&](uint32_t t, uint32_t& off) {
            uint32_t r = 0;
#pragma unroll
            for (uint32_t step = 16; step > 0; step >>= 1)
                if (r + step < 27 && pre[w][r + step] <= t) r += step;
            off = t - pre[w][r];
            return r;
        };
        uint32_t offN = 0, rN = lane < T ? locate(lane, offN) : 0u;
        uint32_t slotN = lane < T ? cst[w][rN] + offN : 0u;
        float4 xjN = lane < T ? pos[slotN] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        for (uint32_t c0 = 0; c0 < T; c0 += kWave) {
            const uint32_t slot = slotN, code = (rN << 11) | offN;
            const float4 xj = xjN;
            const bool has = c0 + lane < T;
            if (c0 + kWave < T) {  // the next chunk's candidates are requested before this chunk is tested
                const uint32_t tn = c0 + kWave + lane;
                rN = tn < T ? locate(tn, offN) : 0u;
                slotN = tn < T ? cst[w][rN] + offN : 0u;
                xjN = tn < T ? pos[slotN] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            }
            for (uint32_t q = b; q < e; ++q) {  // wave-uniform: the segment's particles in slot order
                const float xqx = rdlane(x.x, q), xqy = rdlane(x.y, q), xqz = rdlane(x.z, q);
                const float ddx = xj.x - xqx, ddy = xj.y - xqy, ddz = xj.z - xqz;
                const float d2 = fmaf(ddz, ddz, fmaf(ddy, ddy, ddx * ddx));
                const bool acc = has && d2 < R2 && d2 > 0.0f;
                const unsigned long long bal = __ballot(acc);
                const uint32_t cq = rdlane(cnt, q);
                if (acc) {
                    const uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32),
                                                                     __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u));
                    const uint32_t rank = cq + below;  // accepted lanes before this one
                    if (rank < kMaxNeighbors) {
                        if (staged)
                            st[rank * kWinPitch + q] = (uint16_t)code;
                        else
                            nbrList[(size_t)rank * capacity + win0 + q] = slot;
                    }
                }
                if (lane == q) cnt += (uint32_t)__popcll(bal);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    // the window's rows, whole: row r holds entry r of each of the 64 slots
    const uint32_t n = valid ? min(cnt, kMaxNeighbors) : 0u;
    uint32_t rows = n;
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) rows = max(rows, (uint32_t)__shfl_xor(rows, o, kWave));
    const uint32_t* ts = tabStart[w][mySeg < kWinSegs ? mySeg : 0u];
    for (uint32_t r = 0; r < rows; ++r) {
        if (r < n && mySeg != kWinDirect) {
            const uint32_t code = st[r * kWinPitch + lane];
            nbrList[(size_t)r * capacity + i] = ts[code >> 11] + (code & 2047u);
        }
    }
    if (valid) nbrCount[i] = n;
    // per-scene total and maximum neighbour count (debug statistics)
    unsigned long long t = n;
    uint32_t mx = n;
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) {
        t += __shfl_xor(t, o, kWave);
        mx = max(mx, (uint32_t)__shfl_xor(mx, o, kWave));
    }
    if (lane == 0 && t > 0) {
        atomicAdd(&nbStats[2 * s], t);
        atomicMax(&nbStats[2 * s + 1], (unsigned long long)mx);
    }
}

// ---- exclusive scan of uint32 (cell counts) --------------------------------------------------------------------
constexpr unsigned kScanThreads = 256, kScanItems = 8, kScanChunk = kScanThreads * kScanItems;
static_assert(kStaticScanWords >= (kStaticCells + kScanChunk - 1) / kScanChunk + 1,
              "static scratch must hold the chunk sums of kStaticCells cells");

__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t& total) {
    __shared__ uint32_t wsum[kScanThreads / kWave];
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    uint32_t incl = v;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const uint32_t t = __shfl_up(incl, o, kWave);
        if (lane >= o) incl += t;
    }
    if (lane == kWave - 1) wsum[w] = incl;
    __syncthreads();
    uint32_t before = 0, tot = 0;
    for (int j = 0; j < (int)(kScanThreads / kWave); ++j) {
        if (j < w) before += wsum[j];
        tot += wsum[j];
    }
    __syncthreads();
    total = tot;
    return before + incl - v;
}

__global__ void __launch_bounds__(kScanThreads) k_scan_chunks(const uint32_t* __restrict__ in, uint32_t* out,
                                                              uint32_t* chunkSums, uint32_t n) {
    const uint32_t base = blockIdx.x * kScanChunk + threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    uint32_t local = 0;
#pragma unroll
    for (unsigned k = 0; k < kScanItems; ++k) {
        v[k] = base + k < n ? in[base + k] : 0u;
        local += v[k];
    }
    uint32_t total;
    uint32_t run = block_exclusive_scan(local, total);
#pragma unroll
    for (unsigned k = 0; k < kScanItems; ++k) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
    if (threadIdx.x == 0) chunkSums[blockIdx.x] = total;
}

// single block: exclusive scan of the chunk sums in place (any count, sequential over blocks of kScanThreads)
__global__ void __launch_bounds__(kScanThreads) k_scan_sums(uint32_t* sums, uint32_t n) {
    uint32_t carry = 0;
    for (uint32_t b = 0; b < n; b += kScanThreads) {
        const uint32_t i = b + threadIdx.x;
        const uint32_t v = i < n ? sums[i] : 0u;
        uint32_t total;
        const uint32_t ex = block_exclusive_scan(v, total);
        if (i < n) sums[i] = carry + ex;
        carry += total;
    }
}

__global__ void __launch_bounds__(kScanThreads) k_scan_add(uint32_t* out, const uint32_t* __restrict__ chunkSums,
                                                           uint32_t n) {
    const uint32_t add = chunkSums[blockIdx.x];
    const uint32_t base = blockIdx.x * kScanChunk;
    for (unsigned k = threadIdx.x; k < kScanChunk; k += kScanThreads)
        if (base + k < n) out[base + k] += add;
}

uint32_t scan_scratch_words(uint32_t n) { return (n + kScanChunk - 1) / kScanChunk + 1; }

void exclusive_scan(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* scratch, hipStream_t st) {
    if (n == 0) return;
    const uint32_t chunks = (n + kScanChunk - 1) / kScanChunk;
    hipLaunchKernelGGL(k_scan_chunks, dim3(chunks), dim3(kScanThreads), 0, st, in, out, scratch, n);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kScanThreads), 0, st, scratch, chunks);
    hipLaunchKernelGGL(k_scan_add, dim3(chunks), dim3(kScanThreads), 0, st, out, (const uint32_t*)scratch, n);
}

// ---- host launchers ---------------------------------------------------------------------------------------------
void launch_minmax(const Launch& L, const float4* pos, int* part, int* mm) {
    hipLaunchKernelGGL(k_minmax, dim3(L.tiles), dim3(kTile), 0, L.stream, pos, L.P, L.tileScene, part);
    hipLaunchKernelGGL(k_minmax_finish, dim3(L.sceneCount), dim3(kTile), 0, L.stream, L.P, (const int*)part, mm);
}

void launch_sort(const Launch& L, const SceneGrid* G, uint32_t totalCells, const float4* posIn, const float4* velIn,
                 const float4* dvIn, const float* dkIn, const uint32_t* idIn, float4* posOut, float4* velOut,
                 float4* dvOut, float* dkOut, uint32_t* idOut, uint32_t* cellOf, uint32_t* rank, uint32_t* tmpId,
                 uint32_t* cellCount, uint32_t* cellStart, uint32_t* scanScratch) {
    hipLaunchKernelGGL(k_fill_u32, dim3((totalCells + 255) / 256), dim3(256), 0, L.stream, cellCount, 0u, totalCells);
    hipLaunchKernelGGL(k_insert, dim3(L.tiles), dim3(kTile), 0, L.stream, posIn, L.P, G, L.tileScene, cellOf, rank,
                       cellCount);
    exclusive_scan(cellCount, cellStart, totalCells, scanScratch, L.stream);
    hipLaunchKernelGGL(k_scatter_ids, dim3(L.tiles), dim3(kTile), 0, L.stream, L.P, G, L.tileScene,
                       (const uint32_t*)cellOf, (const uint32_t*)rank, (const uint32_t*)cellStart, idIn, tmpId);
    hipLaunchKernelGGL(k_reorder, dim3(L.tiles), dim3(kTile), 0, L.stream, L.P, G, L.tileScene,
                       (const uint32_t*)cellOf, (const uint32_t*)cellStart, (const uint32_t*)cellCount,
                       (const uint32_t*)tmpId, posIn, velIn, dvIn, dkIn, idIn, posOut, velOut, dvOut, dkOut, idOut);
}

__global__ void k_fill_u64(unsigned long long* p, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

// ---- halos of the neighbour groups ------------------------------------------------------------------------------
// One block per group. The distinct particle slots among the lists of its kGroupSize particles are collected in an
// LDS hash set of 32-slot buckets (bucket = slot >> 5, Fibonacci-hashed, linear probing) with a presence mask each;
// halo numbers are handed out bucket by bucket, ascending inside a bucket, so runs of consecutive particle slots stay
// consecutive in the halo. Every list entry is then rewritten as its halo number, packed with the pair's
// kernel-table code (gcode: the table index GetGradientW reads, or kGCodeZero beyond the cutoff) in the upper 16
// bits: positions do not change between here and the step's last neighbour loop, so the loops read the code instead
// of recomputing the distance, its square root and the index in every pass. A group with more than `limit` distinct
// neighbours (or more than kBucketLimit buckets) gets kHaloNone and is served by the global lists.
// haloStats = {groups staged, groups gathered, largest halo}.
constexpr uint32_t kBucketSlots = 2048;
constexpr uint32_t kBucketLimit = 896;  // + one bucket per thread in flight < kBucketSlots: probing always ends
constexpr uint32_t kHashEmpty = 0xFFFFFFFFu;
static_assert(kBucketLimit + kGroupSize < kBucketSlots, "bucket table must never fill");

__device__ __forceinline__ uint32_t bucket_hash(uint32_t b) { return (b * 0x9E3779B1u) >> 21; }  // 11 bits
constexpr size_t kLdsHaloPos = (size_t)kHaloMax * 3 * sizeof(float);  // halo positions of k_build_halo

__global__ void __launch_bounds__(kGroupSize) k_build_halo(const SceneParams* __restrict__ P,
                                                           const uint32_t* __restrict__ tileScene, uint32_t tiles,
                                                           const uint32_t* __restrict__ nbrCount,
                                                           const uint32_t* __restrict__ nbrList, uint32_t cap,
                                                           uint32_t limit, uint32_t* haloCount, uint32_t* haloIds,
                                                           uint32_t* nbrLocal, uint32_t* haloStats,
                                                           const float4* __restrict__ pos, float R, float invStep) {
    __shared__ uint32_t bkey[kBucketSlots];
    __shared__ uint32_t bmask[kBucketSlots];
    __shared__ uint16_t bbase[kBucketSlots];
    __shared__ uint32_t waveTotal[kGroupSize / kWave];
    __shared__ uint32_t used, rows;
    extern __shared__ float hpos[];  // [kHaloMax * 3]: positions of the halo, by halo number
    const uint32_t g = blockIdx.x;
    const uint32_t tile = g * kGroupTiles + threadIdx.x / kTile;
    const uint32_t i = g * kGroupSize + threadIdx.x;
    uint32_t n = 0;
    if (tile < tiles) {
        const SceneParams& p = P[tileScene[tile]];
        if (i < p.base + p.count) n = nbrCount[i];
    }
    for (uint32_t k = threadIdx.x; k < kBucketSlots; k += blockDim.x) {
        bkey[k] = kHashEmpty;
        bmask[k] = 0u;
    }
    if (threadIdx.x == 0) {
        used = 0;
        rows = 0;
    }
    __syncthreads();
    if (n) atomicMax(&rows, n);
    __syncthreads();
    const uint32_t nr = rows;
    volatile uint32_t* vkey = bkey;
    volatile uint32_t* vmask = bmask;
    volatile uint32_t* vused = &used;
    // the list entries are loaded kHaloBatch at a time (their latencies overlap), then inserted in list order
    constexpr uint32_t kHaloBatch = 8;
    for (uint32_t j0 = 0; j0 < nr; j0 += kHaloBatch) {
        uint32_t tb[kHaloBatch];
#pragma unroll
        for (uint32_t u = 0; u < kHaloBatch; ++u) tb[u] = j0 + u < n ? nbrList[(size_t)(j0 + u) * cap + i] : 0u;
#pragma unroll
        for (uint32_t u = 0; u < kHaloBatch; ++u) {
            if (j0 + u >= n || *vused > kBucketLimit) continue;
            const uint32_t t = tb[u];
            const uint32_t b = t >> 5, bit = 1u << (t & 31u);
            uint32_t h = bucket_hash(b);
            for (;;) {
                uint32_t cur = vkey[h];
                if (cur == kHashEmpty) {
                    cur = atomicCAS(&bkey[h], kHashEmpty, b);
                    if (cur == kHashEmpty) {
                        atomicAdd(&used, 1u);
                        cur = b;
                    }
                }
                if (cur == b) {
                    if (!(vmask[h] & bit)) atomicOr(&bmask[h], bit);
                    break;
                }
                h = (h + 1) & (kBucketSlots - 1);
            }
        }
    }
    __syncthreads();
    // halo numbers: thread t owns bucket slots [2t, 2t + 2)
    constexpr uint32_t kPer = kBucketSlots / kGroupSize;
    uint32_t own = 0;
#pragma unroll
    for (uint32_t k = 0; k < kPer; ++k) own += __popc(bmask[kPer * threadIdx.x + k]);
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    uint32_t incl = own;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const uint32_t v = __shfl_up(incl, o, kWave);
        if (lane >= o) incl += v;
    }
    if (lane == kWave - 1) waveTotal[w] = incl;
    __syncthreads();
    uint32_t m = 0;
    for (int k = 0; k < (int)(kGroupSize / kWave); ++k) m += waveTotal[k];
    if (used > kBucketLimit || m > limit) {  // block-uniform
        if (threadIdx.x == 0) {
            haloCount[g] = kHaloNone;
            atomicAdd(&haloStats[1], 1u);
        }
        return;
    }
    uint32_t base = incl - own;
    for (int k = 0; k < w; ++k) base += waveTotal[k];
    uint32_t* ids = haloIds + (size_t)g * kHaloMax;
#pragma unroll
    for (uint32_t k = 0; k < kPer; ++k) {
        const uint32_t slot = kPer * threadIdx.x + k;
        uint32_t mask = bmask[slot];
        bbase[slot] = (uint16_t)base;
        const uint32_t b0 = bkey[slot] << 5;
        while (mask) {
            const uint32_t bit = (uint32_t)__ffs(mask) - 1u;
            ids[base++] = b0 + bit;
            mask &= mask - 1u;
        }
    }
    if (threadIdx.x == 0) {
        haloCount[g] = m;
        if (m) atomicAdd(&haloStats[0], 1u);
        atomicMax(&haloStats[2], m);
    }
    __syncthreads();
    // the halo's positions, by halo number (ids written above by the block, visible after the barrier)
    if (m) {  // up to four entries per thread (kHaloMax <= 4 kGroupSize): all id loads, then all gathers
        uint32_t t[4];
        float4 x[4];
#pragma unroll
        for (uint32_t u = 0; u < 4; ++u) t[u] = ids[min(threadIdx.x + u * kGroupSize, m - 1u)];
#pragma unroll
        for (uint32_t u = 0; u < 4; ++u) x[u] = pos[t[u]];
#pragma unroll
        for (uint32_t u = 0; u < 4; ++u) {
            const uint32_t k = threadIdx.x + u * kGroupSize;
            if (k < m) {
                hpos[3 * k] = x[u].x;
                hpos[3 * k + 1] = x[u].y;
                hpos[3 * k + 2] = x[u].z;
            }
        }
    }
    __syncthreads();
    // staged entries: the halo number of each neighbour | its table code << 16, in list order (the neighbour's
    // position from the halo's copy in LDS), rows in pairs ([j/2][cap][2]: one 8-byte store per row pair; the entry
    // after a list's last one is 0 and never read)
    static_assert(kHaloBatch % 2 == 0 && kMaxNeighbors % 2 == 0, "staged rows are stored in pairs");
    const f3 xi = n ? ld3(pos[i]) : mk3(0.0f, 0.0f, 0.0f);
    for (uint32_t j0 = 0; j0 < nr; j0 += kHaloBatch) {
        uint32_t tb[kHaloBatch];
        // the list's last read in the step: non-temporal loads (it would only displace cached positions)
#pragma unroll
        for (uint32_t u = 0; u < kHaloBatch; ++u)
            tb[u] = j0 + u < n ? __builtin_nontemporal_load(nbrList + (size_t)(j0 + u) * cap + i) : 0u;
        uint32_t en[kHaloBatch];
#pragma unroll
        for (uint32_t u = 0; u < kHaloBatch; ++u) {
            en[u] = 0u;
            if (j0 + u >= n) continue;
            const uint32_t t = tb[u];
            const uint32_t b = t >> 5, below = (1u << (t & 31u)) - 1u;
            uint32_t h = bucket_hash(b);
            while (bkey[h] != b) h = (h + 1) & (kBucketSlots - 1);
            const uint32_t s = bbase[h] + __popc(bmask[h] & below);
            const uint32_t code = gcode<false>(R, invStep, xi - mk3(hpos[3 * s], hpos[3 * s + 1], hpos[3 * s + 2]));
            en[u] = s | (code << 16);
        }
        uint2* out = reinterpret_cast<uint2*>(nbrLocal) + (size_t)(j0 >> 1) * cap + i;
#pragma unroll
        for (uint32_t u = 0; u < kHaloBatch; u += 2)
            if (j0 + u < n) out[(size_t)(u >> 1) * cap] = make_uint2(en[u], en[u + 1]);
    }
}

void launch_build_halo(const Launch& L, const uint32_t* nbrCount, const uint32_t* nbrList, uint32_t haloLimit,
                       uint32_t* haloCount, uint32_t* haloIds, uint32_t* nbrLocal, uint32_t* haloStats,
                       const float4* pos, const KTabHost& kt) {
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, L.stream, haloStats, 0u, 3u);
    hipLaunchKernelGGL(k_build_halo, dim3(L.groups), dim3(kGroupSize), kLdsHaloPos, L.stream, L.P, L.tileScene, L.tiles,
                       nbrCount, nbrList, L.capacity, haloLimit < kHaloMax ? haloLimit : kHaloMax, haloCount, haloIds,
                       nbrLocal, haloStats, pos, kt.R, kt.invStep);
}

void launch_build_lists(const Launch& L, const SceneGrid* G, const float4* pos, const uint32_t* cellStart,
                        const uint32_t* cellCount, uint32_t* nbrCount, uint32_t* nbrList,
                        unsigned long long* nbStats) {
    hipLaunchKernelGGL(k_fill_u64, dim3((2 * L.sceneCount + 63) / 64), dim3(64), 0, L.stream, nbStats,
                       2 * L.sceneCount);
    if (VFD_SEARCH_CELLS)
        hipLaunchKernelGGL(k_build_lists_cells, dim3(L.gridAll), dim3(kTile), 0, L.stream, pos, L.P, G, L.tileScene,
                           L.tiles, L.tg.per, cellStart, cellCount, nbrCount, nbrList, L.capacity, nbStats);
    else
        hipLaunchKernelGGL(k_build_lists, dim3(L.tiles), dim3(kTile), kLdsListStage, L.stream, pos, L.P, G,
                           L.tileScene, cellStart, cellCount, nbrCount, nbrList, L.capacity, nbStats);
}

int configure_search_kernels() {
    if (hipFuncSetAttribute((const void*)k_build_halo, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)kLdsHaloPos) != hipSuccess)
        return 1;
    if (kLdsListStage == 0) return 0;
    return hipFuncSetAttribute((const void*)k_build_lists, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)kLdsListStage) == hipSuccess ? 0 : 1;
}

// ---- static scratch of the allocation-free step -------------------------------------------------------------------
// The grid cell arrays (counts, starts, scan partials) and the surface-tension curvature tables of vfd_step_no_alloc
// live here instead of in buffers the simulation allocates. Each device holds its own copy for the life of the
// process; every no-alloc step on a device shares it (overlapping calls on one device race). Its address is looked up
// once per device with that device current, and only that address is handed to kernels and copies: in host code
// `scratch` names a host-side stand-in, not device memory.
static __device__ __align__(256) uint32_t scratch[kStaticScratchWords];

namespace {
std::mutex g_scratchMu;
uint32_t* g_scratchPtr[kMaxStaticScratchDevices] = {};
}  // namespace

int static_scratch_address(int device, uint32_t** out) {
    *out = nullptr;
    if (device < 0 || device >= (int)kMaxStaticScratchDevices) return 1;
    std::lock_guard<std::mutex> lk(g_scratchMu);
    if (!g_scratchPtr[device]) {
        int cur = -1;
        if (hipGetDevice(&cur) != hipSuccess || cur != device) return 1;
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || p == nullptr) {
            (void)hipGetLastError();
            return 1;
        }
        g_scratchPtr[device] = static_cast<uint32_t*>(p);
    }
    *out = g_scratchPtr[device];
    return 0;
}

}  // namespace vfd
