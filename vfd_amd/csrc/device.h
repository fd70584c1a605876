// device.h - device-side helpers shared by the HIP kernels (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <float.h>
#include <limits.h>

#include "common.h"

namespace vfd {

constexpr int kWave = 64;

// ---- vec3 arithmetic on float4 storage (w ignored) -------------------------------------------------------------
struct f3 {
    float x, y, z;
};
__device__ __forceinline__ f3 mk3(float x, float y, float z) { return f3{x, y, z}; }
__device__ __forceinline__ f3 ld3(const float4& v) { return f3{v.x, v.y, v.z}; }
__device__ __forceinline__ float4 st4(f3 v) { return make_float4(v.x, v.y, v.z, 0.0f); }
__device__ __forceinline__ f3 operator+(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ f3 operator-(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ f3 operator-(f3 a) { return {-a.x, -a.y, -a.z}; }
__device__ __forceinline__ f3 operator*(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ f3 operator*(float s, f3 a) { return {s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ f3 operator/(f3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ f3 cmul(f3 a, f3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
// a + s*b with explicit fused multiply-adds (used in accumulations)
__device__ __forceinline__ f3 fma3(float s, f3 b, f3 a) {
    return {fmaf(s, b.x, a.x), fmaf(s, b.y, a.y), fmaf(s, b.z, a.z)};
}
__device__ __forceinline__ float dot3(f3 a, f3 b) { return fmaf(a.z, b.z, fmaf(a.y, b.y, a.x * b.x)); }
__device__ __forceinline__ float length2(f3 a) { return dot3(a, a); }
__device__ __forceinline__ float length(f3 a) { return sqrtf(dot3(a, a)); }
__device__ __forceinline__ bool is_zero(f3 a) { return a.x == 0.0f && a.y == 0.0f && a.z == 0.0f; }
__device__ __forceinline__ f3 cross(f3 a, f3 b) {
    return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
__device__ __forceinline__ f3 normalize(f3 a) { return a * (1.0f / sqrtf(dot3(a, a))); }

// |x| > EPS for a float x, where the reference compares in double against the double literal 1e-5
// (Core/Math/Math.h:6): (double)x > 1e-5 holds exactly when x >= the smallest float above 1e-5.
constexpr float kEpsAbove = 1.00000008e-05f;  // 0x3727c5ad: nextafterf((float)1e-5, inf); (float)1e-5 < 1e-5
__device__ __forceinline__ bool above_eps(float ax) { return ax >= kEpsAbove; }

// ---- kernel table ----------------------------------------------------------------------------------------------
// tab[p] = {0.5*(W[p]+W[p+1]), 0.5*(G[p]+G[p+1])}, p in [0, 9998], exactly the values GetW / GetGradientW form
// (Kernel/DFSPHKernels.h:43-80). G stores gradW(x)_x / x, so gradW(r) = G * r.
struct KTab {
    const float2* __restrict__ tab;
    float R, R2, invStep, W0;
};
__device__ __forceinline__ unsigned table_index(const KTab& k, float rl) {
    const unsigned p = (unsigned)(rl * k.invStep);
    return p < kTableEntries - 1u ? p : kTableEntries - 1u;
}
// One lookup serving both W (cutoff r2 <= R^2) and gradW (cutoff |r| <= R) for the pair offset r.
struct Pair {
    float W, G;  // G * r = gradW(r)
};
__device__ __forceinline__ Pair pair_lookup(const KTab& k, f3 r) {
    const float r2 = dot3(r, r);
    const float rl = sqrtf(r2);
    const float2 t = k.tab[table_index(k, rl)];
    Pair o;
    o.W = r2 <= k.R2 ? t.x : 0.0f;
    o.G = rl <= k.R ? t.y : 0.0f;
    return o;
}
__device__ __forceinline__ f3 gradW(const KTab& k, f3 r) {
    const float rl = sqrtf(dot3(r, r));
    if (rl <= k.R) return k.tab[table_index(k, rl)].y * r;
    return mk3(0.0f, 0.0f, 0.0f);
}
__device__ __forceinline__ float getW(const KTab& k, f3 r) {
    const float r2 = dot3(r, r);
    if (r2 <= k.R2) return k.tab[table_index(k, sqrtf(r2))].x;
    return 0.0f;
}
__device__ __forceinline__ float getW_scalar(const KTab& k, float r) {
    if (r <= k.R) return k.tab[table_index(k, r)].x;
    return 0.0f;
}

// ---- wave / block reductions (deterministic order) -------------------------------------------------------------
// A 64-bit lane value moved by the DPP pattern kCtrl (two 32-bit moves); lanes of rows outside kRowMask read 0.
template <int kCtrl, int kRowMask = 0xF>
__device__ __forceinline__ double dpp_f64(double v) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)b, kCtrl, kRowMask, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(b >> 32), kCtrl, kRowMask, 0xF, false);
    return __longlong_as_double((long long)(((unsigned long long)(uint32_t)hi << 32) | (uint32_t)lo));
}
// Sum over the wave's 64 lanes through DPP lane moves (no LDS), the same value in every lane; called with every lane
// active. Within each row of 16 lanes: quad_perm [1,0,3,2], quad_perm [2,3,0,1], row_half_mirror, row_mirror (each
// lane then holds its row's sum: the pairs added are the same in both lanes, so every lane has the same bits); then
// row_bcast:15 adds row 0 into row 1 and row 2 into row 3, row_bcast:31 adds rows 0-1 into row 3, and lane 63 holds
// the wave's sum. A fixed tree, so the per-tile partial sums are deterministic.
__device__ __forceinline__ double wave_sum(double v) {
    v += dpp_f64<0xB1>(v);        // quad_perm [1,0,3,2]
    v += dpp_f64<0x4E>(v);        // quad_perm [2,3,0,1]
    v += dpp_f64<0x141>(v);       // row_half_mirror
    v += dpp_f64<0x140>(v);       // row_mirror
    v += dpp_f64<0x142, 0xA>(v);  // row_bcast:15 into rows 1 and 3
    v += dpp_f64<0x143, 0xC>(v);  // row_bcast:31 into rows 2 and 3
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, kWave - 1);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), kWave - 1);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, kWave));
    return v;
}
__device__ __forceinline__ int wave_min_i(int v) {
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, kWave));
    return v;
}
__device__ __forceinline__ int wave_max_i(int v) {
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, kWave));
    return v;
}

// Block (kThreads threads) sum of NV doubles; result valid in thread 0. Fixed order: wave tree, then the waves in
// order.
template <int NV, int kThreads = kTile>
__device__ __forceinline__ void block_sum(double (&v)[NV]) {
    __shared__ double sh[kThreads / kWave][NV];
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
#pragma unroll
    for (int i = 0; i < NV; ++i) v[i] = wave_sum(v[i]);
    if (lane == 0)
#pragma unroll
        for (int i = 0; i < NV; ++i) sh[w][i] = v[i];
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            double s = sh[0][i];
            for (int j = 1; j < kThreads / kWave; ++j) s += sh[j][i];
            v[i] = s;
        }
    }
}

// Per-tile partial sums: the block reduces NV doubles in a fixed order and thread 0 stores them at
// partials[tile * kPartialSlots + i]. A separate finalize kernel (one block per scene) adds a scene's tiles in tile
// order, so every reduction is deterministic and needs no device-scope fence (on the multi-XCD MI355X an agent-scope
// release writes back the L2; one per block serialises thousands of them).
template <int NV>
__device__ __forceinline__ void tile_partials(double (&v)[NV], double* partials, uint32_t tile) {
    block_sum<NV>(v);
    if (threadIdx.x == 0)
#pragma unroll
        for (int i = 0; i < NV; ++i) partials[(size_t)tile * kPartialSlots + i] = v[i];
    __syncthreads();  // the block-sum scratch is reused by the next tile of a persistent block
}

// Inside a finalize block (kThreads threads) for scene s: sum of the scene's tile partials, valid in thread 0. Each
// thread adds the tiles threadIdx.x, threadIdx.x + kThreads, ... in that order; the loads of eight of them are issued
// together before they are added, so the latency is paid once per eight (once in all for a 1024-thread block and up
// to 8192 tiles, 2M particles).
template <int NV, int kThreads = kTile>
__device__ __forceinline__ void scene_sum(const double* __restrict__ partials, uint32_t tileBegin, uint32_t tileEnd,
                                          double (&out)[NV]) {
    constexpr uint32_t kBatch = 8;
#pragma unroll
    for (int i = 0; i < NV; ++i) out[i] = 0.0;
    for (uint32_t t0 = tileBegin + threadIdx.x; t0 < tileEnd; t0 += kBatch * kThreads) {
        double v[kBatch][NV];
#pragma unroll
        for (uint32_t u = 0; u < kBatch; ++u) {
            const uint32_t t = t0 + u * kThreads;
#pragma unroll
            for (int i = 0; i < NV; ++i) v[u][i] = t < tileEnd ? partials[(size_t)t * kPartialSlots + i] : 0.0;
        }
#pragma unroll
        for (uint32_t u = 0; u < kBatch; ++u)
            if (t0 + u * kThreads < tileEnd)
#pragma unroll
                for (int i = 0; i < NV; ++i) out[i] += v[u][i];
    }
    block_sum<NV, kThreads>(out);
}

__device__ __forceinline__ float block_max(float v) {
    __shared__ float sh[kTile / kWave];
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    v = wave_max(v);
    if (lane == 0) sh[w] = v;
    __syncthreads();
    float m = sh[0];
    for (int j = 1; j < (int)(kTile / kWave); ++j) m = fmaxf(m, sh[j]);
    __syncthreads();
    return m;
}

// Tile bookkeeping shared by every per-particle kernel.
struct TileCtx {
    uint32_t s;     // scene
    uint32_t i;     // storage slot
    bool valid;     // slot holds a particle
};
__device__ __forceinline__ TileCtx tile_ctx(const uint32_t* tileScene, const SceneParams* P) {
    TileCtx c;
    c.s = tileScene[blockIdx.x];
    c.i = blockIdx.x * kTile + threadIdx.x;
    c.valid = c.i < P[c.s].base + P[c.s].count;
    return c;
}
__device__ __forceinline__ TileCtx tile_ctx(uint32_t tile, const uint32_t* tileScene, const SceneParams* P) {
    TileCtx c;
    c.s = tileScene[tile];
    c.i = tile * kTile + threadIdx.x;
    c.valid = c.i < P[c.s].base + P[c.s].count;
    return c;
}

// XCD-aware tile loop. The tiles (Morton-ordered particle ranges) are cut into 8 contiguous chunks of `per` tiles,
// one per XCD; work-group q runs on XCD q % 8 (round-robin dispatch), so chunk q % 8 stays in one XCD's L2. With a
// grid smaller than 8 * per the blocks loop (persistent); the grid must be a multiple of 8.
struct TileGrid {
    uint32_t tiles, per;
    __host__ __device__ TileGrid() = default;
    template <class T> __host__ __device__ TileGrid(const T& t) : tiles(t.tiles), per(t.per) {}
};
template <class Body>
__device__ __forceinline__ void tile_loop(const TileGrid& tg, Body&& body) {
    for (uint32_t q = blockIdx.x; q < 8u * tg.per; q += gridDim.x) {
        const uint32_t k = q >> 3, tile = (q & 7u) * tg.per + k;
        if (tile < tg.tiles) body(tile);
    }
}

// ---- neighbour groups (kGroupTiles tiles per kGroupSize-thread block) ---------------------------------------------
// Thread group sub = threadIdx.x / kTile of a block works on tile kGroupTiles * g + sub of group g.
__device__ __forceinline__ TileCtx tile_ctx4(uint32_t tile, bool ok, const uint32_t* tileScene,
                                             const SceneParams* P) {
    TileCtx c;
    c.s = ok ? tileScene[tile] : 0u;
    c.i = tile * kTile + (threadIdx.x & (kTile - 1));
    c.valid = ok && c.i < P[c.s].base + P[c.s].count;
    return c;
}
// Per-tile sums inside a kGroupSize-thread block: wave sums, then the tile's four waves in order (the same association
// as block_sum), stored by the tile's first thread when `store`. Called by every thread.
template <int NV>
__device__ __forceinline__ void tile_partials4(double (&v)[NV], double* partials, uint32_t tile, bool store) {
    __shared__ double sh[kGroupSize / kWave][NV];
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
#pragma unroll
    for (int i = 0; i < NV; ++i) v[i] = wave_sum(v[i]);
    if (lane == 0)
#pragma unroll
        for (int i = 0; i < NV; ++i) sh[w][i] = v[i];
    __syncthreads();
    if ((threadIdx.x & (kTile - 1)) == 0 && store) {
        const int w0 = (threadIdx.x / kTile) * (kTile / kWave);
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            double s = sh[w0][i];
            for (int j = 1; j < (int)(kTile / kWave); ++j) s += sh[w0 + j][i];
            partials[(size_t)tile * kPartialSlots + i] = s;
        }
    }
    __syncthreads();
}

// Neighbour j of particle i in the transposed lists; f(t) is called in list order, four list entries fetched per
// trip so that their gathers can be in flight together.
template <class Fn>
__device__ __forceinline__ void each_neighbor(const uint32_t* __restrict__ list, uint32_t cap, uint32_t i, uint32_t n,
                                              Fn&& f) {
    uint32_t j = 0;
    for (; j + 4 <= n; j += 4) {
        const uint32_t t0 = list[(size_t)j * cap + i], t1 = list[(size_t)(j + 1) * cap + i];
        const uint32_t t2 = list[(size_t)(j + 2) * cap + i], t3 = list[(size_t)(j + 3) * cap + i];
        f(t0);
        f(t1);
        f(t2);
        f(t3);
    }
    for (; j < n; ++j) f(list[(size_t)j * cap + i]);
}

// Kernel table staged in LDS: tab[p] (float2 {W, G}) or G only, loaded once per block. The sizes are rounded up to
// 16 bytes so that the halo records can follow the table.
// both LDS tables hold one entry past the kernel table, for code kGCodeZero (see gcode)
constexpr unsigned kLdsG = ((kTableEntries + 1u) * sizeof(float) + 15u) / 16u * 16u;
constexpr unsigned kLdsWG = ((kTableEntries + 1u) * sizeof(float2) + 15u) / 16u * 16u;
// Each thread's loads are all issued before its first LDS store (kTableBatch per round: one round for a 1024-thread
// block), so the copy pays one memory latency per launch instead of one per entry it handles.
constexpr uint32_t kTableBatch = 16;
__device__ __forceinline__ void load_table_g(float* sG, const float2* __restrict__ tab) {
    for (uint32_t p0 = threadIdx.x; p0 < kTableEntries; p0 += kTableBatch * blockDim.x) {
        float v[kTableBatch];
#pragma unroll
        for (uint32_t u = 0; u < kTableBatch; ++u) {
            const uint32_t p = p0 + u * blockDim.x;
            v[u] = tab[min(p, kTableEntries - 1u)].y;  // clamped: every load in bounds, none behind a branch
        }
#pragma unroll
        for (uint32_t u = 0; u < kTableBatch; ++u) {
            const uint32_t p = p0 + u * blockDim.x;
            if (p < kTableEntries) sG[p] = v[u];
        }
    }
    if (threadIdx.x == 0) sG[kTableEntries] = 0.0f;  // code kGCodeZero: beyond the gradient cutoff
    __syncthreads();
}
// {W, G} pairs; entry kGCodeZero = {W of the last entry (what GetW reads beyond R), 0}
__device__ __forceinline__ void load_table_wg(float2* sWG, const float2* __restrict__ tab) {
    for (uint32_t p0 = threadIdx.x; p0 < kTableEntries; p0 += kTableBatch * blockDim.x) {
        float2 v[kTableBatch];
#pragma unroll
        for (uint32_t u = 0; u < kTableBatch; ++u) {
            const uint32_t p = p0 + u * blockDim.x;
            v[u] = tab[min(p, kTableEntries - 1u)];
        }
#pragma unroll
        for (uint32_t u = 0; u < kTableBatch; ++u) {
            const uint32_t p = p0 + u * blockDim.x;
            if (p < kTableEntries) sWG[p] = v[u];
        }
    }
    if (threadIdx.x == 0) sWG[kTableEntries] = make_float2(tab[kTableEntries - 1u].x, 0.0f);
    __syncthreads();
}
// gradW(r) with the G table in LDS (cutoff |r| <= R), same values as gradW(KTab, r)
struct KLds {
    const float* sG;  // kTableEntries + 1 entries, the last one 0
    float R, invStep;
};
// Correctly rounded square root (the value of IEEE sqrtf) in fewer instructions for x >= 2^-64: the hardware
// estimate s (within 1 ulp) and its two neighbours s-, s+ (one ulp apart); the root rounds to s- when x - s- s <= 0
// and to s+ when x - s+ s > 0. Both residuals are exact-sign single-rounding fmas, and x - s- s is a multiple of
// ulp(s)^2, so it cannot fall strictly between 0 and the midpoint gap ulp(s)^2 / 4 - the tests decide the rounding
// exactly. Above 2^-64 the residuals stay normal numbers. A smaller input is scaled by 2^126 first (exact, and the
// scaled value is at least 2^-23) and the root by 2^-63 after (exact: the root is a normal number), which commutes
// with the rounding; zero gives zero and NaN gives NaN. No branch, so that the loops can issue the LDS reads of a
// whole chunk before the first root.
// The neighbour loops use it where it pays: the loops of one-record-and-scalar and viscosity records (LoadXS,
// LoadXSV: 5% and 1% faster on the 2M dam break); the two-record and density loops were 2-3% slower with it and keep
// sqrtf (both give the same value).
#ifndef VFD_FAST_SQRT
#define VFD_FAST_SQRT 1
#endif
__device__ __forceinline__ float sqrt_rn(float x) {
#if VFD_FAST_SQRT
    const bool small = x < 0x1.0p-64f;
    const float xs = small ? x * 0x1.0p126f : x;
    const float s = __builtin_amdgcn_sqrtf(xs);
    const float sd = __uint_as_float(__float_as_uint(s) - 1u), su = __uint_as_float(__float_as_uint(s) + 1u);
    float r = fmaf(-sd, s, xs) <= 0.0f ? sd : s;
    r = fmaf(-su, s, xs) > 0.0f ? su : r;
    return small ? r * 0x1.0p-63f : r;
#else
    return sqrtf(x);
#endif
}

// Table code of a pair offset r for the gradient: the table index GetGradientW reads, or kGCodeZero beyond its
// cutoff |r| <= R (Kernel/DFSPHKernels.h:67-80). G[code] * r is gradW(r), up to the sign of a zero component.
constexpr uint32_t kGCodeZero = kTableEntries;
template <bool kShortSqrt = false>
__device__ __forceinline__ uint32_t gcode(float R, float invStep, f3 r) {
    const float r2 = dot3(r, r);
    const float rl = kShortSqrt ? sqrt_rn(r2) : sqrtf(r2);
    const unsigned p0 = (unsigned)(rl * invStep);
    const unsigned p = p0 < kTableEntries - 1u ? p0 : kTableEntries - 1u;
    return rl <= R ? p : kGCodeZero;
}
__device__ __forceinline__ f3 gradW(const KLds& k, f3 r) {
    const float rl = sqrtf(dot3(r, r));
    if (rl <= k.R) {
        const unsigned p0 = (unsigned)(rl * k.invStep);
        const unsigned p = p0 < kTableEntries - 1u ? p0 : kTableEntries - 1u;
        return k.sG[p] * r;
    }
    return mk3(0.0f, 0.0f, 0.0f);
}

// ---- neighbour sources: halo records in LDS or gathers from the global arrays -----------------------------------
// A neighbour-loop kernel sees each neighbour t of particle i as one or two float4 records {lo, hi} that a loader
// assembles from the global arrays (e.g. lo = {x_t, kappa_t}, hi = {a_t}). For a staged group the records of its
// halo are copied to LDS once and the lists are read as 16-bit halo slots; otherwise the loader gathers them per
// list entry. Both paths hand the same values to the same arithmetic in the same list order.
struct NbrSrc {
    const float* rec;                      // LDS records of the halo (layout: HaloRec); nullptr: global gathers
    const uint32_t* __restrict__ local;    // [j * cap + i] halo slot | table code << 16 (staged groups)
    const uint32_t* __restrict__ list;     // [j * cap + i] particle slots
    const uint32_t* __restrict__ count;    // [i]
    uint32_t cap;
};

struct LoadX {  // lo = hi = {x_t, -}
    static constexpr int kLo = 3, kHi = 0;  // floats of lo and hi that the neighbour terms read
    static constexpr bool kShortSqrt = false;  // the table code's square root: sqrt_rn (true) or sqrtf
    const float4* __restrict__ pos;
    __device__ __forceinline__ void operator()(uint32_t t, float4& lo, float4& hi) const {
        lo = pos[t];
        hi = lo;
    }
};
struct LoadXS {  // lo = hi = {x_t, s_t}
    static constexpr int kLo = 4, kHi = 0;
    static constexpr bool kShortSqrt = true;
    const float4* __restrict__ pos;
    const float* __restrict__ s;
    __device__ __forceinline__ void operator()(uint32_t t, float4& lo, float4& hi) const {
        const float4 x = pos[t];
        lo = make_float4(x.x, x.y, x.z, s[t]);
        hi = lo;
    }
};
struct LoadXV {  // lo = {x_t, -}, hi = {v_t, -}
    static constexpr int kLo = 3, kHi = 3;
    static constexpr bool kShortSqrt = false;
    const float4* __restrict__ pos;
    const float4* __restrict__ v;
    __device__ __forceinline__ void operator()(uint32_t t, float4& lo, float4& hi) const {
        lo = pos[t];
        hi = v[t];
    }
};
// Viscosity records: lo = {x_t, m / rho_t} (hi = {v_t, -}). The division is the one the neighbour terms take per pair
// (m / rho_j, DFSPHKernels.cu:638, :790), formed once per halo record instead; m is the same for every scene of a
// batch (all scenes share the particle radius).
struct LoadXQ {
    static constexpr int kLo = 4, kHi = 0;
    static constexpr bool kShortSqrt = true;
    const float4* __restrict__ pos;
    const float* __restrict__ rho;
    float m;
    __device__ __forceinline__ void operator()(uint32_t t, float4& lo, float4& hi) const {
        const float4 x = pos[t];
        lo = make_float4(x.x, x.y, x.z, m / rho[t]);
        hi = lo;
    }
};
struct LoadXQV {
    static constexpr int kLo = 4, kHi = 3;
    static constexpr bool kShortSqrt = true;
    const float4* __restrict__ pos;
    const float* __restrict__ rho;
    const float4* __restrict__ v;
    float m;
    __device__ __forceinline__ void operator()(uint32_t t, float4& lo, float4& hi) const {
        const float4 x = pos[t];
        lo = make_float4(x.x, x.y, x.z, m / rho[t]);
        hi = v[t];
    }
};
struct LoadXSV {  // lo = {x_t, s_t}, hi = {v_t, -}
    static constexpr int kLo = 4, kHi = 3;
    static constexpr bool kShortSqrt = true;
    const float4* __restrict__ pos;
    const float* __restrict__ s;
    const float4* __restrict__ v;
    __device__ __forceinline__ void operator()(uint32_t t, float4& lo, float4& hi) const {
        const float4 x = pos[t];
        lo = make_float4(x.x, x.y, x.z, s[t]);
        hi = v[t];
    }
};

// f(lo, hi, t) for every neighbour of particle i in list order, t = tab[code] and code = gcode(xi - x_j) (the
// kernel-table entry of the pair). Staged groups read their list as halo slot | code << 16 (4 bytes per neighbour;
// k_build_halo computes the code once per step) in chunks of ListShape<R>::chunk entries: the chunk's
// list loads are issued together (rows past the end repeat the last entry, so every load and LDS read stays in
// bounds), then all of its LDS record and table reads, and only then its f calls, so that neither the list latency
// nor the LDS latency is paid once per neighbour. With a lookahead the next chunks' list loads are issued before
// the current chunk is consumed. The global path fetches four list entries per trip so that their gathers can be
// in flight together, and computes the code itself.
// Chunk length and lookahead are chosen per record count R: one-record loops (16 B per neighbour) profit from long
// chunks; two-record loops hold twice the registers per entry and take short ones. VFD_LIST_PREFETCH(2) is the
// number of further chunks whose list loads are in flight while a chunk is consumed (0: none), since the list
// latency, not the LDS, bounds these loops at one 1024-thread block per CU.
#ifndef VFD_LIST_CHUNK
#define VFD_LIST_CHUNK 8
#endif
#ifndef VFD_LIST_PREFETCH
#define VFD_LIST_PREFETCH 1
#endif
#ifndef VFD_LIST_CHUNK2
#define VFD_LIST_CHUNK2 4
#endif
#ifndef VFD_LIST_PREFETCH2
#define VFD_LIST_PREFETCH2 1
#endif
template <int R> struct ListShape {
    static constexpr uint32_t chunk = R == 1 ? VFD_LIST_CHUNK : VFD_LIST_CHUNK2;
    static constexpr uint32_t depth = 1u + (R == 1 ? VFD_LIST_PREFETCH : VFD_LIST_PREFETCH2);  // chunks in flight
};
// Halo records in LDS, as two arrays of kHaloMax records: lo of slot s at rec[L * s], hi at rec[L * kHaloMax + H * s]
// (H = 0 for one-record loops). With VFD_PACK_RECORDS a half that the neighbour terms read as three floats is stored
// in 12 bytes (L or H = 3), which spreads a wave's random reads over all 32 banks instead of 8 groups of 4; otherwise
// every half is a 16-byte float4. VFD_PACK_RECORDS = 2 stores a record of two three-float halves as 16 + 8 bytes
// instead ({lo.xyz, hi.x} and {hi.y, hi.z}: L = 4, H = 2), read with one 16-byte and one 8-byte LDS load rather than
// two 12-byte halves of two loads each.
#ifndef VFD_PACK_RECORDS
#define VFD_PACK_RECORDS 1
#endif
#ifndef VFD_SCHED_PHASES
#define VFD_SCHED_PHASES 1
#endif
template <class Load> struct HaloRec {
    static constexpr bool kSplit = VFD_PACK_RECORDS == 2 && Load::kLo == 3 && Load::kHi == 3;
    static constexpr int L = kSplit ? 4 : (VFD_PACK_RECORDS ? Load::kLo : 4);
    static constexpr int H = Load::kHi == 0 ? 0 : (kSplit ? 2 : (VFD_PACK_RECORDS ? Load::kHi : 4));
    static constexpr size_t kBytes = (size_t)(L + H) * sizeof(float) * kHaloMax;  // LDS of the records
    static __device__ __forceinline__ float4 get(const float* p, int n) {
        if (n == 4) return *reinterpret_cast<const float4*>(p);
        return make_float4(p[0], p[1], p[2], 0.0f);
    }
    // both halves of slot s
    static __device__ __forceinline__ void load(const float* rec, uint32_t s, float4& lo, float4& hi) {
        if (kSplit) {
            lo = *reinterpret_cast<const float4*>(rec + 4 * s);
            const float2 t = *reinterpret_cast<const float2*>(rec + 4 * kHaloMax + 2 * s);
            hi = make_float4(lo.w, t.x, t.y, 0.0f);
        } else {
            lo = get(rec + L * s, L);
            hi = H ? get(rec + L * kHaloMax + H * s, H) : lo;
        }
    }
    static __device__ __forceinline__ void put(float* p, int n, const float4& v) {
        if (n == 4) {
            *reinterpret_cast<float4*>(p) = v;
        } else {
            p[0] = v.x;
            p[1] = v.y;
            p[2] = v.z;
        }
    }
    static __device__ __forceinline__ float4 lo(const float* rec, uint32_t s) { return get(rec + L * s, L); }
    static __device__ __forceinline__ float4 hi(const float* rec, uint32_t s) {
        return get(rec + L * kHaloMax + H * s, H);
    }
    static __device__ __forceinline__ void store(float* rec, uint32_t s, const float4& l, const float4& h) {
        if (kSplit) {
            *reinterpret_cast<float4*>(rec + 4 * s) = make_float4(l.x, l.y, l.z, h.x);
            *reinterpret_cast<float2*>(rec + 4 * kHaloMax + 2 * s) = make_float2(h.y, h.z);
            return;
        }
        put(rec + L * s, L, l);
        if (H) put(rec + L * kHaloMax + H * s, H, h);
    }
};
// Word 3 of a CDNA buffer resource for raw (stride 0) 32-bit data: data format 32, no swizzle, no index stride.
constexpr int kRawBufferWord3 = 0x00020000;
// An offset at or past a raw buffer's num_records reads 0 without touching memory; adding this to a lane's offset
// keeps the load in the instruction stream (so the compiler still counts it in flight) but takes it off the memory
// path. Valid offsets stay below 2^31 bytes (a simulation holds at most 2^28 slots: 8-byte row-pair entries).
constexpr uint32_t kBufferOutOfRange = 0x80000000u;
// Cache policy bits of a buffer load on gfx950: bit 1 = non-temporal.
constexpr int kBufferNonTemporal = 2;

template <int R, class TV, class Load, class Fn>
__device__ __forceinline__ void each_neighbor_rec(const NbrSrc& src, uint32_t i, float cutR, float invStep, f3 xi,
                                                  const TV* tab, const Load& load, Fn&& f) {
    constexpr uint32_t C = ListShape<R>::chunk;
    const uint32_t n = src.count[i];
    const size_t cap = src.cap;
    uint32_t j = 0;
    if (src.rec) {
        if (n == 0) return;
        const float* rec = src.rec;
        using HR = HaloRec<Load>;
        const uint32_t last = n - 1u;
        // the list is read once per launch: non-temporal loads keep it from displacing the halo records, which
        // neighbouring groups gather again, from the caches. The staged lists hold rows in pairs ([j/2][cap][2]):
        // rows r, r + 1 of the wave (r even, the same in every lane) are one buffer of cap 8-byte entries, read with
        // one load per lane; a lane whose list ended before row r reads past the buffer's end, which returns 0
        // without a memory access, so the rows past a short list cost the memory path nothing.
        static_assert(C % 2 == 0, "list chunks cover whole row pairs");
        auto fetch = [&](uint32_t (&e)[C], uint32_t b) {
#pragma unroll
            for (uint32_t u = 0; u < C / 2; ++u) {
                const uint32_t r = b + 2 * u;
                const __amdgpu_buffer_rsrc_t rows = __builtin_amdgcn_make_buffer_rsrc(
                    const_cast<uint32_t*>(src.local + (size_t)(r >> 1) * cap * 2), (short)0, (int)(cap * 8),
                    kRawBufferWord3);
                const uint32_t off = i * 8u + (r <= last ? 0u : kBufferOutOfRange);
                const auto v = __builtin_amdgcn_raw_buffer_load_b64(rows, (int)off, 0, kBufferNonTemporal);
                e[2 * u] = v[0];
                e[2 * u + 1] = v[1];
            }
        };
        auto consume = [&](const uint32_t (&e)[C], uint32_t b) {
            float4 lo[C], hi[C];
            TV tv[C];
            // every record read of the chunk is issued before the first root, every table read before the first
            // neighbour term (VFD_SCHED_PHASES; the arithmetic and its order are the same either way)
#pragma unroll
            for (uint32_t u = 0; u < C; ++u) {
                HR::load(rec, e[u] & 0xFFFFu, lo[u], hi[u]);
                tv[u] = tab[e[u] >> 16];  // the pair's table code, computed once per step by k_build_halo
            }
            if (VFD_SCHED_PHASES) __builtin_amdgcn_sched_barrier(0);
            if (b + C <= n) {
#pragma unroll
                for (uint32_t u = 0; u < C; ++u) f(lo[u], hi[u], tv[u]);
            } else {
#pragma unroll
                for (uint32_t u = 0; u < C; ++u)
                    if (b + u < n) f(lo[u], hi[u], tv[u]);
            }
        };
        // ring of D chunk buffers: chunk c lives in e[c % D]; chunk c + D - 1 is requested before chunk c is consumed.
        // The requests are unconditional (rows past a list's end are out-of-range reads that return 0): under a
        // per-lane condition the compiler cannot count the loads in flight and waits for all of them, the prefetch
        // included, before the current chunk's first use.
        constexpr uint32_t D = ListShape<R>::depth;
        uint32_t e[D][C];
#pragma unroll
        for (uint32_t k = 0; k + 1 < D; ++k) fetch(e[k], k * C);
        for (; j < n; j += D * C) {
#pragma unroll
            for (uint32_t k = 0; k < D; ++k) {
                const uint32_t b = j + k * C;
                fetch(e[(k + D - 1) % D], b + (D - 1) * C);
                if (b < n) consume(e[k], b);
            }
        }
    } else {
        for (; j + 4 <= n; j += 4) {
            uint32_t t[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) t[u] = src.list[(j + u) * cap + i];
            float4 lo[4], hi[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) load(t[u], lo[u], hi[u]);
#pragma unroll
            for (int u = 0; u < 4; ++u) f(lo[u], hi[u], tab[gcode<Load::kShortSqrt>(cutR, invStep, xi - ld3(lo[u]))]);
        }
        for (; j < n; ++j) {
            float4 lo, hi;
            load(src.list[j * cap + i], lo, hi);
            f(lo, hi, tab[gcode<Load::kShortSqrt>(cutR, invStep, xi - ld3(lo))]);
        }
    }
}

struct HaloView {
    const uint32_t* __restrict__ count;  // [groups]
    const uint32_t* __restrict__ ids;    // [groups][kHaloMax]
};

// Persistent, XCD-aware loop over groups: XCD x (= block index mod 8) takes groups [x * per/4, (x + 1) * per/4) of
// the tiles it owns, so that its L2 holds the data of one contiguous stretch of the Morton order. For a staged group
// every thread first copies halo records to LDS; body(tile, ok, src) is then called by every thread of the block
// (ok = false past the last tile) and may contain block barriers.
template <int R, class Load, class Body>
__device__ __forceinline__ void group_loop(const TileGrid& tg, const HaloView& H, float* rec, const NbrSrc& global,
                                           const Load& load, Body&& body) {
    const uint32_t sub = threadIdx.x / kTile;
    const uint32_t perG = tg.per / kGroupTiles;
    for (uint32_t q = blockIdx.x; q < 8u * perG; q += gridDim.x) {
        const uint32_t g = (q & 7u) * perG + (q >> 3);
        if (g * kGroupTiles >= tg.tiles) continue;  // block-uniform
        const uint32_t tile = g * kGroupTiles + sub;
        const uint32_t m = H.count[g];
        if (m <= kHaloMax) {
            // every thread copies up to four records: the four slot loads first (indices past the halo repeat its
            // last entry, so they stay in bounds), then the four record gathers, then the LDS stores
            static_assert(kHaloMax <= 4 * kGroupSize, "fill covers four records per thread");
            if (m) {
                const uint32_t* ids = H.ids + (size_t)g * kHaloMax;
                const uint32_t k0 = threadIdx.x, k1 = k0 + kGroupSize, k2 = k1 + kGroupSize, k3 = k2 + kGroupSize;
                const uint32_t mm = m - 1u;
                const uint32_t t0 = ids[min(k0, mm)], t1 = ids[min(k1, mm)];
                const uint32_t t2 = ids[min(k2, mm)], t3 = ids[min(k3, mm)];
                float4 lo0, hi0, lo1, hi1, lo2, hi2, lo3, hi3;
                load(t0, lo0, hi0);
                if (k1 < m) load(t1, lo1, hi1);
                if (k2 < m) load(t2, lo2, hi2);
                if (k3 < m) load(t3, lo3, hi3);
                auto put = [&](uint32_t k, const float4& lo, const float4& hi) { HaloRec<Load>::store(rec, k, lo, hi); };
                if (k0 < m) put(k0, lo0, hi0);
                if (k1 < m) put(k1, lo1, hi1);
                if (k2 < m) put(k2, lo2, hi2);
                if (k3 < m) put(k3, lo3, hi3);
            }
            __syncthreads();
            NbrSrc src = global;
            src.rec = rec;
            body(tile, tile < tg.tiles, src);
            __syncthreads();
        } else {
            body(tile, tile < tg.tiles, global);
        }
    }
}

}  // namespace vfd
