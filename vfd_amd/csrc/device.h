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
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
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

// Block (kTile threads) sum of NV doubles; result valid in thread 0. Fixed order: wave tree, then waves 0..3.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV]) {
    __shared__ double sh[kTile / kWave][NV];
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
            for (int j = 1; j < (int)(kTile / kWave); ++j) s += sh[j][i];
            v[i] = s;
        }
    }
}

// Per-scene "last tile finalises" reduction. Every tile of scene s writes its NV partials to
// partials[tile*kPartialSlots + i], then bumps counters[s]; the tile that completes the count sums the scene's
// partials in tile order (deterministic) and returns true in thread 0 with the totals in out[]. The counter is reset
// for the next use. Must be called by all threads of the block.
template <int NV>
__device__ __forceinline__ bool tile_reduce_last(double (&v)[NV], double* partials, uint32_t* counters, uint32_t s,
                                                 uint32_t tileBegin, uint32_t tileEnd, double (&out)[NV]) {
    __shared__ int isLast;
    block_sum<NV>(v);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < NV; ++i) partials[(size_t)blockIdx.x * kPartialSlots + i] = v[i];
        __threadfence();
        const uint32_t done = atomicAdd(&counters[s], 1u);
        isLast = (done == tileEnd - tileBegin - 1u) ? 1 : 0;
    }
    __syncthreads();
    if (!isLast) return false;
    __threadfence();
    // sum partials of the scene's tiles: thread t takes tiles t, t+kTile, ...; then fixed-order block sum
    double acc[NV];
#pragma unroll
    for (int i = 0; i < NV; ++i) acc[i] = 0.0;
    for (uint32_t t = tileBegin + threadIdx.x; t < tileEnd; t += kTile)
#pragma unroll
        for (int i = 0; i < NV; ++i) acc[i] += partials[(size_t)t * kPartialSlots + i];
    block_sum<NV>(acc);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < NV; ++i) out[i] = acc[i];
        counters[s] = 0u;
    }
    return threadIdx.x == 0;
}

// Same pattern for a float max.
__device__ __forceinline__ bool tile_max_last(float v, float* partials, uint32_t* counters, uint32_t s,
                                              uint32_t tileBegin, uint32_t tileEnd, float& out) {
    __shared__ float sh[kTile / kWave];
    __shared__ int isLast;
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    v = wave_max(v);
    if (lane == 0) sh[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        float m = sh[0];
        for (int j = 1; j < (int)(kTile / kWave); ++j) m = fmaxf(m, sh[j]);
        partials[blockIdx.x] = m;
        __threadfence();
        const uint32_t done = atomicAdd(&counters[s], 1u);
        isLast = (done == tileEnd - tileBegin - 1u) ? 1 : 0;
    }
    __syncthreads();
    if (!isLast) return false;
    __threadfence();
    float m = -FLT_MAX;
    for (uint32_t t = tileBegin + threadIdx.x; t < tileEnd; t += kTile)
        m = fmaxf(m, partials[t]);
    m = wave_max(m);
    __syncthreads();
    if (lane == 0) sh[w] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        float r = sh[0];
        for (int j = 1; j < (int)(kTile / kWave); ++j) r = fmaxf(r, sh[j]);
        out = r;
        counters[s] = 0u;
    }
    return threadIdx.x == 0;
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

}  // namespace vfd
