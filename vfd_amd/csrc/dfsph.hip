// dfsph.hip - the DFSPH solver kernels (density, factor, divergence and density Jacobi solves, implicit viscosity
// PCG, Monte-Carlo surface tension, CFL time step, integration).
//
// Reference: DFSPHImplementation.cu:63-170,395-839 and DFSPHKernels.cu:6-1046 (Goubermouche/VFD). Each kernel
// below names the reference kernel(s) it replaces. Differences by design:
//   - structure-of-arrays in Morton slot order, transposed neighbour lists (see common.h);
//   - density and factor share one neighbour loop (one table fetch per pair gives W and gradW);
//   - solver control lives on the device: every reduction is a deterministic FP64 per-tile sum that a one-block-per-
//     scene finalize kernel adds up in tile order; that kernel also evaluates the reference's loop condition, so the
//     host only polls every few iterations instead of copying a scalar back per iteration;
//   - ClearAcceleration is fused into the divergence finaliser; the position update into the pressure finaliser.
// Floating-point expressions keep the reference's association order (no implicit contraction; -ffp-contract=off).
#include "device.h"
#include "kernels.h"

namespace vfd {

__device__ __forceinline__ KTab ktab(const KTabHost& h) { return KTab{h.tab, h.R, h.R2, h.invStep, h.W0}; }
__device__ __forceinline__ uint32_t nbr(const Fields& F, uint32_t cap, uint32_t i, uint32_t j) {
    return F.nbrList[(size_t)j * cap + i];
}
__device__ __forceinline__ HaloView halo(const Fields& F) { return HaloView{F.haloCount, F.haloIds}; }
__device__ __forceinline__ NbrSrc global_src(const Fields& F, uint32_t cap) {
    return NbrSrc{nullptr, F.nbrLocal, F.nbrList, F.nbrCount, cap};
}

// Dynamic LDS of the staged kernels: the kernel table, then the halo records of the current group.
constexpr size_t kLdsDensity = kLdsWG + HaloRec<LoadX>::kBytes;
constexpr size_t kLdsXS = kLdsG + HaloRec<LoadXS>::kBytes;
constexpr size_t kLdsXV = kLdsG + HaloRec<LoadXV>::kBytes;
constexpr size_t kLdsXSV = kLdsG + HaloRec<LoadXQV>::kBytes;
static_assert(HaloRec<LoadXQV>::kBytes == HaloRec<LoadXSV>::kBytes && HaloRec<LoadXQ>::kBytes == HaloRec<LoadXS>::kBytes,
              "viscosity records are the size of the records they replace");
static_assert(kLdsDensity <= 160 * 1024 && kLdsXS <= 160 * 1024 && kLdsXV <= 160 * 1024 && kLdsXSV <= 160 * 1024,
              "staged kernels must fit the 160 KiB LDS of a CU");

// ---- ComputeDensityKernel (DFSPHKernels.cu:152-187) + ComputeDFSPHFactorKernel (:189-231) ----------------------
// One table fetch per pair gives W (cutoff r^2 <= R^2) and gradW (cutoff |r| <= R); the {W, G} table sits in LDS.
__global__ void __launch_bounds__(kGroupSize, kGroupsPerCU) k_density_factor(const SceneParams* __restrict__ P,
                                                           const uint32_t* __restrict__ tileScene, KTabHost kh,
                                                           Fields F, uint32_t cap, TileGrid tg) {
    extern __shared__ float4 lds[];
    float2* sWG = reinterpret_cast<float2*>(lds);
    load_table_wg(sWG, kh.tab);
    const KTab k{sWG, kh.R, kh.R2, kh.invStep, kh.W0};
    float* rec = reinterpret_cast<float*>(lds + kLdsWG / sizeof(float4));
    group_loop<1>(tg, halo(F), rec, global_src(F, cap), LoadX{F.pos}, [&](uint32_t tile, bool ok, const NbrSrc& src) {
        const TileCtx c = tile_ctx4(tile, ok, tileScene, P);
        if (!c.valid) return;
        const SceneParams& p = P[c.s];
        const f3 xi = ld3(F.pos[c.i]);
        const float V = p.V;
        float dens = V * k.W0;
        f3 gi = mk3(0.0f, 0.0f, 0.0f);
        float sum = 0.0f;
        each_neighbor_rec<1>(src, c.i, k.R, k.invStep, xi, k.tab, LoadX{F.pos}, [&](float4 lo, float4, float2 t) {
            // W: cutoff r^2 <= R^2 always holds for a listed pair (0 < d^2 < R^2); beyond the gradient cutoff the
            // index GetW reads is the last entry (|r| > R gives |r| / step > 9998), which the table's entry
            // kGCodeZero repeats with G = 0
            const f3 r = xi - ld3(lo);
            dens += V * t.x;
            const f3 gj = (-V) * (t.y * r);
            sum += dot3(gj, gj);
            gi = gi - gj;
        });
        for (uint32_t b = 0; b < p.bodyCount; ++b) {
            const float vb = F.vb[(size_t)b * cap + c.i];
            if (vb > 0.0f) {
                const f3 r = xi - ld3(F.xb[(size_t)b * cap + c.i]);
                dens += vb * getW(k, r);
                const f3 gj = (-vb) * gradW(k, r);
                gi = gi - gj;
            }
        }
        sum += dot3(gi, gi);
        F.density[c.i] = dens * p.rho0;
        F.factor[c.i] = above_eps(sum) ? 1.0f / sum : 0.0f;
    });
}

void launch_density_factor(const Launch& L, const KTabHost& kt, const Fields& F) {
    hipLaunchKernelGGL(k_density_factor, dim3(L.gridGroups), dim3(kGroupSize), kLdsDensity,
                       L.stream, L.P, L.tileScene, kt, F, L.capacity, TileGrid(L.tg));
}

// ---- shared neighbour sums --------------------------------------------------------------------------------------
// v_i-weighted divergence sum_j (v_i - v_j).gradW_ij * V + sum_b V_b v_i.gradW_ib (:233-278, :483-529);
// records lo = {x_j}, hi = {v_j}
__device__ __forceinline__ float velocity_divergence(const SceneParams& p, const KLds& k, const Fields& F,
                                                     uint32_t cap, uint32_t i, const NbrSrc& src) {
    const f3 xi = ld3(F.pos[i]), vi = ld3(F.vel[i]);
    float d = 0.0f;
    each_neighbor_rec<2>(src, i, k.R, k.invStep, xi, k.sG, LoadXV{F.pos, F.vel}, [&](float4 lo, float4 hi, float g) {
        d += dot3(vi - ld3(hi), g * (xi - ld3(lo)));
    });
    d *= p.V;
    for (uint32_t b = 0; b < p.bodyCount; ++b) {
        const float vb = F.vb[(size_t)b * cap + i];
        if (vb > 0.0f) d += vb * dot3(vi, gradW(k, xi - ld3(F.xb[(size_t)b * cap + i])));
    }
    return d;
}

// pressure acceleration from kappa (ComputePressureAccelerationKernel :342-386 and its three siblings);
// records lo = {x_j, kappa_j}
__device__ __forceinline__ f3 pressure_accel(const SceneParams& p, const KLds& k, const Fields& F, uint32_t cap,
                                             uint32_t i, const float* __restrict__ kap, const NbrSrc& src) {
    const f3 xi = ld3(F.pos[i]);
    const float ki = kap[i];
    const float negV = -p.V;  // read before the loop: a global read inside it would wait for the list prefetch
    f3 a = mk3(0.0f, 0.0f, 0.0f);
    each_neighbor_rec<1>(src, i, k.R, k.invStep, xi, k.sG, LoadXS{F.pos, kap}, [&](float4 lo, float4, float g) {
        const float s = ki + lo.w;
        if (above_eps(fabsf(s))) {
            const f3 gj = negV * (g * (xi - ld3(lo)));
            a = a + s * gj;
        }
    });
    if (above_eps(fabsf(ki))) {
        for (uint32_t b = 0; b < p.bodyCount; ++b) {
            const float vb = F.vb[(size_t)b * cap + i];
            if (vb > 0.0f) {
                const f3 gj = (-vb) * gradW(k, xi - ld3(F.xb[(size_t)b * cap + i]));
                a = a + ki * gj;
            }
        }
    }
    return a;
}

// ComputeDensityPressureForce (:280-315); records lo = {x_j}, hi = {a_j}
__device__ __forceinline__ float density_pressure_force(const SceneParams& p, const KLds& k, const Fields& F,
                                                        uint32_t cap, uint32_t i, const NbrSrc& src) {
    const f3 xi = ld3(F.pos[i]), ai = ld3(F.ap[i]);
    float f = 0.0f;
    each_neighbor_rec<2>(src, i, k.R, k.invStep, xi, k.sG, LoadXV{F.pos, F.ap}, [&](float4 lo, float4 hi, float g) {
        f += dot3(ai - ld3(hi), g * (xi - ld3(lo)));
    });
    f *= p.V;
    for (uint32_t b = 0; b < p.bodyCount; ++b) {
        const float vb = F.vb[(size_t)b * cap + i];
        if (vb > 0.0f) f += vb * dot3(ai, gradW(k, xi - ld3(F.xb[(size_t)b * cap + i])));
    }
    return f;
}

// Every gradient-loop kernel stages the G table in LDS, followed by the halo records of the current group.
#define VFD_HALO_KERNEL_PROLOGUE(kh)                                   \
    extern __shared__ float4 lds[];                                    \
    float* sG = reinterpret_cast<float*>(lds);                         \
    load_table_g(sG, (kh).tab);                                        \
    const KLds K{sG, (kh).R, (kh).invStep};                            \
    float* rec = reinterpret_cast<float*>(lds + kLdsG / sizeof(float4));

// ---- Jacobi solver control (DFSPHImplementation.cu:455-461, 526-532) -------------------------------------------
// The per-scene finalisers (Jacobi, CG) run one 1024-thread block per scene: the tile partials of a 2M scene are then
// loaded in one round.
constexpr int kFinishThreads = 1024;
// which = 0: divergence (eta = (1/dt) tol 1e-4 rho0), 1: pressure (eta = tol 1e-4 rho0)
// One block; also counts the scenes that iterate into live[which] (and clears live[2], the viscosity count, at the
// start of a step).
__global__ void __launch_bounds__(kTile) k_jacobi_begin(const SceneParams* __restrict__ P, SceneState* state,
                                                        uint32_t* live, uint32_t sceneCount, int which) {
    __shared__ uint32_t count;
    if (threadIdx.x == 0) count = 0;
    __syncthreads();
    for (uint32_t s = threadIdx.x; s < sceneCount; s += blockDim.x) {
        const SceneParams& p = P[s];
        SceneState& st = state[s];
        JacobiState& J = which == 0 ? st.div : st.pres;
        J.it = 0;
        J.err = 0.0f;
        if (which == 0) {
            J.eta = st.dtInv * p.divTol * 0.0001f * p.rho0;
            J.active = p.enableDiv && ((J.err > J.eta || 0u < p.minDiv) && 0u < p.maxDiv) ? 1u : 0u;
        } else {
            J.eta = p.presTol * 0.0001f * p.rho0;
            J.active = ((J.err > J.eta || 0u < p.minPres) && 0u < p.maxPres) ? 1u : 0u;
        }
        if (J.active && p.count) atomicAdd(&count, 1u);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        live[which] = count;
        if (which == 0) live[2] = 0u;
    }
}

// Residual error -rho0 * sum(r) / N, iteration count and the loop condition (DFSPHImplementation.cu:461, 483-492,
// 532, 554-563), one block per scene after every iteration kernel.
__global__ void __launch_bounds__(kFinishThreads) k_jacobi_finish(const SceneParams* __restrict__ P, Fields F, int which) {
    const uint32_t s = blockIdx.x;
    const SceneParams& p = P[s];
    JacobiState& J = which == 0 ? F.state[s].div : F.state[s].pres;
    if (!J.active || p.count == 0) return;  // uniform per block
    double out[1];
    scene_sum<1, kFinishThreads>(F.partials, p.tileBegin, p.tileEnd, out);
    if (threadIdx.x == 0) {
        const uint32_t minIt = which == 0 ? p.minDiv : p.minPres, maxIt = which == 0 ? p.maxDiv : p.maxPres;
        const float err = (float)(-out[0]) / (float)p.count;
        const uint32_t it = J.it + 1u;
        J.err = err;
        J.it = it;
        J.active = ((err > J.eta || it < minIt) && it < maxIt) ? 1u : 0u;
        if (!J.active) atomicSub(&F.live[which], 1u);
    }
}

// ---- divergence solve (ComputeDivergence, DFSPHImplementation.cu:506-575) ---------------------------------------
// ComputeDensityChangeKernel (DFSPHKernels.cu:483-529)
__global__ void __launch_bounds__(kGroupSize, kGroupsPerCU) k_div_init(const SceneParams* __restrict__ P,
                                                    const uint32_t* __restrict__ tileScene, KTabHost kh, Fields F,
                                                    uint32_t cap, TileGrid tg) {
    VFD_HALO_KERNEL_PROLOGUE(kh)
    group_loop<2>(tg, halo(F), rec, global_src(F, cap), LoadXV{F.pos, F.vel}, [&](uint32_t tile, bool ok, const NbrSrc& src) {
        const TileCtx c = tile_ctx4(tile, ok, tileScene, P);
        const SceneParams& p = P[c.s];
        if (!c.valid || !p.enableDiv) return;
        float d = velocity_divergence(p, K, F, cap, c.i, src);
        d = F.nbrCount[c.i] < 20u ? 0.0f : fmaxf(d, 0.0f);
        F.densAdv[c.i] = d;
        const float f = F.factor[c.i] * F.state[c.s].dtInv;
        F.factor[c.i] = f;
        F.kappaV[c.i] = d * f;
    });
}

// ComputePressureAccelerationAndDivergenceKernel (:388-432)
__global__ void __launch_bounds__(kGroupSize, kGroupsPerCU) k_div_accel(const SceneParams* __restrict__ P,
                                                     const uint32_t* __restrict__ tileScene, KTabHost kh, Fields F,
                                                     uint32_t cap, TileGrid tg) {
    if (F.live[0] == 0u) return;  // every scene has left this solve (block-uniform)
    VFD_HALO_KERNEL_PROLOGUE(kh)
    group_loop<1>(tg, halo(F), rec, global_src(F, cap), LoadXS{F.pos, F.kappaV}, [&](uint32_t tile, bool ok, const NbrSrc& src) {
        const TileCtx c = tile_ctx4(tile, ok, tileScene, P);
        if (!F.state[c.s].div.active || !c.valid) return;
        F.ap[c.i] = st4(pressure_accel(P[c.s], K, F, cap, c.i, F.kappaV, src));
    });
}

// DivergenceSolveIterationKernel (:531-553) + per-tile residual partial sums (DFSPHImplementation.cu:554-563)
__global__ void __launch_bounds__(kGroupSize, kGroupsPerCU) k_div_iter(const SceneParams* __restrict__ P,
                                                    const uint32_t* __restrict__ tileScene, KTabHost kh, Fields F,
                                                    uint32_t cap, TileGrid tg) {
    if (F.live[0] == 0u) return;  // every scene has left this solve (block-uniform)
    VFD_HALO_KERNEL_PROLOGUE(kh)
    group_loop<2>(tg, halo(F), rec, global_src(F, cap), LoadXV{F.pos, F.ap}, [&](uint32_t tile, bool ok, const NbrSrc& src) {
        const TileCtx c = tile_ctx4(tile, ok, tileScene, P);
        const SceneState& st = F.state[c.s];
        const bool act = ok && st.div.active;
        const SceneParams& p = P[c.s];
        double part = 0.0;
        if (act && c.valid) {
            const float f = density_pressure_force(p, K, F, cap, c.i, src) * st.dt;
            const float r = F.nbrCount[c.i] < 20u ? 0.0f : fminf(-F.densAdv[c.i] - f, 0.0f);
            F.residuum[c.i] = r;
            F.kappaV[c.i] -= r * F.factor[c.i];
            part = (double)(p.rho0 * r);
        }
        double v[1] = {part};
        tile_partials4<1>(v, F.partials, tile, act);
    });
}

// ComputePressureAccelerationAndFactorKernel (:555-602) + ClearAccelerationKernel (:24-37)
__global__ void __launch_bounds__(kGroupSize, kGroupsPerCU) k_div_final(const SceneParams* __restrict__ P,
                                                     const uint32_t* __restrict__ tileScene, KTabHost kh, Fields F,
                                                     uint32_t cap, TileGrid tg) {
    VFD_HALO_KERNEL_PROLOGUE(kh)
    group_loop<1>(tg, halo(F), rec, global_src(F, cap), LoadXS{F.pos, F.kappaV}, [&](uint32_t tile, bool ok, const NbrSrc& src) {
        const TileCtx c = tile_ctx4(tile, ok, tileScene, P);
        if (!c.valid) return;
        const SceneParams& p = P[c.s];
        if (p.enableDiv) {
            const float dt = F.state[c.s].dt;
            const f3 a = pressure_accel(p, K, F, cap, c.i, F.kappaV, src);
            F.ap[c.i] = st4(a);
            F.vel[c.i] = st4(ld3(F.vel[c.i]) + dt * a);
            F.factor[c.i] *= dt;
        }
        F.acc[c.i] = make_float4(p.gx, p.gy, p.gz, 0.0f);
    });
}

void launch_scene_init_div(const Launch& L, const Fields& F) {
    hipLaunchKernelGGL(k_jacobi_begin, dim3(1), dim3(kTile), 0, L.stream, L.P, F.state, F.live, L.sceneCount, 0);
}
void launch_div_init(const Launch& L, const KTabHost& kt, const Fields& F) {
    hipLaunchKernelGGL(k_div_init, dim3(L.gridGroups), dim3(kGroupSize), kLdsXV, L.stream, L.P, L.tileScene, kt, F, L.capacity,
                       TileGrid(L.tg));
}
void launch_div_accel(const Launch& L, const KTabHost& kt, const Fields& F) {
    hipLaunchKernelGGL(k_div_accel, dim3(L.gridGroups), dim3(kGroupSize), kLdsXS, L.stream, L.P, L.tileScene, kt, F, L.capacity,
                       TileGrid(L.tg));
}
void launch_div_iter(const Launch& L, const KTabHost& kt, const Fields& F) {
    hipLaunchKernelGGL(k_div_iter, dim3(L.gridGroups), dim3(kGroupSize), kLdsXV, L.stream, L.P, L.tileScene, kt, F, L.capacity,
                       TileGrid(L.tg));
}
// the residual reduction and loop condition after a Jacobi iteration; which = 0 divergence, 1 pressure
void launch_jacobi_finish(const Launch& L, const Fields& F, int which) {
    hipLaunchKernelGGL(k_jacobi_finish, dim3(L.sceneCount), dim3(kFinishThreads), 0, L.stream, L.P, F, which);
}
void launch_div_final(const Launch& L, const KTabHost& kt, const Fields& F) {
    hipLaunchKernelGGL(k_div_final, dim3(L.gridGroups), dim3(kGroupSize), kLdsXS, L.stream, L.P, L.tileScene, kt, F, L.capacity,
                       TileGrid(L.tg));
}

// ---- pressure solve (ComputePressure, DFSPHImplementation.cu:443-504) -------------------------------------------
// ComputeDensityAdvectionKernel (DFSPHKernels.cu:233-278)
__global__ void __launch_bounds__(kGroupSize, kGroupsPerCU) k_pres_init(const SceneParams* __restrict__ P,
                                                     const uint32_t* __restrict__ tileScene, KTabHost kh, Fields F,
                                                     uint32_t cap, TileGrid tg) {
    VFD_HALO_KERNEL_PROLOGUE(kh)
    group_loop<2>(tg, halo(F), rec, global_src(F, cap), LoadXV{F.pos, F.vel}, [&](uint32_t tile, bool ok, const NbrSrc& src) {
        const TileCtx c = tile_ctx4(tile, ok, tileScene, P);
        if (!c.valid) return;
        const SceneParams& p = P[c.s];
        const SceneState& st = F.state[c.s];
        const float d = velocity_divergence(p, K, F, cap, c.i, src);
        const float da = F.density[c.i] / p.rho0 + st.dt * d;
        F.densAdv[c.i] = da;
        const float f = F.factor[c.i] * st.dt2Inv;
        F.factor[c.i] = f;
        const float r = fminf(1.0f - da, 0.0f);
        F.kappa[c.i] = -r * f;
    });
}

// ComputePressureAccelerationKernel (:342-386)
__global__ void __launch_bounds__(kGroupSize, kGroupsPerCU) k_pres_accel(const SceneParams* __restrict__ P,
                                                      const uint32_t* __restrict__ tileScene, KTabHost kh, Fields F,
                                                      uint32_t cap, TileGrid tg) {
    if (F.live[1] == 0u) return;  // every scene has left this solve (block-uniform)
    VFD_HALO_KERNEL_PROLOGUE(kh)
    group_loop<1>(tg, halo(F), rec, global_src(F, cap), LoadXS{F.pos, F.kappa}, [&](uint32_t tile, bool ok, const NbrSrc& src) {
        const TileCtx c = tile_ctx4(tile, ok, tileScene, P);
        if (!F.state[c.s].pres.active || !c.valid) return;
        F.ap[c.i] = st4(pressure_accel(P[c.s], K, F, cap, c.i, F.kappa, src));
    });
}

// PressureSolveIterationKernel (:317-340) + per-tile residual partial sums (DFSPHImplementation.cu:483-491)
__global__ void __launch_bounds__(kGroupSize, kGroupsPerCU) k_pres_iter(const SceneParams* __restrict__ P,
                                                     const uint32_t* __restrict__ tileScene, KTabHost kh, Fields F,
                                                     uint32_t cap, TileGrid tg) {
    if (F.live[1] == 0u) return;  // every scene has left this solve (block-uniform)
    VFD_HALO_KERNEL_PROLOGUE(kh)
    group_loop<2>(tg, halo(F), rec, global_src(F, cap), LoadXV{F.pos, F.ap}, [&](uint32_t tile, bool ok, const NbrSrc& src) {
        const TileCtx c = tile_ctx4(tile, ok, tileScene, P);
        const SceneState& st = F.state[c.s];
        const bool act = ok && st.pres.active;
        const SceneParams& p = P[c.s];
        double part = 0.0;
        if (act && c.valid) {
            const float f = density_pressure_force(p, K, F, cap, c.i, src) * st.dt2;
            const float r = fminf(1.0f - F.densAdv[c.i] - f, 0.0f);
            F.residuum[c.i] = r;
            F.kappa[c.i] -= r * F.factor[c.i];
            part = (double)(p.rho0 * r);
        }
        double v[1] = {part};
        tile_partials4<1>(v, F.partials, tile, act);
    });
}

// ComputePressureAccelerationAndVelocityKernel (:434-481) + ComputePositionKernel (:54-67) into posNext
__global__ void __launch_bounds__(kGroupSize, kGroupsPerCU) k_pres_final(const SceneParams* __restrict__ P,
                                                      const uint32_t* __restrict__ tileScene, KTabHost kh, Fields F,
                                                      uint32_t cap, TileGrid tg, int writeNext) {
    VFD_HALO_KERNEL_PROLOGUE(kh)
    group_loop<1>(tg, halo(F), rec, global_src(F, cap), LoadXS{F.pos, F.kappa}, [&](uint32_t tile, bool ok, const NbrSrc& src) {
        const TileCtx c = tile_ctx4(tile, ok, tileScene, P);
        if (!c.valid) return;
        const SceneParams& p = P[c.s];
        const float dt = F.state[c.s].dt;
        const f3 a = pressure_accel(p, K, F, cap, c.i, F.kappa, src);
        F.ap[c.i] = st4(a);
        const f3 v = ld3(F.vel[c.i]) + dt * a;
        F.vel[c.i] = st4(v);
        const f3 x = ld3(F.pos[c.i]) + dt * v;
        if (writeNext) F.posNext[c.i] = st4(x);
        // per-step non-finite check (no reference counterpart): any lane of the scene may raise the flag
        if (!(isfinite(v.x) && isfinite(v.y) && isfinite(v.z) && isfinite(x.x) && isfinite(x.y) && isfinite(x.z)))
            F.state[c.s].nonfinite = 1u;
    });
}

void launch_scene_init_pres(const Launch& L, const Fields& F) {
    hipLaunchKernelGGL(k_jacobi_begin, dim3(1), dim3(kTile), 0, L.stream, L.P, F.state, F.live, L.sceneCount, 1);
}
void launch_pres_init(const Launch& L, const KTabHost& kt, const Fields& F) {
    hipLaunchKernelGGL(k_pres_init, dim3(L.gridGroups), dim3(kGroupSize), kLdsXV, L.stream, L.P, L.tileScene, kt, F, L.capacity,
                       TileGrid(L.tg));
}
void launch_pres_accel(const Launch& L, const KTabHost& kt, const Fields& F) {
    hipLaunchKernelGGL(k_pres_accel, dim3(L.gridGroups), dim3(kGroupSize), kLdsXS, L.stream, L.P, L.tileScene, kt, F,
                       L.capacity, TileGrid(L.tg));
}
void launch_pres_iter(const Launch& L, const KTabHost& kt, const Fields& F) {
    hipLaunchKernelGGL(k_pres_iter, dim3(L.gridGroups), dim3(kGroupSize), kLdsXV, L.stream, L.P, L.tileScene, kt, F, L.capacity,
                       TileGrid(L.tg));
}
void launch_pres_final(const Launch& L, const KTabHost& kt, const Fields& F, int writeNext) {
    hipLaunchKernelGGL(k_pres_final, dim3(L.gridGroups), dim3(kGroupSize), kLdsXS, L.stream, L.P, L.tileScene, kt, F,
                       L.capacity, TileGrid(L.tg), writeNext);
}

// ---- time step (ComputeTimeStepSize, DFSPHImplementation.cu:395-441) --------------------------------------------
__global__ void __launch_bounds__(kTile) k_time_step(const SceneParams* __restrict__ P,
                                                     const uint32_t* __restrict__ tileScene, Fields F, TileGrid tg) {
    tile_loop(tg, [&](uint32_t tile) {
        const TileCtx c = tile_ctx(tile, tileScene, P);
        const SceneState& st = F.state[c.s];
        float m = -FLT_MAX;
        if (c.valid) m = length2(ld3(F.vel[c.i]) + ld3(F.acc[c.i]) * st.dt);
        m = block_max(m);
        if (threadIdx.x == 0) F.partialsF[tile] = m;
    });
}

// max(0.1, max |v + a dt|^2) -> CFL time step and the quantities derived from it, one block per scene
__global__ void __launch_bounds__(kTile) k_time_step_finish(const SceneParams* __restrict__ P, Fields F) {
    const uint32_t s = blockIdx.x;
    const SceneParams& p = P[s];
    if (p.count == 0) return;
    float m = -FLT_MAX;
    for (uint32_t t = p.tileBegin + threadIdx.x; t < p.tileEnd; t += kTile) m = fmaxf(m, F.partialsF[t]);
    m = block_max(m);
    if (threadIdx.x == 0) {
        SceneState& st = F.state[s];
        st.nonfinite = 0u;  // raised again by k_pres_final if this step ends with a non-finite particle
        float mv = fmaxf(0.1f, m);
        st.maxVel = mv;
        if (mv < 1.0e-9f) mv = 1.0e-9f;
        float dt = 0.4f * (p.diam / sqrtf(mv));
        dt = fminf(dt, p.dtMax);
        dt = fmaxf(dt, p.dtMin);
        st.dt = dt;
        st.dt2 = dt * dt;
        st.dtInv = 1.0f / dt;
        st.dt2Inv = 1.0f / st.dt2;
        st.S = p.csdFix > 0 ? (uint32_t)p.csdFix : (uint32_t)(int32_t)((float)p.csd * dt);
    }
}

// ConvertParticlesToBuffer (ParticleBuffer/DFSPHParticleBuffer.cu:6-22): {x, v, a} of the particles of one scene,
// 9 floats each (DFSPHParticleSimple), written at their original index so that the frame is in input order
__global__ void __launch_bounds__(kTile) k_frame_pack(uint32_t base, uint32_t count, const uint32_t* __restrict__ id,
                                                      const float4* __restrict__ x, const float4* __restrict__ v,
                                                      const float4* __restrict__ a, float* out) {
    const uint32_t k = blockIdx.x * kTile + threadIdx.x;
    if (k >= count) return;
    const uint32_t i = base + k;
    const uint32_t j = id[i];
    if (j >= count) return;  // ids of a scene are 0..count-1; anything else is not written
    const float4 xi = x[i], vi = v[i], ai = a[i];
    float* o = out + (size_t)9 * j;
    o[0] = xi.x, o[1] = xi.y, o[2] = xi.z;
    o[3] = vi.x, o[4] = vi.y, o[5] = vi.z;
    o[6] = ai.x, o[7] = ai.y, o[8] = ai.z;
}
void launch_frame_pack(hipStream_t s, uint32_t base, uint32_t count, const uint32_t* id, const float4* x,
                       const float4* v, const float4* a, float* out) {
    if (count) hipLaunchKernelGGL(k_frame_pack, dim3((count + kTile - 1) / kTile), dim3(kTile), 0, s, base, count, id,
                                  x, v, a, out);
}

// ComputeVelocityKernel (DFSPHKernels.cu:39-52)
__global__ void __launch_bounds__(kTile) k_velocity(const SceneParams* __restrict__ P,
                                                    const uint32_t* __restrict__ tileScene, Fields F, TileGrid tg) {
    tile_loop(tg, [&](uint32_t tile) {
        const TileCtx c = tile_ctx(tile, tileScene, P);
        if (!c.valid) return;
        F.vel[c.i] = st4(ld3(F.vel[c.i]) + F.state[c.s].dt * ld3(F.acc[c.i]));
    });
}

void launch_time_step(const Launch& L, const Fields& F) {
    hipLaunchKernelGGL(k_time_step, dim3(L.gridAll), dim3(kTile), 0, L.stream, L.P, L.tileScene, F, TileGrid(L.tg));
}
void launch_time_step_finish(const Launch& L, const Fields& F) {
    hipLaunchKernelGGL(k_time_step_finish, dim3(L.sceneCount), dim3(kTile), 0, L.stream, L.P, F);
}
void launch_velocity(const Launch& L, const Fields& F) {
    hipLaunchKernelGGL(k_velocity, dim3(L.gridAll), dim3(kTile), 0, L.stream, L.P, L.tileScene, F, TileGrid(L.tg));
}

// ---- implicit viscosity (ComputeViscosity, DFSPHImplementation.cu:577-615; Weiler et al. 2018) ------------------
struct M3 {
    float m[9];  // row-major
};
__device__ __forceinline__ M3 inverse3(const M3& A) {
    const float a = A.m[0], b = A.m[1], c = A.m[2];
    const float d = A.m[3], e = A.m[4], f = A.m[5];
    const float g = A.m[6], h = A.m[7], i = A.m[8];
    const float c00 = e * i - f * h, c01 = f * g - d * i, c02 = d * h - e * g;
    const float det = a * c00 + b * c01 + c * c02;
    const float s = 1.0f / det;
    M3 r;
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
__device__ __forceinline__ f3 minv_mul(const float* __restrict__ minv, uint32_t cap, uint32_t i, f3 v) {
    float m[9];
#pragma unroll
    for (int e = 0; e < 9; ++e) m[e] = minv[(size_t)e * cap + i];
    return mk3(m[0] * v.x + m[1] * v.y + m[2] * v.z, m[3] * v.x + m[4] * v.y + m[5] * v.z,
               m[6] * v.x + m[7] * v.y + m[8] * v.z);
}

// GetOrthogonalVectors (Core/Math/Math.h:19-33)
__device__ __forceinline__ void orthogonal_vectors(f3 vec, f3& x, f3& y) {
    f3 v = mk3(1.0f, 0.0f, 0.0f);
    if ((double)fabsf(dot3(v, vec)) > 0.999) v = mk3(0.0f, 1.0f, 0.0f);
    x = cross(vec, v);
    y = cross(vec, x);
    x = normalize(x);
    y = normalize(y);
}

// four tangential boundary samples (DFSPHKernels.cu:643-671, 713-742, 800-829)
struct Tangential {
    f3 pd[4];
    f3 gw[4];
    float vol;
};
__device__ __forceinline__ bool tangential(const SceneParams& p, const KLds& k, f3 xi, f3 xb, float vb, Tangential& t) {
    f3 normal = -(xi - xb);
    const float nl = length(normal);
    if (!(nl > 0.0001f)) return false;
    normal = normal / nl;
    f3 t1, t2;
    orthogonal_vectors(normal, t1, t2);
    const f3 pos[4] = {xb - t1 * p.td, xb + t1 * p.td, xb - t2 * p.td, xb + t2 * p.td};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        t.pd[q] = xi - pos[q];
        t.gw[q] = gradW(k, t.pd[q]);
    }
    t.vol = 0.25f * vb;
    return true;
}

// ComputeViscosityPreconditionerKernel (:604-684) + ComputeViscosityGradientKernel (:686-756): inverse diagonal
// block, right-hand side b = v (the boundary rhs term is dot(0, .) = 0) and initial guess x0 = v + dv.
__global__ void __launch_bounds__(kGroupSize, kGroupsPerCU) k_visc_precond(const SceneParams* __restrict__ P,
                                                        const uint32_t* __restrict__ tileScene, KTabHost kh,
                                                        Fields F, uint32_t cap, TileGrid tg) {
    VFD_HALO_KERNEL_PROLOGUE(kh)
    const LoadXQ load{F.pos, F.density, P[0].m};  // lo.w = m / rho_j
    group_loop<1>(tg, halo(F), rec, global_src(F, cap), load, [&](uint32_t tile, bool ok, const NbrSrc& src) {
        const TileCtx c = tile_ctx4(tile, ok, tileScene, P);
        const SceneParams& p = P[c.s];
        if (!c.valid || !p.enableVisc) return;
        const f3 xi = ld3(F.pos[c.i]);
        const float d001 = 0.01f * p.R2;
        const float mu10 = 10.0f * p.mu;  // read before the loop (see pressure_accel)
        float M[9];
#pragma unroll
        for (int e = 0; e < 9; ++e) M[e] = 0.0f;
        each_neighbor_rec<1>(src, c.i, K.R, K.invStep, xi, K.sG, load, [&](float4 lo, float4, float gt) {
            const f3 xij = xi - ld3(lo);
            const f3 gw = gt * xij;
            const float s = mu10 * lo.w / (length2(xij) + d001);
            const float a[3] = {xij.x, xij.y, xij.z}, g[3] = {gw.x, gw.y, gw.z};
#pragma unroll
            for (int r = 0; r < 3; ++r)
#pragma unroll
                for (int cc = 0; cc < 3; ++cc) M[3 * r + cc] += s * (a[r] * g[cc]);
        });
        if (p.muB != 0.0f) {
            for (uint32_t b = 0; b < p.bodyCount; ++b) {
                const float vb = F.vb[(size_t)b * cap + c.i];
                if (!(vb > 0.0f)) continue;
                Tangential t;
                if (!tangential(p, K, xi, ld3(F.xb[(size_t)b * cap + c.i]), vb, t)) continue;
                const float a[3] = {t.pd[0].x, t.pd[0].y, t.pd[0].z};  // the reference uses direction 1 for all four
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const float s = 10.0f * p.muB * t.vol / (length2(t.pd[q]) + d001);
                    const float g[3] = {t.gw[q].x, t.gw[q].y, t.gw[q].z};
#pragma unroll
                    for (int r = 0; r < 3; ++r)
#pragma unroll
                        for (int cc = 0; cc < 3; ++cc) M[3 * r + cc] += s * (a[r] * g[cc]);
                }
            }
        }
        const float f = F.state[c.s].dt / F.density[c.i];
        M3 A;
#pragma unroll
        for (int e = 0; e < 9; ++e) A.m[e] = (e % 4 == 0 ? 1.0f : 0.0f) - f * M[e];
        const M3 inv = inverse3(A);
#pragma unroll
        for (int e = 0; e < 9; ++e) F.minv[(size_t)e * cap + c.i] = inv.m[e];
        F.cgX[c.i] = st4(ld3(F.vel[c.i]) + ld3(F.dv[c.i]));
    });
}

// ComputeMatrixVecProdFunctionKernel (:758-842): y_i = x_i - dt/rho_i * (viscous acceleration of x)
// records lo = {x_j, m / rho_j}, hi = {in_j}
__device__ __forceinline__ f3 visc_matvec(const SceneParams& p, const KLds& k, const Fields& F, uint32_t cap,
                                          uint32_t i, const float4* __restrict__ in, float dt, const NbrSrc& src,
                                          const LoadXQV& load) {
    const f3 xi = ld3(F.pos[i]), vi = ld3(in[i]);
    const float d001 = 0.01f * p.R2;
    const float mu10 = 10.0f * p.mu;  // read before the loop (see pressure_accel)
    f3 acc = mk3(0.0f, 0.0f, 0.0f);
    each_neighbor_rec<2>(src, i, k.R, k.invStep, xi, k.sG, load, [&](float4 lo, float4 hi, float g) {
        const f3 xij = xi - ld3(lo);
        const f3 gw = g * xij;
        const float s = mu10 * lo.w * dot3(vi - ld3(hi), xij) / (length2(xij) + d001);
        acc = acc + s * gw;
    });
    if (p.muB != 0.0f) {
        for (uint32_t b = 0; b < p.bodyCount; ++b) {
            const float vb = F.vb[(size_t)b * cap + i];
            if (!(vb > 0.0f)) continue;
            Tangential t;
            if (!tangential(p, k, xi, ld3(F.xb[(size_t)b * cap + i]), vb, t)) continue;
            f3 s4;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const float s = 10.0f * p.muB * t.vol * dot3(vi, t.pd[q]) / (length2(t.pd[q]) + d001);
                s4 = q == 0 ? s * t.gw[q] : s4 + s * t.gw[q];
            }
            acc = acc + s4;
        }
    }
    return vi - (dt / F.density[i]) * acc;
}

// SolveViscosity set-up (DFSPHImplementation.cu:627-691): r = b - A x0, ||b||^2, ||r||^2, p = M^-1 r, |r.p|
__global__ void __launch_bounds__(kGroupSize, kGroupsPerCU) k_cg_begin(const SceneParams* __restrict__ P,
                                                    const uint32_t* __restrict__ tileScene, KTabHost kh, Fields F,
                                                    uint32_t cap, TileGrid tg) {
    VFD_HALO_KERNEL_PROLOGUE(kh)
    const LoadXQV load{F.pos, F.density, F.cgX, P[0].m};
    group_loop<2>(tg, halo(F), rec, global_src(F, cap), load, [&](uint32_t tile, bool ok, const NbrSrc& src) {
        const TileCtx c = tile_ctx4(tile, ok, tileScene, P);
        const SceneParams& p = P[c.s];
        const bool act = ok && p.enableVisc;
        const SceneState& st = F.state[c.s];
        double v[3] = {0.0, 0.0, 0.0};
        if (act && c.valid) {
            const f3 b = ld3(F.vel[c.i]);
            const f3 ax = visc_matvec(p, K, F, cap, c.i, F.cgX, st.dt, src, load);
            const f3 r = b - ax;
            const f3 z = minv_mul(F.minv, cap, c.i, r);
            F.cgR[c.i] = st4(r);
            F.cgP[c.i] = st4(z);
            v[0] = (double)dot3(b, b);
            v[1] = (double)dot3(r, r);
            v[2] = (double)dot3(r, z);
        }
        tile_partials4<3>(v, F.partials, tile, act);
    });
}

__global__ void __launch_bounds__(kFinishThreads) k_cg_begin_finish(const SceneParams* __restrict__ P, Fields F) {
    const uint32_t s = blockIdx.x;
    const SceneParams& p = P[s];
    if (!p.enableVisc || p.count == 0) return;
    double out[3];
    scene_sum<3, kFinishThreads>(F.partials, p.tileBegin, p.tileEnd, out);
    if (threadIdx.x == 0) {
        CgState& G = F.state[s].cg;
        G.it = 0;
        G.zero = 0;
        G.active = 0;
        G.loops = 0;
        const float rhs2 = (float)out[0];
        G.rhs2 = rhs2;
        if (rhs2 == 0.0f) {
            G.zero = 1;
            G.err = 0.0f;
            return;
        }
        G.thr = fmaxf(p.viscTol * p.viscTol * 0.0001f * rhs2, FLT_MIN);
        G.res2 = (float)out[1];
        G.err = sqrtf(G.res2 / rhs2);
        if (G.res2 < G.thr) return;
        G.absNew = fabsf((float)out[2]);
        G.active = (0u >= p.minVisc && 0u < p.maxVisc) ? 1u : 0u;
        if (G.active) atomicAdd(&F.live[2], 1u);
    }
}

// CG matvec Ap and p.Ap -> alpha (:695-720)
__global__ void __launch_bounds__(kGroupSize, kGroupsPerCU) k_cg_matvec(const SceneParams* __restrict__ P,
                                                     const uint32_t* __restrict__ tileScene, KTabHost kh, Fields F,
                                                     uint32_t cap, TileGrid tg) {
    if (F.live[2] == 0u) return;  // every scene has left this solve (block-uniform)
    VFD_HALO_KERNEL_PROLOGUE(kh)
    const LoadXQV load{F.pos, F.density, F.cgP, P[0].m};
    group_loop<2>(tg, halo(F), rec, global_src(F, cap), load, [&](uint32_t tile, bool ok, const NbrSrc& src) {
        const TileCtx c = tile_ctx4(tile, ok, tileScene, P);
        const SceneState& st = F.state[c.s];
        const bool act = ok && st.cg.active;
        double v[1] = {0.0};
        if (act && c.valid) {
            const f3 ap = visc_matvec(P[c.s], K, F, cap, c.i, F.cgP, st.dt, src, load);
            F.cgAp[c.i] = st4(ap);
            v[0] = (double)dot3(ld3(F.cgP[c.i]), ap);
        }
        tile_partials4<1>(v, F.partials, tile, act);
    });
}

__global__ void __launch_bounds__(kFinishThreads) k_cg_alpha(const SceneParams* __restrict__ P, Fields F) {
    if (F.live[2] == 0u) return;  // every scene has left this solve (block-uniform)
    const uint32_t s = blockIdx.x;
    const SceneParams& p = P[s];
    CgState& G = F.state[s].cg;
    if (!G.active || p.count == 0) return;
    double out[1];
    scene_sum<1, kFinishThreads>(F.partials, p.tileBegin, p.tileEnd, out);
    if (threadIdx.x == 0) G.alpha = G.absNew / (float)out[0];
}

// x += alpha p, r -= alpha Ap, ||r||^2, z = M^-1 r, |r.z| -> break test, beta (:722-804)
__global__ void __launch_bounds__(kTile) k_cg_update(const SceneParams* __restrict__ P,
                                                     const uint32_t* __restrict__ tileScene, Fields F, uint32_t cap,
                                                     TileGrid tg) {
    if (F.live[2] == 0u) return;  // every scene has left this solve (block-uniform)
    tile_loop(tg, [&](uint32_t tile) {
        const TileCtx c = tile_ctx(tile, tileScene, P);
        const SceneState& st = F.state[c.s];
        if (!st.cg.active) return;
        double v[2] = {0.0, 0.0};
        if (c.valid) {
            const float alpha = st.cg.alpha;
            F.cgX[c.i] = st4(ld3(F.cgX[c.i]) + ld3(F.cgP[c.i]) * alpha);
            const f3 r = ld3(F.cgR[c.i]) - ld3(F.cgAp[c.i]) * alpha;
            F.cgR[c.i] = st4(r);
            const f3 z = minv_mul(F.minv, cap, c.i, r);
            F.cgZ[c.i] = st4(z);
            v[0] = (double)dot3(r, r);
            v[1] = (double)dot3(r, z);
        }
        tile_partials<2>(v, F.partials, tile);
    });
}

__global__ void __launch_bounds__(kFinishThreads) k_cg_update_finish(const SceneParams* __restrict__ P, Fields F) {
    const uint32_t s = blockIdx.x;
    const SceneParams& p = P[s];
    CgState& G = F.state[s].cg;
    if (!G.active || p.count == 0) return;
    double out[2];
    scene_sum<2, kFinishThreads>(F.partials, p.tileBegin, p.tileEnd, out);
    if (threadIdx.x == 0) {
        G.loops += 1u;
        G.res2 = (float)out[0];
        G.err = sqrtf(G.res2 / G.rhs2);
        if (G.res2 < G.thr) {
            G.active = 0;
            atomicSub(&F.live[2], 1u);
            return;
        }
        const float absOld = G.absNew;
        G.absNew = fabsf((float)out[1]);
        G.beta = G.absNew / absOld;
        G.it += 1u;
        G.active = (G.it >= p.minVisc && G.it < p.maxVisc) ? 1u : 0u;
        if (!G.active) atomicSub(&F.live[2], 1u);
    }
}

// p = p * beta + z (:788-802)
__global__ void __launch_bounds__(kTile) k_cg_direction(const SceneParams* __restrict__ P,
                                                        const uint32_t* __restrict__ tileScene, Fields F,
                                                        TileGrid tg) {
    if (F.live[2] == 0u) return;  // every scene has left this solve (block-uniform)
    tile_loop(tg, [&](uint32_t tile) {
        const TileCtx c = tile_ctx(tile, tileScene, P);
        const SceneState& st = F.state[c.s];
        if (!st.cg.active || !c.valid) return;
        F.cgP[c.i] = st4(ld3(F.cgP[c.i]) * st.cg.beta + ld3(F.cgZ[c.i]));
    });
}

// ApplyViscosityForceKernel (:844-862); a zero right-hand side zeroes the solution first (:654-659)
__global__ void __launch_bounds__(kTile) k_visc_apply(const SceneParams* __restrict__ P,
                                                      const uint32_t* __restrict__ tileScene, Fields F, TileGrid tg) {
    tile_loop(tg, [&](uint32_t tile) {
        const TileCtx c = tile_ctx(tile, tileScene, P);
        const SceneParams& p = P[c.s];
        if (!c.valid || !p.enableVisc) return;
        const SceneState& st = F.state[c.s];
        f3 x = ld3(F.cgX[c.i]);
        if (st.cg.zero) {
            x = mk3(0.0f, 0.0f, 0.0f);
            F.cgX[c.i] = st4(x);
        }
        const f3 d = x - ld3(F.vel[c.i]);
        F.acc[c.i] = st4(ld3(F.acc[c.i]) + (1.0f / st.dt) * d);
        F.dv[c.i] = st4(d);
    });
}

void launch_visc_precond(const Launch& L, const KTabHost& kt, const Fields& F) {
    hipLaunchKernelGGL(k_visc_precond, dim3(L.gridGroups), dim3(kGroupSize), kLdsXS, L.stream, L.P, L.tileScene, kt, F,
                       L.capacity, TileGrid(L.tg));
}
void launch_cg_begin(const Launch& L, const KTabHost& kt, const Fields& F) {
    hipLaunchKernelGGL(k_cg_begin, dim3(L.gridGroups), dim3(kGroupSize), kLdsXSV, L.stream, L.P, L.tileScene, kt, F, L.capacity,
                       TileGrid(L.tg));
}
void launch_cg_begin_finish(const Launch& L, const Fields& F) {
    hipLaunchKernelGGL(k_cg_begin_finish, dim3(L.sceneCount), dim3(kFinishThreads), 0, L.stream, L.P, F);
}
void launch_cg_matvec(const Launch& L, const KTabHost& kt, const Fields& F) {
    hipLaunchKernelGGL(k_cg_matvec, dim3(L.gridGroups), dim3(kGroupSize), kLdsXSV, L.stream, L.P, L.tileScene, kt, F,
                       L.capacity, TileGrid(L.tg));
}
void launch_cg_alpha(const Launch& L, const Fields& F) {
    hipLaunchKernelGGL(k_cg_alpha, dim3(L.sceneCount), dim3(kFinishThreads), 0, L.stream, L.P, F);
}
void launch_cg_update(const Launch& L, const Fields& F) {
    hipLaunchKernelGGL(k_cg_update, dim3(L.gridAll), dim3(kTile), 0, L.stream, L.P, L.tileScene, F, L.capacity, TileGrid(L.tg));
}
void launch_cg_update_finish(const Launch& L, const Fields& F) {
    hipLaunchKernelGGL(k_cg_update_finish, dim3(L.sceneCount), dim3(kFinishThreads), 0, L.stream, L.P, F);
}
void launch_cg_direction(const Launch& L, const Fields& F) {
    hipLaunchKernelGGL(k_cg_direction, dim3(L.gridAll), dim3(kTile), 0, L.stream, L.P, L.tileScene, F, TileGrid(L.tg));
}
void launch_visc_apply(const Launch& L, const Fields& F) {
    hipLaunchKernelGGL(k_visc_apply, dim3(L.gridAll), dim3(kTile), 0, L.stream, L.P, L.tileScene, F, TileGrid(L.tg));
}

// ---- Monte-Carlo surface tension (SolveSurfaceTension, DFSPHImplementation.cu:810-839; Zorilla et al. 2020) -----
// ComputeSurfaceTensionClassificationKernel (DFSPHKernels.cu:881-958) with ClassifyParticleConfigurable (:864-872)
__global__ void __launch_bounds__(kTile) k_st_classify(const SceneParams* __restrict__ P,
                                                       const uint32_t* __restrict__ tileScene, Fields F,
                                                       uint32_t cap) {
    const TileCtx c = tile_ctx(tileScene, P);
    const SceneParams& p = P[c.s];
    if (!c.valid || !p.enableST) return;
    const float R = p.R;
    const f3 xi = ld3(F.pos[c.i]);
    const uint32_t n = F.nbrCount[c.i];
    F.nMC[c.i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (n == 0) {
        F.kMC[c.i] = 1.0f / R;
        return;
    }
    // neighbour positions are loaded kStBatch at a time so that their latencies overlap; sums stay in list order
    constexpr uint32_t kStBatch = 8;
    f3 com = mk3(0.0f, 0.0f, 0.0f);
    for (uint32_t j0 = 0; j0 < n; j0 += kStBatch) {
        f3 xb[kStBatch];
#pragma unroll
        for (uint32_t u = 0; u < kStBatch; ++u) xb[u] = ld3(F.pos[nbr(F, cap, c.i, min(j0 + u, n - 1u))]);
#pragma unroll
        for (uint32_t u = 0; u < kStBatch; ++u)
            if (j0 + u < n) com = com + (xb[u] - xi);
    }
    com = com / R;
    const float ci = length(com) / (float)n;
    if (!((float)n <= p.slope * ci + p.constant + 0.0f)) return;
    const uint32_t S = F.state[c.s].S;
    const uint32_t s = F.id[c.i] * S / 3u * 3u;
    const float rr = p.NR / p.r;
    const float lim = rr * rr * p.R2;
    uint32_t count = 0;
    f3 nsum = mk3(0.0f, 0.0f, 0.0f);
    for (uint32_t q = 0; q < S; ++q) {
        const uint32_t i3 = s + 3u * q;
        const f3 pt = R * mk3(F.halton[i3 % kHaltonFloats], F.halton[(i3 + 1u) % kHaltonFloats],
                              F.halton[(i3 + 2u) % kHaltonFloats]);
        bool covered = false;
        for (uint32_t j0 = 0; j0 < n && !covered; j0 += kStBatch) {
            f3 xb[kStBatch];
#pragma unroll
            for (uint32_t u = 0; u < kStBatch; ++u) xb[u] = ld3(F.pos[nbr(F, cap, c.i, min(j0 + u, n - 1u))]);
#pragma unroll
            for (uint32_t u = 0; u < kStBatch; ++u)
                if (j0 + u < n && length2(pt - (xb[u] - xi)) <= lim) covered = true;
        }
        if (covered) continue;
        nsum = nsum + pt;
        ++count;
    }
    if (count > 0) {
        if (length2(nsum) > 0.0f) nsum = normalize(nsum);
        F.nMC[c.i] = st4(nsum);
        F.kMC[c.i] = 1.0f / R * -2.0f * sqrtf(1.0f - p.NR * p.NR / (p.r * p.r)) * F.mcTab[F.mcOff[c.s] + count];
    } else {
        F.kMC[c.i] = 0.0f;
    }
}

// ComputeSurfaceTensionNormalsAndCurvatureKernel (:960-1006)
__global__ void __launch_bounds__(kTile) k_st_normals(const SceneParams* __restrict__ P,
                                                      const uint32_t* __restrict__ tileScene, Fields F,
                                                      uint32_t cap) {
    const TileCtx c = tile_ctx(tileScene, P);
    const SceneParams& p = P[c.s];
    if (!c.valid || !p.enableST) return;
    const f3 ni = ld3(F.nMC[c.i]);
    if (is_zero(ni)) return;
    const float R = p.R, sf = p.smoothing;
    const f3 xi = ld3(F.pos[c.i]);
    f3 nc = mk3(0.0f, 0.0f, 0.0f);
    float cc = 0.0f, cf = 0.0f;
    const uint32_t n = F.nbrCount[c.i];
    for (uint32_t j = 0; j < n; ++j) {
        const uint32_t t = nbr(F, cap, c.i, j);
        const f3 nj = ld3(F.nMC[t]);
        if (is_zero(nj)) continue;
        const float dist = length(ld3(F.pos[t]) - xi);
        const float num = 1.0f - dist / R;
        nc = nc + nj * num;
        cc += F.kMC[t] * num;
        cf += num;
    }
    if (length2(nc) > 0.0f) nc = normalize(nc);
    f3 ns = (1.0f - sf) * ni + sf * nc;
    if (length2(ns) > 0.0f) ns = normalize(ns);
    F.nS[c.i] = st4(ns);
    F.kS[c.i] = ((1.0f - sf) * F.kMC[c.i] + sf * cc) / (1.0f - sf + sf * cf);
}

// ComputeSurfaceTensionBlendingKernel (:1008-1046), applied SurfaceTensionSmoothPassCount times
__global__ void __launch_bounds__(kTile) k_st_blend(const SceneParams* __restrict__ P,
                                                    const uint32_t* __restrict__ tileScene, Fields F) {
    const TileCtx c = tile_ctx(tileScene, P);
    const SceneParams& p = P[c.s];
    if (!c.valid || !p.enableST) return;
    float dk = F.dk[c.i];
    f3 a = ld3(F.acc[c.i]);
    const bool surface = !is_zero(ld3(F.nMC[c.i]));
    const f3 ns = ld3(F.nS[c.i]);
    const float ks = F.kS[c.i];
    for (uint32_t pass = 0; pass < p.stPasses; ++pass) {
        if (surface) {
            float fk = ks;
            if (p.temporal) fk = 0.05f * fk + 0.95f * dk;
            const f3 force = ns * p.sigma * fk;
            a = a - p.mInv * force;
            dk = fk;
        } else {
            float fk = 0.0f;
            if (p.temporal) fk = 0.95f * dk;
            dk = fk;
        }
    }
    F.acc[c.i] = st4(a);
    F.dk[c.i] = dk;
}

void launch_surface_tension(const Launch& L, const Fields& F) {
    hipLaunchKernelGGL(k_st_classify, dim3(L.tiles), dim3(kTile), 0, L.stream, L.P, L.tileScene, F, L.capacity);
    hipLaunchKernelGGL(k_st_normals, dim3(L.tiles), dim3(kTile), 0, L.stream, L.P, L.tileScene, F, L.capacity);
    hipLaunchKernelGGL(k_st_blend, dim3(L.tiles), dim3(kTile), 0, L.stream, L.P, L.tileScene, F);
}

}  // namespace vfd

namespace vfd {
// Dynamic LDS above the 64 KB default: kernel table plus one group's halo records.
int configure_kernels() {
    struct {
        const void* fn;
        size_t bytes;
    } const k[] = {
        {(const void*)k_density_factor, kLdsDensity},
        {(const void*)k_div_init, kLdsXV},
        {(const void*)k_div_accel, kLdsXS},
        {(const void*)k_div_iter, kLdsXV},
        {(const void*)k_div_final, kLdsXS},
        {(const void*)k_pres_init, kLdsXV},
        {(const void*)k_pres_accel, kLdsXS},
        {(const void*)k_pres_iter, kLdsXV},
        {(const void*)k_pres_final, kLdsXS},
        {(const void*)k_visc_precond, kLdsXS},
        {(const void*)k_cg_begin, kLdsXSV},
        {(const void*)k_cg_matvec, kLdsXSV},
    };
    for (const auto& e : k)
        if (hipFuncSetAttribute(e.fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)e.bytes) != hipSuccess) return 1;
    return configure_search_kernels() || configure_adjoint_kernels();
}
}  // namespace vfd
